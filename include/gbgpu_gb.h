/*
 * gbgpu_gb.h -- drop-in replacements for gb's linear-algebra function
 * pointers, on gb's own data structures.
 *
 * gb (v0.16) calls its finite-field linear algebra through two global
 * function pointers set by set_function_pointers (src/io.c:581-780):
 *
 *   void (*linear_algebra)(mat_t *mat, const bs_t * const bs, stat_t *st);
 *                                                     (src/data.h:269-273)
 *   void (*interreduce_matrix_rows)(mat_t *mat, bs_t *bs, const stat_t *st);
 *                                                     (src/data.h:275-279)
 *
 * The functions below have exactly these signatures (struct mat_t,
 * struct bs_t and struct stat_t are gb's types; the library is built
 * against a layout-identical mirror of them) and keep the reference's
 * contract:
 *   - input: mat->r[0..nr) rows [cf id, len % 4, len, columns...] sorted by
 *     (lead, length), coefficients in bs->cf_{8,16,32}[row[0]];
 *   - output: mat->np new rows in mat->r, each [i, len % 4, len, strictly
 *     ascending columns >= ncl] with coefficient array mat->cf_*[i] (unit
 *     lead), in decreasing lead order (la_ff_32.c:1046, :1687);
 *     mat->nr = mat->sz = np;
 *   - every input row array is freed, basis coefficient arrays are not;
 *   - st->la_ctime / st->la_rtime / st->num_zerored are updated and the
 *     "new / zero" counts printed at info_level > 1 (la_ff_32.c:1840-1850).
 * The kernel form follows st->laopt: 101 = plain integer kernels, any other
 * value = the MFMA form.  There is no CPU fallback: a device error prints
 * a message and aborts, as an unchecked malloc failure would in gb.
 *
 * The interreduce variants replace interreduce_matrix_rows_ff_*
 * (la_ff_32.c:1949-2017): they free the basis elements like the reference
 * (bs->ld becomes 0) and return the rows in ascending lead order.
 */
#ifndef GBGPU_GB_H
#define GBGPU_GB_H

#ifdef __cplusplus
extern "C" {
#endif

struct mat_t;
struct bs_t;
struct stat_t;

/* replaces exact_sparse_dense_linear_algebra_ff_32 / exact_sparse_linear_
 * algebra_ff_32 (la_ff_32.c:1774-1851); p < 2^31 */
void gbgpu_linear_algebra_ff_32(struct mat_t *mat, const struct bs_t *const bs, struct stat_t *st);
/* replaces the same routines of la_ff_16.c (:1430-1508); p < 2^15 */
void gbgpu_linear_algebra_ff_16(struct mat_t *mat, const struct bs_t *const bs, struct stat_t *st);
/* replaces the same routines of la_ff_8.c (:1430-1508); p < 2^7 */
void gbgpu_linear_algebra_ff_8(struct mat_t *mat, const struct bs_t *const bs, struct stat_t *st);

/* replace interreduce_matrix_rows_ff_{32,16,8} (la_ff_32.c:1949-2017,
 * la_ff_16.c:1605-1673) */
void gbgpu_interreduce_matrix_rows_ff_32(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st);
void gbgpu_interreduce_matrix_rows_ff_16(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st);
void gbgpu_interreduce_matrix_rows_ff_8(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st);

#ifdef __cplusplus
}
#endif
#endif
