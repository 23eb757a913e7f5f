/*
 * la_ff_oracle.h -- CPU restatement of gb's finite-field F4 linear algebra.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libgbgpu.so) never links it
 * and never falls back to it.
 *
 * Reference followed: gb v0.16, src/la_ff_32.c, src/la_ff_16.c,
 * src/la_ff_8.c, src/tools.c, src/io.c (field-size selector).
 *
 * Matrix format: one flat CSR per matrix, the same wire format the GPU
 * library takes (see include/gbgpu.h).  Row i has columns
 * ci[rp[i] .. rp[i+1]) (first one is the lead = minimum column, the rest in
 * any order) and coefficients cf[cp[rcf[i]] + k] for k < rp[i+1]-rp[i]:
 * rows built from one basis element share one coefficient array, exactly
 * as hm_t rows share bs->cf_32[row[0]] in gb (hash.c:903-926).
 */
#ifndef GB_LA_FF_ORACLE_H
#define GB_LA_FF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t nr;          /* rows */
    int32_t nc;          /* columns */
    int32_t ncl;         /* left (known pivot) columns; ncr = nc - ncl */
    const int64_t *rp;   /* nr + 1 row offsets into ci */
    const int32_t *ci;   /* column indices */
    const int32_t *rcf;  /* coefficient array id per row */
    const int64_t *cp;   /* coefficient array offsets into cf */
    const uint32_t *cf;  /* coefficient pool, entries in [0, p) */
} orc_mat_t;

typedef struct {
    int32_t np;          /* number of rows */
    int64_t *rp;         /* np + 1 */
    int32_t *ci;         /* strictly ascending columns per row */
    uint32_t *cf;        /* coefficients; first of each row is 1 */
} orc_rows_t;

/* arithmetic variant chosen exactly like io.c:815-829 + io.c:710-729:
 * 8 (p < 2^7), 16 (p < 2^15), 17 (p < 2^17), 31 (otherwise) */
int32_t orc_arith_bits(int32_t p);

/* tools.c:83-110 */
int32_t orc_mod_inverse(int32_t val, int32_t p);

/* la_ff_32.c:24-60 (normalize_dense_matrix_row_ff_32), in place */
void orc_normalize_dense_row(int32_t *row, int32_t len, int32_t p);

/* la_ff_32.c:1081-1179 (sparse_AB_CD_linear_algebra_ff_32): the plain
 * sparse pass.  Writes the nonzero reduced rows (dense, ncr entries each)
 * into out (capacity nrl rows) in the reference's drs order and returns
 * their number; src_row[k] receives the input row index that produced
 * output row k.  Returns -1 if the input violates the pivot contract. */
int64_t orc_sparse_pass(const orc_mat_t *m, int32_t p, int32_t nthreads,
                        uint32_t *out, int32_t *src_row);

/* la_ff_32.c:1217-1333 + 1181-1215 + 1666-1736: exact dense echelon form,
 * interreduction and conversion of an m x n dense block (entries are
 * reduced mod p on entry).  Output rows are in decreasing lead column; row
 * k has lead lead[k] and is written full width to out[k*n .. k*n+n).
 * Either output may be NULL.  Returns the rank. */
int32_t orc_dense_rref(int32_t m, int32_t n, const uint32_t *dm, int32_t p,
                       int32_t nthreads, uint32_t *out, int32_t *lead);

/* la_ff_32.c:1804-1851 (la_option 1, exact_sparse_dense_linear_algebra). */
int32_t orc_la_sparse_dense(const orc_mat_t *m, int32_t p, int32_t nthreads,
                            orc_rows_t *out);

/* la_ff_32.c:1774-1802 + 945-1079 (la_option 2, exact sparse RREF). */
int32_t orc_la_sparse(const orc_mat_t *m, int32_t p, int32_t nthreads,
                      orc_rows_t *out);

/* la_ff_32.c:1949-2017 (interreduce_matrix_rows): ncl = 0, all rows have
 * distinct leads; output in ascending lead order. */
int32_t orc_interreduce(const orc_mat_t *m, int32_t p, orc_rows_t *out);

/* per to-reduce row statistics of the sparse pass: number of pivots
 * applied (|S(r)|) and the pivot entries streamed (sum of len(P_j)). */
int64_t orc_sparse_pass_work(const orc_mat_t *m, int32_t p, int32_t nthreads,
                             int64_t *npiv_applied, int64_t *entries_streamed);

void orc_rows_free(orc_rows_t *r);

#ifdef __cplusplus
}
#endif
#endif
