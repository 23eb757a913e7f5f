/* gbgpu_f4.h -- the F4 host driving libgbgpu's linear algebra.
 *
 * A C implementation of the F4 loop around the linear-algebra step: the
 * steps gb runs in f4_julia (f4.c:190-372) -- pair selection by minimal
 * degree, symbolic preprocessing, matrix construction in gb's row format,
 * linear algebra, conversion of the new rows to basis elements and the
 * Gebauer-Moeller update, optionally the final interreduction
 * (reduce_basis, f4.c:90-179).  Linear algebra is a callback, so the same
 * host runs on the GPU (gbgpu_f4_gpu_la) or on any CPU routine with the
 * same contract.
 *
 * Input and output follow f4_julia's arrays (f4.c:180-212, io.c:332-392):
 * lens[i] terms per generator, exps nr_vars exponents per term, cfs the
 * coefficients; the basis comes back as bld elements with blen, bexp and
 * bcf (coefficients in [0, p), leading coefficient 1, terms in decreasing
 * monomial order).  Only prime fields are handled (field_char > 1). */
#ifndef GBGPU_F4_H
#define GBGPU_F4_H

#include <stdint.h>

#include "gbgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* variables per system: gb has no limit of its own; this bound only keeps
 * an exponent vector (2 bytes per variable) a sane size */
#define GBGPU_F4_MAX_VARS 4096

typedef struct {
    int32_t nr_vars;     /* 1 .. GBGPU_F4_MAX_VARS */
    int32_t nr_gens;
    const int32_t *lens; /* nr_gens */
    const int32_t *exps; /* nr_vars per term, each in [0, 32767] (gb's int16 exp_t);
                            a run whose products leave that range returns GBGPU_ERANGE */
    const int32_t *cfs;  /* one per term, any sign (taken mod field_char) */
    int32_t field_char;  /* prime below 2^31 */
    int32_t mon_order;   /* 0: degree reverse lexicographical, 1: lexicographical */
    int32_t la_option;   /* accepted as f4_julia does; every exact option gives the same rows */
    int32_t reduce_gb;   /* 1: return the reduced basis */
    int32_t info_level;  /* > 1: one line per round on stdout */
} gbgpu_f4_in_t;

typedef struct {
    int32_t bld;         /* basis elements */
    int32_t *blen;       /* bld */
    int32_t *bexp;       /* nterms x nr_vars */
    uint32_t *bcf;       /* nterms */
    int64_t nterms;
    /* statistics */
    int32_t rounds;
    int64_t max_rows, max_cols, max_nnz;  /* largest matrix */
    int64_t rows_total;                   /* rows to reduce, summed over rounds */
    double la_ms;                         /* wall time inside the callback */
    double total_ms;
    double select_ms, symbolic_ms, build_ms, update_ms;  /* host phases */
    /* one record of GBGPU_F4_TRACE_W values per round (rounds records):
     * degree, pairs selected, rows, columns, pivot columns, nonzeros, new
     * rows -- the matrix shape and result gb prints at info_level 2
     * (symbol.c:72-75, convert.c:131, la_ff_32.c:1845) -- and a 64-bit
     * FNV-1a hash of the round's echelon form (the LA result rows) */
    int64_t *trace;
} gbgpu_f4_out_t;

#define GBGPU_F4_TRACE_W 8

/* Linear algebra of one round: interreduce == 0 is gb's linear_algebra on
 * the matrix m (data.h:269-273: known pivots = first row per lead, result
 * = the new pivots in decreasing lead order, columns >= ncl); interreduce
 * == 1 is interreduce_matrix_rows (ncl = 0, every row, result in
 * increasing lead order).  out must be freeable by gbgpu_rows_free.
 * Returns 0 on success. */
typedef int (*gbgpu_f4_la_t)(void *user, const gbgpu_csr_t *m, int interreduce,
                             gbgpu_rows_t *out);

/* Runs F4 on the input system.  Returns GBGPU_OK or an error code. */
int gbgpu_f4(const gbgpu_f4_in_t *in, gbgpu_f4_la_t la, void *user, gbgpu_f4_out_t *out);
void gbgpu_f4_out_free(gbgpu_f4_out_t *out);

/* The GPU back end: user points to a gbgpu_f4_gpu_t. */
typedef struct {
    gbgpu_ctx *ctx;
    int32_t form;        /* GBGPU_FORM_PLAIN or GBGPU_FORM_MFMA */
    double device_ms;    /* summed device time of the LA calls (phase 2) */
} gbgpu_f4_gpu_t;
int gbgpu_f4_gpu_la(void *user, const gbgpu_csr_t *m, int interreduce, gbgpu_rows_t *out);

/* F4 of one system modulo several primes (the images of a multimodular run),
 * in lockstep: ins[i] is image i (its own field_char), outs[i] its result.
 * Each image runs the F4 host in a host thread of its own; in every round
 * the matrices of all images still running go through one batched launch
 * (gbgpu_batch_la: la_option 1 on every matrix, each modulo its prime), in
 * `form`; the final interreduction (reduce_gb) runs image by image.  The
 * host's OpenMP threads are shared out over the images.  No counterpart in
 * the reference, whose rational path is fraction-free over GMP
 * (la_qq.c:271-362).  Returns GBGPU_OK or the first error; outs are freed
 * with gbgpu_f4_out_free. */
typedef struct {
    int64_t batches;     /* batched launches */
    int64_t matrices;    /* matrices they reduced */
    double device_ms;    /* their graphs' device time plus the interreductions' */
} gbgpu_f4_multi_stats_t;
int gbgpu_f4_multimod(gbgpu_ctx *ctx, const gbgpu_f4_in_t *ins, int32_t nimg, int32_t form,
                      gbgpu_f4_out_t *outs, gbgpu_f4_multi_stats_t *stats);

/* f4_julia's signature (f4.c:190-212) on the GPU back end: device 0 (or
 * GBGPU_DEVICE), form from la_option (101 plain, otherwise MFMA, which
 * also maps la_option 1/2/42/43/44 to the exact routine).  ht_size,
 * nr_threads, max_nr_pairs, reset_ht and pbm_file are accepted and
 * ignored.  Returns the number of terms of the basis (0 on failure); the
 * arrays are malloc'ed. */
int64_t gbgpu_f4_julia(int32_t *bld, int32_t **blen, int32_t **bexp, void **bcf,
                       const int32_t *lens, const int32_t *exps, const void *cfs,
                       const int32_t field_char, const int32_t mon_order,
                       const int32_t nr_vars, const int32_t nr_gens, const int32_t ht_size,
                       const int32_t nr_threads, const int32_t max_nr_pairs,
                       const int32_t reset_ht, const int32_t la_option,
                       const int32_t reduce_gb, const int32_t pbm_file,
                       const int32_t info_level);

#ifdef __cplusplus
}
#endif
#endif
