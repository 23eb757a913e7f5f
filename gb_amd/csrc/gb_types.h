/*
 * gb_types.h -- layout mirror of the gb (v0.16) structures the linear
 * algebra touches: mat_t, bs_t, stat_t (src/data.h:52-211).
 *
 * These are an interface, not an implementation: the field order, the
 * types and hence the C layout must be those of gb's data.h, so that the
 * adapter functions can be stored in gb's function pointers.  GMP's
 * mpz_t ** members are represented by void ** (same size).  tests/ check
 * the sizes and offsets against values computed from the field list.
 */
#ifndef GBGPU_GB_TYPES_H
#define GBGPU_GB_TYPES_H

#include <stdint.h>

typedef int8_t cf8_t;     /* data.h:53 */
typedef int16_t cf16_t;   /* data.h:54 */
typedef int32_t cf32_t;   /* data.h:55 */
typedef int32_t hm_t;     /* data.h:56-58 (val_t -> hl_t -> hm_t) */
typedef int32_t len_t;    /* data.h:62 */

/* data.h:117-135 */
struct bs_t {
    int32_t ld;
    int32_t sz;
    int32_t lo;
    int32_t mltdeg;
    int32_t *lmps;
    int32_t *lm;
    int32_t lml;
    int8_t *red;
    hm_t **hm;
    cf8_t **cf_8;
    cf16_t **cf_16;
    cf32_t **cf_32;
    void **cf_qq;
};

/* data.h:137-155 */
struct mat_t {
    hm_t **r;
    cf8_t **cf_8;
    cf16_t **cf_16;
    cf32_t **cf_32;
    void **cf_qq;
    void **cf_ab_qq;
    len_t sz;
    len_t np;
    len_t nr;
    len_t nc;
    len_t nru;
    len_t nrl;
    len_t ncl;
    len_t ncr;
};

/* data.h:158-211 */
struct stat_t {
    double round_ctime, select_ctime, symbol_ctime, la_ctime, update_ctime,
        convert_ctime, overall_ctime, reduce_gb_ctime, rht_ctime;
    double round_rtime, select_rtime, symbol_rtime, la_rtime, update_rtime,
        convert_rtime, overall_rtime, reduce_gb_rtime, rht_rtime;
    int64_t num_pairsred, num_gb_crit, num_redundant_old, num_redundant, num_rht,
        num_rowsred, num_zerored;
    int32_t ngens, nvars, mnsel, homogeneous, fc, mo, laopt, init_hts, nthrds, reset_ht,
        current_rd, current_deg;
    int64_t max_bht_size, max_sht_size, max_uht_size, nterms_basis;
    int32_t size_basis, ff_bits, reduce_gb;
    int32_t info_level, gen_pbm_file;
};

#endif
