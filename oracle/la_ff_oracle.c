/*
 * la_ff_oracle.c -- CPU restatement of gb's finite-field F4 linear algebra.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
 *
 * Every function below names the reference routine it restates (gb v0.16).
 * The arithmetic variants are kept as in the reference:
 *   - 8/16/17-bit (la_ff_8.c / la_ff_16.c / la_ff_32.c *_17_bit): the
 *     multiplier is p - v and products are added into an int64 row, which
 *     is reduced lazily when a column is visited;
 *   - 31-bit (la_ff_32.c *_31_bit): the multiplier is v, products are
 *     subtracted and the entry is lifted by p^2 when it turns negative.
 * The OpenMP row parallelism and the compare-and-swap pivot publication of
 * the reference are kept, so the timing of this file is a fair stand-in for
 * the reference's CPU path.
 */
#include "la_ff_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- tools */

/* io.c:815-829 picks the storage width, io.c:710-729 the 17/31-bit
 * kernels of the 32-bit file. */
int32_t orc_arith_bits(int32_t p)
{
    if (p < (1 << 7))  return 8;
    if (p < (1 << 15)) return 16;
    if (p < (1 << 17)) return 17;
    return 31;
}

/* tools.c:83-110: extended Euclid on (p, val mod p); 0 maps to 0. */
int32_t orc_mod_inverse(int32_t val, int32_t p)
{
    int32_t r_prev = p, r = val % p;
    int32_t s_prev = 0, s = 1;
    if (r < 0) r += p;
    while (r != 0) {
        const int32_t q = r_prev / r;
        int32_t t = r_prev - q * r;
        r_prev = r;
        r = t;
        t = s_prev - q * s;
        s_prev = s;
        s = t;
    }
    if (s_prev < 0) s_prev += p;
    return s_prev;
}

/* la_ff_32.c:24-60: scale entries 1..len-1 by row[0]^-1, set row[0] = 1.
 * Inputs need not be reduced (the reference's own known-answer test,
 * test/normalize_row.c, feeds values above p). */
void orc_normalize_dense_row(int32_t *row, int32_t len, int32_t p)
{
    const int64_t inv = orc_mod_inverse(row[0], p);
    for (int32_t i = 1; i < len; ++i) {
        int64_t t = ((int64_t)row[i] * inv) % p;
        t += (t >> 63) & p;
        row[i] = (int32_t)t;
    }
    row[0] = 1;
}

static void normalize_u32(uint32_t *row, int64_t len, int32_t p)
{
    const uint64_t inv = (uint64_t)orc_mod_inverse((int32_t)row[0], p);
    for (int64_t i = 1; i < len; ++i)
        row[i] = (uint32_t)(((uint64_t)row[i] * inv) % (uint64_t)p);
    row[0] = 1;
}

/* ------------------------------------------------- sparse pivot rows */

/* A sparse row: len entries, columns in any order except that entry 0 is
 * the lead.  Known pivots point into the input, new ones are owned. */
typedef struct {
    int32_t len;
    const int32_t *col;
    const uint32_t *cf;
    int owned;
} srow_t;

static srow_t input_row(const orc_mat_t *m, int32_t i)
{
    srow_t r;
    r.len = (int32_t)(m->rp[i + 1] - m->rp[i]);
    r.col = m->ci + m->rp[i];
    r.cf = m->cf + m->cp[m->rcf[i]];
    r.owned = 0;
    return r;
}

static void srow_free(srow_t *r)
{
    if (r && r->owned) {
        free((void *)r->col);
        free((void *)r->cf);
    }
    free(r);
}

/* dr[c] updated by one pivot row with value v = dr[lead] mod p (v != 0):
 * la_ff_32.c:528-536 (17-bit: += (p - v) * cf) and :593-606 (31-bit:
 * -= v * cf, lifted by p^2 when negative). */
static inline void apply_sparse(int64_t *dr, const srow_t *pv, int64_t v,
                                int64_t p, int bits)
{
    if (bits != 31) {
        const int64_t mul = p - v;
        for (int32_t j = 0; j < pv->len; ++j)
            dr[pv->col[j]] += mul * (int64_t)pv->cf[j];
    } else {
        const int64_t p2 = p * p;
        for (int32_t j = 0; j < pv->len; ++j) {
            int64_t t = dr[pv->col[j]] - v * (int64_t)pv->cf[j];
            t += (t >> 63) & p2;
            dr[pv->col[j]] = t;
        }
    }
}

static inline void apply_dense(int64_t *dr, const uint32_t *pv, int32_t from,
                               int32_t to, int64_t v, int64_t p, int bits)
{
    if (bits != 31) {
        const int64_t mul = p - v;
        for (int32_t j = from; j < to; ++j)
            dr[j] += mul * (int64_t)pv[j - from];
    } else {
        const int64_t p2 = p * p;
        for (int32_t j = from; j < to; ++j) {
            int64_t t = dr[j] - v * (int64_t)pv[j - from];
            t += (t >> 63) & p2;
            dr[j] = t;
        }
    }
}

static srow_t *load_piv(srow_t *const *pivs, int32_t i)
{
    return __atomic_load_n(&pivs[i], __ATOMIC_ACQUIRE);
}

/* ---------------------------------------------------- a4: old pivots */

/* la_ff_32.c:496-558 / :560-630 / la_ff_16.c:289-351: reduce the dense
 * row by the known pivots of columns [dpiv, ncl) and return its right part
 * (ncr entries, reduced) or NULL if it is zero.  The two counters feed the
 * algorithmic-work figures of the benchmark. */
static uint32_t *reduce_by_old_pivots(int64_t *dr, const orc_mat_t *m,
                                      srow_t *const *pivs, int32_t dpiv,
                                      int64_t p, int bits, int64_t *napp,
                                      int64_t *nent)
{
    const int32_t ncl = m->ncl, nc = m->nc;
    for (int32_t i = dpiv; i < ncl; ++i) {
        if (dr[i] != 0) dr[i] %= p;
        if (dr[i] == 0) continue;
        const srow_t *pv = pivs[i];
        if (pv == NULL) continue;
        apply_sparse(dr, pv, dr[i], p, bits);
        dr[i] = 0;
        if (napp) { ++*napp; *nent += pv->len; }
    }
    uint32_t *row = (uint32_t *)calloc((size_t)(nc - ncl), sizeof(uint32_t));
    int64_t nz = 0;
    for (int32_t i = ncl; i < nc; ++i) {
        if (dr[i] != 0) {
            dr[i] %= p;
            if (dr[i] != 0) { row[i - ncl] = (uint32_t)dr[i]; ++nz; }
        }
    }
    if (nz == 0) { free(row); row = NULL; }
    return row;
}

/* The known-pivot split shared by la_option 1 and 2
 * (la_ff_32.c:1106-1116, :964-978): the first row of each lead column is
 * the pivot, every later row is to be reduced, and the to-reduce list is
 * filled from its end ("upivs[nrl-j]"), i.e. it is the reversed order of
 * appearance.  Returns the number of to-reduce rows or -1 if a lead lies
 * outside [0, lim). */
static int32_t split_rows(const orc_mat_t *m, int32_t lim, srow_t **pivs,
                          int32_t *upiv)
{
    int32_t nup = 0;
    for (int32_t i = 0; i < m->nr; ++i) {
        if (m->rp[i + 1] == m->rp[i]) return -1;
        const int32_t lead = m->ci[m->rp[i]];
        if (lead < 0 || lead >= lim) return -1;
        if (pivs[lead] == NULL) {
            srow_t *r = (srow_t *)malloc(sizeof(srow_t));
            *r = input_row(m, i);
            pivs[lead] = r;
        } else {
            upiv[nup++] = i;
        }
    }
    for (int32_t a = 0, b = nup - 1; a < b; ++a, --b) {
        const int32_t t = upiv[a]; upiv[a] = upiv[b]; upiv[b] = t;
    }
    return nup;
}

static void scatter_input(int64_t *dr, const orc_mat_t *m, int32_t i)
{
    const int64_t b = m->rp[i], e = m->rp[i + 1];
    const uint32_t *cf = m->cf + m->cp[m->rcf[i]];
    for (int64_t k = b; k < e; ++k) dr[m->ci[k]] = (int64_t)cf[k - b];
}

/* a3: la_ff_32.c:1081-1179 (sparse_AB_CD_linear_algebra_ff_32). */
static int64_t sparse_pass_impl(const orc_mat_t *m, int32_t p, int32_t nthreads,
                                uint32_t **drs_out, int32_t *src_out,
                                int64_t *napp_tot, int64_t *nent_tot,
                                uint32_t *dense_out)
{
    const int bits = orc_arith_bits(p);
    const int32_t nc = m->nc, ncl = m->ncl, ncr = nc - ncl;
    srow_t **pivs = (srow_t **)calloc((size_t)(ncl > 0 ? ncl : 1), sizeof(srow_t *));
    int32_t *upiv = (int32_t *)malloc((size_t)(m->nr > 0 ? m->nr : 1) * sizeof(int32_t));
    const int32_t nrl = split_rows(m, ncl, pivs, upiv);
    if (nrl < 0) {
        for (int32_t i = 0; i < ncl; ++i) free(pivs[i]);
        free(pivs); free(upiv);
        return -1;
    }
    if (nthreads < 1) nthreads = 1;
    uint32_t **drs = (uint32_t **)calloc((size_t)(nrl > 0 ? nrl : 1), sizeof(uint32_t *));
    int64_t *dr = (int64_t *)malloc((size_t)nthreads * (size_t)(nc > 0 ? nc : 1) * sizeof(int64_t));
    int64_t napp = 0, nent = 0;
    const int chunk = nrl / nthreads / 100 == 0 ? 1 : nrl / nthreads / 100;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, chunk) reduction(+:napp, nent)
    for (int32_t i = 0; i < nrl; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t *drl = dr + (size_t)tid * (size_t)nc;
        memset(drl, 0, (size_t)nc * sizeof(int64_t));
        scatter_input(drl, m, upiv[i]);
        const int32_t sc = m->ci[m->rp[upiv[i]]];
        int64_t a = 0, e = 0;
        drs[i] = reduce_by_old_pivots(drl, m, pivs, sc, p, bits,
                                      napp_tot ? &a : NULL, &e);
        napp += a; nent += e;
    }
    free(dr);
    int64_t np = 0;
    for (int32_t i = 0; i < nrl; ++i) {
        if (drs[i] == NULL) continue;
        if (src_out) src_out[np] = upiv[i];
        if (dense_out) {
            memcpy(dense_out + (size_t)np * (size_t)ncr, drs[i], (size_t)ncr * sizeof(uint32_t));
            free(drs[i]);
        } else {
            drs[np] = drs[i];
        }
        ++np;
    }
    for (int32_t i = 0; i < ncl; ++i) free(pivs[i]);
    free(pivs);
    free(upiv);
    if (napp_tot) { *napp_tot = napp; *nent_tot = nent; }
    if (drs_out) *drs_out = (uint32_t *)drs; else free(drs);
    return np;
}

int64_t orc_sparse_pass(const orc_mat_t *m, int32_t p, int32_t nthreads,
                        uint32_t *out, int32_t *src_row)
{
    return sparse_pass_impl(m, p, nthreads, NULL, src_row, NULL, NULL, out);
}

int64_t orc_sparse_pass_work(const orc_mat_t *m, int32_t p, int32_t nthreads,
                             int64_t *npiv_applied, int64_t *entries_streamed)
{
    uint32_t **drs = NULL;
    const int64_t np = sparse_pass_impl(m, p, nthreads, (uint32_t **)&drs, NULL,
                                        npiv_applied, entries_streamed, NULL);
    if (np > 0) for (int64_t i = 0; i < np; ++i) free(drs[i]);
    free(drs);
    return np;
}

/* ------------------------------------------------- a5-a8: dense part */

/* la_ff_32.c:632-688 / :690-753 / la_ff_16.c:353-409: reduce a dense row
 * (ncr entries, column index = position) from column *pc on by the dense
 * pivots nps[] (nps[c] holds columns c..ncr-1).  Returns the normalised
 * remainder starting at its first free column, stored in *pc, or NULL with
 * *pc = -1 when nothing is left. */
static uint32_t *reduce_by_dense_pivots(int64_t *dr, int32_t *pc,
                                        uint32_t *const *nps, int32_t ncr,
                                        int64_t p, int bits)
{
    int32_t np = -1;
    int64_t k = 0;
    for (int32_t i = *pc; i < ncr; ++i) {
        if (dr[i] != 0) dr[i] %= p;
        if (dr[i] == 0) continue;
        const uint32_t *pv = __atomic_load_n(&nps[i], __ATOMIC_ACQUIRE);
        if (pv == NULL) {
            if (np == -1) np = i;
            ++k;
            continue;
        }
        apply_dense(dr, pv, i, ncr, dr[i], p, bits);
    }
    if (k == 0) { *pc = -1; return NULL; }
    uint32_t *row = (uint32_t *)calloc((size_t)(ncr - np), sizeof(uint32_t));
    for (int32_t i = np; i < ncr; ++i) {
        if (dr[i] != 0) dr[i] %= p;
        row[i - np] = (uint32_t)dr[i];
    }
    if (row[0] != 1) normalize_u32(row, ncr - np, (int32_t)p);
    *pc = np;
    return row;
}

/* a5: la_ff_32.c:1217-1333 (exact_dense_linear_algebra_ff_32).  The rows
 * in dm (nrows of them, ncr entries each, none zero) are consumed. */
static uint32_t **dense_echelon(uint32_t **dm, int32_t nrows, int32_t ncr,
                                int64_t p, int bits, int32_t nthreads)
{
    uint32_t **nps = (uint32_t **)calloc((size_t)(ncr > 0 ? ncr : 1), sizeof(uint32_t *));
    uint32_t **tbr = (uint32_t **)malloc((size_t)(nrows > 0 ? nrows : 1) * sizeof(uint32_t *));
    int32_t ntr = 0;
    for (int32_t i = 0; i < nrows; ++i) {
        int32_t k = 0;
        while (dm[i][k] == 0) ++k;
        if (nps[k] == NULL) {
            memmove(dm[i], dm[i] + k, (size_t)(ncr - k) * sizeof(uint32_t));
            if (dm[i][0] != 1) normalize_u32(dm[i], ncr - k, (int32_t)p);
            nps[k] = dm[i];
        } else {
            tbr[ntr++] = dm[i];
        }
    }
    if (nthreads < 1) nthreads = 1;
    int64_t *dr = (int64_t *)malloc((size_t)nthreads * (size_t)(ncr > 0 ? ncr : 1) * sizeof(int64_t));
    const int chunk = ntr / nthreads / 100 == 0 ? 1 : ntr / nthreads / 100;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, chunk)
    for (int32_t i = 0; i < ntr; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t *drl = dr + (size_t)tid * (size_t)ncr;
        for (int32_t j = 0; j < ncr; ++j) drl[j] = (int64_t)tbr[i][j];
        free(tbr[i]);
        int32_t npc = 0;
        for (;;) {
            uint32_t *npiv = reduce_by_dense_pivots(drl, &npc, nps, ncr, p, bits);
            if (npc == -1) break;
            uint32_t *expect = NULL;
            if (__atomic_compare_exchange_n(&nps[npc], &expect, npiv, 0,
                                            __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE))
                break;
            free(npiv); /* another thread won this column: reduce further */
        }
    }
    free(dr);
    free(tbr);
    return nps;
}

/* a7: la_ff_32.c:1181-1215 (interreduce_dense_matrix_ff_32), right to
 * left, so every pivot is reduced by already reduced pivots. */
static void dense_interreduce(uint32_t **nps, int32_t ncr, int64_t p, int bits)
{
    int64_t *dr = (int64_t *)malloc((size_t)(ncr > 0 ? ncr : 1) * sizeof(int64_t));
    for (int32_t i = ncr - 1; i >= 0; --i) {
        if (nps[i] == NULL) continue;
        memset(dr, 0, (size_t)ncr * sizeof(int64_t));
        for (int32_t j = i; j < ncr; ++j) dr[j] = (int64_t)nps[i][j - i];
        free(nps[i]);
        nps[i] = NULL;
        int32_t pc = i;
        nps[i] = reduce_by_dense_pivots(dr, &pc, nps, ncr, p, bits);
    }
    free(dr);
}

int32_t orc_dense_rref(int32_t m, int32_t n, const uint32_t *dmin, int32_t p,
                       int32_t nthreads, uint32_t *out, int32_t *lead)
{
    const int bits = orc_arith_bits(p);
    uint32_t **dm = (uint32_t **)malloc((size_t)(m > 0 ? m : 1) * sizeof(uint32_t *));
    int32_t nrows = 0;
    for (int32_t i = 0; i < m; ++i) {
        const uint32_t *src = dmin + (size_t)i * (size_t)n;
        int32_t nz = 0;
        for (int32_t j = 0; j < n; ++j) if (src[j] % (uint32_t)p) { nz = 1; break; }
        if (!nz) continue;
        uint32_t *r = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
        for (int32_t j = 0; j < n; ++j) r[j] = src[j] % (uint32_t)p;
        dm[nrows++] = r;
    }
    uint32_t **nps = dense_echelon(dm, nrows, n, p, bits, nthreads);
    free(dm);
    dense_interreduce(nps, n, p, bits);
    int32_t rank = 0;
    for (int32_t c = n - 1; c >= 0; --c) {
        if (nps[c] == NULL) continue;
        if (out) {
            uint32_t *o = out + (size_t)rank * (size_t)n;
            memset(o, 0, (size_t)c * sizeof(uint32_t));
            memcpy(o + c, nps[c], (size_t)(n - c) * sizeof(uint32_t));
        }
        if (lead) lead[rank] = c;
        ++rank;
        free(nps[c]);
    }
    free(nps);
    return rank;
}

/* a8: la_ff_32.c:1666-1736 -- emit the dense pivots as sparse rows in
 * decreasing lead column, column indices shifted by ncl. */
static int32_t dense_to_rows(uint32_t **nps, int32_t ncr, int32_t ncl,
                             orc_rows_t *out)
{
    int32_t np = 0;
    int64_t nnz = 0;
    for (int32_t c = 0; c < ncr; ++c) {
        if (nps[c] == NULL) continue;
        ++np;
        for (int32_t j = 0; j < ncr - c; ++j) nnz += nps[c][j] != 0;
    }
    out->np = np;
    out->rp = (int64_t *)malloc((size_t)(np + 1) * sizeof(int64_t));
    out->ci = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    out->cf = (uint32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(uint32_t));
    int64_t w = 0;
    int32_t r = 0;
    out->rp[0] = 0;
    for (int32_t c = ncr - 1; c >= 0; --c) {
        if (nps[c] == NULL) continue;
        for (int32_t j = 0; j < ncr - c; ++j) {
            if (nps[c][j] != 0) {
                out->ci[w] = ncl + c + j;
                out->cf[w] = nps[c][j];
                ++w;
            }
        }
        out->rp[++r] = w;
    }
    return np;
}

/* a2: la_ff_32.c:1804-1851 (exact_sparse_dense_linear_algebra_ff_32). */
int32_t orc_la_sparse_dense(const orc_mat_t *m, int32_t p, int32_t nthreads,
                            orc_rows_t *out)
{
    const int bits = orc_arith_bits(p);
    const int32_t ncr = m->nc - m->ncl;
    uint32_t **drs = NULL;
    const int64_t np = sparse_pass_impl(m, p, nthreads, (uint32_t **)&drs,
                                        NULL, NULL, NULL, NULL);
    if (np < 0) return -1;
    uint32_t **nps;
    if (np > 0) {
        nps = dense_echelon(drs, (int32_t)np, ncr, p, bits, nthreads);
        dense_interreduce(nps, ncr, p, bits);
    } else {
        nps = (uint32_t **)calloc((size_t)(ncr > 0 ? ncr : 1), sizeof(uint32_t *));
    }
    free(drs);
    const int32_t res = dense_to_rows(nps, ncr, m->ncl, out);
    for (int32_t c = 0; c < ncr; ++c) free(nps[c]);
    free(nps);
    return res;
}

/* ---------------------------------------------- a9/a10: sparse RREF */

/* la_ff_32.c:100-189 / :191-288 / la_ff_16.c:102-190: reduce the dense
 * row from column dpiv on by every published pivot (known ones left of
 * ncl, new ones right of it).  Returns the right part as a new sparse row
 * (ascending columns, not yet normalised) or NULL. */
static srow_t *reduce_by_known_pivots_sparse(int64_t *dr, int32_t nc,
                                             int32_t ncl, srow_t *const *pivs,
                                             int32_t dpiv, int64_t p, int bits)
{
    int64_t k = 0;
    for (int32_t i = dpiv; i < nc; ++i) {
        if (dr[i] != 0) dr[i] %= p;
        if (dr[i] == 0) continue;
        const srow_t *pv = load_piv(pivs, i);
        if (pv == NULL) { ++k; continue; }
        apply_sparse(dr, pv, dr[i], p, bits);
        dr[i] = 0;
    }
    if (k == 0) return NULL;
    int32_t cnt = 0;
    for (int32_t i = ncl; i < nc; ++i) {
        if (dr[i] != 0) {
            dr[i] %= p;
            if (dr[i] != 0) ++cnt;
        }
    }
    if (cnt == 0) return NULL;
    int32_t *col = (int32_t *)malloc((size_t)cnt * sizeof(int32_t));
    uint32_t *cf = (uint32_t *)malloc((size_t)cnt * sizeof(uint32_t));
    int32_t j = 0;
    for (int32_t i = ncl; i < nc; ++i) {
        if (dr[i] != 0) { col[j] = i; cf[j] = (uint32_t)dr[i]; ++j; }
    }
    srow_t *r = (srow_t *)malloc(sizeof(srow_t));
    r->len = cnt; r->col = col; r->cf = cf; r->owned = 1;
    return r;
}

static void scatter_srow(int64_t *dr, const srow_t *r)
{
    for (int32_t j = 0; j < r->len; ++j) dr[r->col[j]] = (int64_t)r->cf[j];
}

static void rows_from_list(srow_t **lst, int32_t n, orc_rows_t *out)
{
    int64_t nnz = 0;
    for (int32_t i = 0; i < n; ++i) nnz += lst[i]->len;
    out->np = n;
    out->rp = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    out->ci = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    out->cf = (uint32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(uint32_t));
    int64_t w = 0;
    out->rp[0] = 0;
    for (int32_t i = 0; i < n; ++i) {
        memcpy(out->ci + w, lst[i]->col, (size_t)lst[i]->len * sizeof(int32_t));
        memcpy(out->cf + w, lst[i]->cf, (size_t)lst[i]->len * sizeof(uint32_t));
        w += lst[i]->len;
        out->rp[i + 1] = w;
    }
}

/* a9: la_ff_32.c:945-1079 (exact_sparse_reduced_echelon_form_ff_32). */
int32_t orc_la_sparse(const orc_mat_t *m, int32_t p, int32_t nthreads,
                      orc_rows_t *out)
{
    const int bits = orc_arith_bits(p);
    const int32_t nc = m->nc, ncl = m->ncl;
    srow_t **pivs = (srow_t **)calloc((size_t)(nc > 0 ? nc : 1), sizeof(srow_t *));
    int32_t *upiv = (int32_t *)malloc((size_t)(m->nr > 0 ? m->nr : 1) * sizeof(int32_t));
    const int32_t nrl = split_rows(m, nc, pivs, upiv);
    if (nrl < 0) {
        for (int32_t i = 0; i < nc; ++i) srow_free(pivs[i]);
        free(pivs); free(upiv);
        return -1;
    }
    if (nthreads < 1) nthreads = 1;
    int64_t *dr = (int64_t *)malloc((size_t)nthreads * (size_t)(nc > 0 ? nc : 1) * sizeof(int64_t));
    const int chunk = nrl / nthreads / 100 == 0 ? 1 : nrl / nthreads / 100;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, chunk)
    for (int32_t i = 0; i < nrl; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t *drl = dr + (size_t)tid * (size_t)nc;
        memset(drl, 0, (size_t)nc * sizeof(int64_t));
        scatter_input(drl, m, upiv[i]);
        int32_t sc = m->ci[m->rp[upiv[i]]];
        for (;;) {
            srow_t *npiv = reduce_by_known_pivots_sparse(drl, nc, ncl, pivs, sc, p, bits);
            if (npiv == NULL) break;
            /* normalise before publishing (la_ff_32.c:1016-1022) */
            if (npiv->cf[0] != 1) normalize_u32((uint32_t *)npiv->cf, npiv->len, (int32_t)p);
            srow_t *expect = NULL;
            if (__atomic_compare_exchange_n(&pivs[npiv->col[0]], &expect, npiv, 0,
                                            __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE))
                break;
            sc = npiv->col[0];
            srow_free(npiv);
        }
    }
    free(upiv);
    for (int32_t i = 0; i < ncl; ++i) { srow_free(pivs[i]); pivs[i] = NULL; }
    /* interreduce new pivots right to left (la_ff_32.c:1046-1071) */
    srow_t **res = (srow_t **)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(srow_t *));
    int32_t np = 0;
    for (int32_t i = nc - 1; i >= ncl; --i) {
        if (pivs[i] == NULL) continue;
        memset(dr, 0, (size_t)nc * sizeof(int64_t));
        scatter_srow(dr, pivs[i]);
        srow_free(pivs[i]);
        pivs[i] = NULL;
        srow_t *r = reduce_by_known_pivots_sparse(dr, nc, ncl, pivs, i, p, bits);
        pivs[i] = r;
        res[np++] = r;
    }
    free(dr);
    rows_from_list(res, np, out);
    for (int32_t i = 0; i < nc; ++i) srow_free(pivs[i]);
    free(pivs);
    free(res);
    return np;
}

/* a12: la_ff_32.c:1949-2017 (interreduce_matrix_rows_ff_32). */
int32_t orc_interreduce(const orc_mat_t *m, int32_t p, orc_rows_t *out)
{
    const int bits = orc_arith_bits(p);
    const int32_t nc = m->nc;
    srow_t **pivs = (srow_t **)calloc((size_t)(nc > 0 ? nc : 1), sizeof(srow_t *));
    for (int32_t i = 0; i < m->nr; ++i) {
        if (m->rp[i + 1] == m->rp[i]) continue;
        const int32_t lead = m->ci[m->rp[i]];
        if (lead < 0 || lead >= nc) {
            for (int32_t c = 0; c < nc; ++c) srow_free(pivs[c]);
            free(pivs);
            return -1;
        }
        /* a later row with the same lead replaces an earlier one, as the
         * reference's pivs[] assignment does */
        srow_free(pivs[lead]);
        srow_t *r = (srow_t *)malloc(sizeof(srow_t));
        *r = input_row(m, i);
        pivs[lead] = r;
    }
    int64_t *dr = (int64_t *)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(int64_t));
    srow_t **res = (srow_t **)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(srow_t *));
    int32_t k = 0;
    for (int32_t i = nc - 1; i >= 0; --i) {
        if (pivs[i] == NULL) continue;
        memset(dr, 0, (size_t)nc * sizeof(int64_t));
        scatter_srow(dr, pivs[i]);
        srow_free(pivs[i]);
        pivs[i] = NULL;
        srow_t *r = reduce_by_known_pivots_sparse(dr, nc, 0, pivs, i, p, bits);
        /* the lead stays (no pivot of its own column while reducing), so r
         * is never NULL; rows of the input may be unnormalised only if the
         * caller broke the contract, normalise to keep the output monic */
        if (r->cf[0] != 1) normalize_u32((uint32_t *)r->cf, r->len, (int32_t)p);
        pivs[i] = r;
        res[k++] = r;
    }
    free(dr);
    /* ascending lead order, as mat->r[k--] fills it */
    for (int32_t a = 0, b = k - 1; a < b; ++a, --b) {
        srow_t *t = res[a]; res[a] = res[b]; res[b] = t;
    }
    rows_from_list(res, k, out);
    for (int32_t i = 0; i < nc; ++i) srow_free(pivs[i]);
    free(pivs);
    free(res);
    return k;
}

void orc_rows_free(orc_rows_t *r)
{
    if (!r) return;
    free(r->rp); free(r->ci); free(r->cf);
    r->rp = NULL; r->ci = NULL; r->cf = NULL; r->np = 0;
}
