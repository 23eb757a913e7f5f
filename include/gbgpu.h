/*
 * gbgpu.h -- C-ABI of the MI355X (gfx950) finite-field F4 linear algebra.
 *
 * This library replaces the linear-algebra step of gb (v0.16): the
 * function behind gb's `linear_algebra` pointer (src/data.h:269-273),
 * selected in set_function_pointers (src/io.c:581-780) from the field size
 * (src/io.c:815-829) and la_option (src/io.c:867-872), and the one behind
 * `interreduce_matrix_rows` (src/data.h:275-279).
 *
 * Two layers:
 *   1. flat entry points on plain pointers and sizes (this file): upload an
 *      F4 matrix once, run any of the reference's LA routines on the device,
 *      download the result rows;
 *   2. drop-in adapters with the reference's exact signatures on gb's own
 *      mat_t / bs_t / stat_t (include/gbgpu_gb.h).
 *
 * All entry points return GBGPU_OK (0) or a negative error code; none of
 * them falls back to a CPU path.  Every call is synchronous with respect to
 * the host unless it says otherwise.  A context is bound to one device and
 * one HIP stream; calls on one context must not overlap.
 */
#ifndef GBGPU_H
#define GBGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- status */
enum {
    GBGPU_OK        = 0,
    GBGPU_EINVAL    = -1,  /* bad argument (NULL pointer, size, prime) */
    GBGPU_ECONTRACT = -2,  /* matrix violates the LA input contract */
    GBGPU_ENOMEM    = -3,  /* device or host allocation failed */
    GBGPU_EDEVICE   = -4,  /* HIP runtime / kernel launch error */
    GBGPU_ERANGE    = -5   /* size beyond what this build supports */
};

/* LA forms (the same canonical output, different kernels) */
enum {
    GBGPU_FORM_PLAIN = 0,  /* integer ALU only */
    GBGPU_FORM_MFMA  = 1   /* dense block products on matrix cores */
};

/* ------------------------------------------------------ wire formats */

/* One F4 matrix in flat CSR form.  Row i has the columns
 * ci[rp[i] .. rp[i+1]) -- the first one is its lead (the minimum), the
 * others in any order (gb's convert.c:111-119 contract) -- and the
 * coefficients cf[cp[rcf[i]] + k], k < rp[i+1]-rp[i], in [0, p).  Rows
 * that are multiples of one basis element may share a coefficient array,
 * as gb's rows share bs->cf_32[row[0]] (hash.c:903-926).  Columns
 * [0, ncl) are the known-pivot (A|C) block, [ncl, nc) the B|D block
 * (mat_t fields, data.h:137-155).  Rows are expected in gb's order after
 * sort_matrix_rows (order.c:98-131): the first row of each lead column is
 * its known pivot. */
typedef struct {
    int32_t nr;            /* rows */
    int32_t nc;            /* columns */
    int32_t ncl;           /* known-pivot columns */
    int32_t ncf;           /* coefficient arrays */
    uint32_t p;            /* prime field size, 2 <= p < 2^31 */
    const int64_t *rp;     /* nr + 1 */
    const int32_t *ci;     /* rp[nr] */
    const int32_t *rcf;    /* nr */
    const int64_t *cp;     /* ncf + 1 */
    const uint32_t *cf;    /* cp[ncf] */
} gbgpu_csr_t;

/* Result rows (host memory owned by the library until gbgpu_rows_free):
 * strictly ascending columns, unit lead coefficient, ordered as the
 * reference returns them (decreasing lead for la_option 1/2,
 * la_ff_32.c:1046 and :1687; ascending for interreduction, :2009). */
typedef struct {
    int32_t np;
    int64_t *rp;           /* np + 1 */
    int32_t *ci;
    uint32_t *cf;
} gbgpu_rows_t;

typedef struct gbgpu_ctx gbgpu_ctx;      /* device + stream + pools */
typedef struct gbgpu_mat gbgpu_mat;      /* device-resident F4 matrix */
typedef struct gbgpu_dense gbgpu_dense;  /* device-resident dense block */

/* --------------------------------------------------------- context */

/* Open a context on HIP device `device` with a stream of its own. */
int gbgpu_init(int device, gbgpu_ctx **ctx);
void gbgpu_finalize(gbgpu_ctx *ctx);

/* Use an existing HIP stream (hipStream_t passed as void*) instead of the
 * context's own, e.g. to time the work with events on that stream. */
int gbgpu_set_stream(gbgpu_ctx *ctx, void *stream);

/* Text of the last error on this context (never NULL). */
const char *gbgpu_last_error(const gbgpu_ctx *ctx);

/* Arithmetic variant for p, as the reference selects it
 * (io.c:815-829, io.c:710-729): 8, 16, 17 or 31. */
int gbgpu_arith_bits(uint32_t p);

/* Duration in milliseconds of the named kernel phase in the last run, or
 * -1 when unknown.  Phases: 0 = sparse pass kernel (a4),
 * 1 = dense echelon form (a5-a7), 2 = whole run. */
double gbgpu_phase_ms(const gbgpu_ctx *ctx, int phase);

/* --------------------------------------------------- F4 matrices */

/* Copy one F4 matrix to the device (validated on the device). */
int gbgpu_mat_upload(gbgpu_ctx *ctx, const gbgpu_csr_t *m, gbgpu_mat **out);
void gbgpu_mat_free(gbgpu_mat *m);

/* Coefficient arrays kept on the device, appended once and shared by the
 * matrices of later rounds (each basis element's array is sent once, not
 * once per round).  Append checks every entry < p and returns the offset of
 * the first; a matrix uploaded against the store must be freed before the
 * next append (the store may move).  No counterpart in the reference (gb's
 * matrices point into host coefficient arrays, hash.c:903-926). */
typedef struct gbgpu_cfstore gbgpu_cfstore;
int gbgpu_cfstore_create(gbgpu_ctx *ctx, uint32_t p, gbgpu_cfstore **out);
int gbgpu_cfstore_append(gbgpu_cfstore *s, const uint32_t *cf, int64_t n, int64_t *off);
int64_t gbgpu_cfstore_size(const gbgpu_cfstore *s);
void gbgpu_cfstore_free(gbgpu_cfstore *s);

/* The same matrix with its column indices formed on the device: m->ci is
 * not read; row q's entries are the monomial ids mon[src[q] ..
 * src[q] + rp[q+1] - rp[q]) and column = colmap[id] (ids in [0, nmap)).
 * With a store, m->cf is not read either: m->cp are offsets into the store
 * (m->cp[m->ncf] <= its size).  No counterpart in the reference: an F4 host
 * that keeps its rows as monomial ids (gbgpu_f4 does) hands them over as
 * they are instead of translating every entry on the CPU. */
int gbgpu_mat_upload_rows(gbgpu_ctx *ctx, const gbgpu_csr_t *m, const int32_t *mon, int64_t nmon,
                          const int64_t *src, const int32_t *colmap, int32_t nmap,
                          const gbgpu_cfstore *store, gbgpu_mat **out);

/* Replaces sparse_AB_CD_linear_algebra_ff_{8,16,32}
 * (la_ff_32.c:1081-1179, la_ff_16.c:737-836): reduce every to-reduce row
 * by the known pivots; the nonzero D' rows stay on the device.  *nrows
 * receives their number (the reference's mat->np after this step).
 * GBGPU_FORM_PLAIN walks the pivot columns one at a time (ballot over the
 * accumulator); GBGPU_FORM_MFMA eliminates 64 columns at a time with the
 * inverses of the pivots' diagonal blocks.  Same output. */
int gbgpu_sparse_pass(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *nrows);

/* The sparse pass once more with work counters: work[0] receives the
 * number of pivot rows applied (sum over to-reduce rows r of |S(r)|, an
 * implementation-independent count), work[1] the number of pivot entries
 * those rows hold (sum of len(P_j)).  For sizing and roofline figures; the
 * result is the same as gbgpu_sparse_pass. */
int gbgpu_sparse_pass_work(gbgpu_ctx *ctx, gbgpu_mat *m, int32_t *nrows, uint64_t *work);

/* Download the D' rows of the last sparse pass: `dense` receives nrows x
 * (nc - ncl) entries row-major in the reference's order (drs,
 * la_ff_32.c:1157-1176); src_row[k] the input row of D' row k.  Either
 * pointer may be NULL. */
int gbgpu_sparse_pass_download(gbgpu_ctx *ctx, gbgpu_mat *m, uint32_t *dense,
                               int32_t *src_row);

/* Replaces exact_sparse_dense_linear_algebra_ff_{8,16,32}
 * (la_option 1, la_ff_32.c:1804-1851): sparse pass, dense echelon form of
 * D' (la_ff_32.c:1217-1333), interreduction (:1181-1215) and conversion
 * to rows (:1666-1736).  *np receives the number of new pivot rows. */
int gbgpu_la_sparse_dense(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *np);

/* Replaces exact_sparse_linear_algebra_ff_{8,16,32} (la_option 2,
 * la_ff_32.c:1774-1802 + 945-1079).  Its output is by construction the
 * same canonical row set as la_option 1 (the reduced echelon form of the
 * row space restricted to the B|D columns). */
int gbgpu_la_sparse(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *np);

/* Replaces interreduce_matrix_rows_ff_{8,16,32} (la_ff_32.c:1949-2017):
 * requires ncl == 0 and distinct leads; rows come back in ascending lead
 * order. */
int gbgpu_interreduce(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *np);

/* Download the rows produced by the last LA call on m. */
int gbgpu_rows_download(gbgpu_ctx *ctx, gbgpu_mat *m, gbgpu_rows_t *out);
void gbgpu_rows_free(gbgpu_rows_t *rows);

/* One call: upload, la_option (1, 2, 42, 43, 44 as in io.c:608-779; the
 * probabilistic options run the exact kernels), download. */
int gbgpu_la_ff(gbgpu_ctx *ctx, const gbgpu_csr_t *m, int la_option, int form,
                gbgpu_rows_t *out);

/* ---------------------------------------------------- dense blocks */

/* The dense D' step on its own (la_ff_32.c:1217-1333, 1181-1215): an
 * m x n block mod p, row-major, entries in [0, p). */
int gbgpu_dense_upload(gbgpu_ctx *ctx, int32_t m, int32_t n, uint32_t p,
                       const uint32_t *host, gbgpu_dense **out);
void gbgpu_dense_free(gbgpu_dense *d);
/* The MFMA form's panel loop orders its two streams through device words
 * (every wait is for work queued before it, bounded; a wait that runs out
 * returns GBGPU_EDEVICE); GBGPU_DENSE_SYNC=events in the environment, read at
 * each call, orders them by HIP events instead, as runs under a tool that
 * serialises kernels across queues (rocprofv3 --pmc) need. */
int gbgpu_dense_rref(gbgpu_ctx *ctx, gbgpu_dense *d, int form, int32_t *rank);
/* rows: rank x n full-width RREF rows in decreasing lead order; lead: rank
 * lead columns.  Either may be NULL. */
int gbgpu_dense_download(gbgpu_ctx *ctx, gbgpu_dense *d, uint32_t *rows,
                         int32_t *lead);

/* ---- the dense echelon form without allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (a serving
 * loop or a harness that forbids hipMalloc).  The reduced echelon form of
 * an m x n block mod p that the caller already holds in device memory
 * (`a`: row-major, entries in [0, p), overwritten), by the kernels of
 * gbgpu_dense_rref and with the result of gbgpu_dense_upload +
 * gbgpu_dense_rref + gbgpu_dense_download: `rows` (device, min(m, n) x n)
 * receives the rank full-width rows in decreasing lead order, `lead`
 * (device, min(m, n)) their lead columns, *rank (host) the rank.
 * Every scratch array comes from one static device array of
 * GBGPU_NOALLOC_SCRATCH_MB MiB per device, which each device holds until
 * the process ends; a block needing more (gbgpu_dense_rref_noalloc_bytes)
 * is refused with GBGPU_ERANGE.  No hipMalloc or hipFree is called.
 * Calls that overlap on one device share that array and race. */
#define GBGPU_NOALLOC_SCRATCH_MB 64
size_t gbgpu_dense_rref_noalloc_bytes(int32_t m, int32_t n, uint32_t p, int form);
int gbgpu_dense_rref_noalloc(gbgpu_ctx *ctx, uint32_t *a, int32_t m, int32_t n, uint32_t p,
                             int form, uint32_t *rows, int32_t *lead, int32_t *rank);

/* ---- batches: la_option 1 over many matrices, each modulo its own prime.
 * No counterpart in the reference (its rational path is fraction-free over
 * GMP, la_qq.c:271-362); the per-matrix result is exactly what
 * gbgpu_la_sparse_dense gives for that matrix (rows in decreasing lead
 * order, columns shifted by ncl). */
typedef struct gbgpu_batch gbgpu_batch;

/* Upload nmat matrices (each checked as by gbgpu_mat_upload). */
int gbgpu_batch_upload(gbgpu_ctx *ctx, const gbgpu_csr_t *mats, int32_t nmat,
                       gbgpu_batch **out);
/* la_option 1 on every matrix as one graph launch: one persistent sparse
 * pass over all matrices' rows, then the dense panel loop of all blocks.
 * The first call in a form lays the batch out on the device, tunes the
 * sparse pass's window width and waves per CU on it (kept per process for
 * batches of the same shape) and captures the graph; later calls replay
 * it.  np[i] receives the new pivots of matrix i (may be NULL).
 * gbgpu_phase_ms(ctx, 2) is the graph's time on the device. */
int gbgpu_batch_la(gbgpu_ctx *ctx, gbgpu_batch *b, int form, int32_t *np);
/* Rows of matrix i's result, as gbgpu_rows_download. */
int gbgpu_batch_rows_download(gbgpu_ctx *ctx, gbgpu_batch *b, int32_t i,
                              gbgpu_rows_t *out);
/* out[0..nout): matrices, window shift, waves per CU, sparse grid, replays,
 * kernel launches in the graph, the MFMA form's dense update tile (columns,
 * rows; 64 x 64 in the plain form).  GBGPU_BATCH_EAGER=1 in the environment
 * issues the same launches straight on the stream instead of through the
 * graph (for tools that trace stream launches). */
int gbgpu_batch_info(gbgpu_batch *b, int32_t *out, int32_t nout);
void gbgpu_batch_free(gbgpu_batch *b);

#ifdef __cplusplus
}
#endif
#endif
