// internal.h -- device-resident objects behind the opaque C-ABI handles.
#pragma once

#include <climits>

#include "common.h"

// One F4 matrix on the device plus everything derived from it.
struct gbgpu_mat {
    gbgpu_ctx *ctx = nullptr;
    int32_t nr = 0, nc = 0, ncl = 0, ncf = 0;
    uint32_t p = 0;
    int64_t nnz = 0, npool = 0;
    // wire format, as uploaded
    int64_t *rp = nullptr;
    int32_t *ci = nullptr;
    int32_t *rcf = nullptr;
    int64_t *cp = nullptr;
    uint32_t *cf = nullptr;
    bool cf_borrowed = false;   // cf belongs to a gbgpu_cfstore
    // known-pivot split (first row per lead = pivot, la_ff_32.c:1106-1116)
    int32_t split_lim = -1;     // leads were checked against [0, split_lim)
    int32_t *first = nullptr;   // nc: first row with that lead, INT32_MAX if none
    int32_t *upiv = nullptr;    // to-reduce rows in the reference's order
    int32_t nrl = 0;
    // pivot rows bucketed by column window (sparse pass)
    int32_t wshift = 0, nwin = 0;
    int64_t *seg = nullptr;     // (nwin + 1) x ncl, window-major
    uint2 *bucket = nullptr;    // (column, coefficient), leads excluded
    uint32_t *pbits = nullptr;  // has-pivot bit per left column
    int64_t nbucket = 0;
    // the same entries grouped per (window, 64-pivot group) for the blocked
    // pass: gseg[w * ngroups + g] .. gseg[w * ngroups + g + 1], each entry
    // x = (pivot index in group << 24) | column, y = coefficient
    int64_t *gseg = nullptr;
    uint2 *gbucket = nullptr;
    // sparse pass output
    uint32_t *dprime = nullptr; // nrl x ncr, row q = upiv[q]'s remainder
    int32_t *dnz = nullptr;     // nrl: 1 if row q is nonzero
    int32_t *dlist = nullptr;   // nd: nonzero rows q, in order
    int32_t nd = 0;
    // result rows (device)
    int64_t *orp = nullptr;
    int32_t *oci = nullptr;
    uint32_t *ocf = nullptr;
    int32_t onp = 0;
    int64_t onnz = 0;
    ~gbgpu_mat();
};

// A dense block and the state of its echelon form.
struct gbgpu_dense {
    gbgpu_ctx *ctx = nullptr;
    int32_t m = 0, n = 0;
    int64_t ld = 0;
    uint32_t p = 0;
    uint32_t *a = nullptr;      // m x ld, row-major
    bool own_a = true;
    // result: rank rows, decreasing lead
    int32_t rank = 0;
    int32_t *orow = nullptr;    // rank: row of a holding the k-th output row
    int32_t *olead = nullptr;   // rank: its lead column
    ~gbgpu_dense();
};

namespace gbgpu {

// sparse.hip
struct SpArgs {
    int32_t nc, ncl, ncr, nrl, nwin;
    const int64_t *rp;
    const int32_t *ci;
    const int32_t *rcf;
    const int64_t *cp;
    const uint32_t *cf;
    const int32_t *upiv;
    const uint32_t *pbits;
    const int64_t *seg;
    const uint2 *bucket;
    uint32_t *xs;       // gridDim.x x ncl multipliers
    uint32_t *xs2;      // their Shoup quotients (blocked pass, 32-bit accumulators)
    uint32_t *dprime;   // nrl x ncr
    int32_t *dnz;       // nrl
    int32_t *queue;     // work counter
    unsigned long long *work;  // COUNT only: [pivots applied, pivot entries]
    Mod md;
};

// what the blocked pass needs of one matrix
struct BlockedArgs {
    SpArgs a;
    const uint32_t *tinv;
    const int64_t *gseg;
    const uint2 *gbucket;
    int32_t ngroups;
};

int prep_split(gbgpu_mat *m, int32_t lim);
// force_wshift: window of 2^force_wshift columns (0: the smallest that fits)
int prep_buckets(gbgpu_mat *m, int force_wshift = 0);
// acc_shift: log2 of the accumulator size the entries' offsets are scaled by
int prep_group_buckets(gbgpu_mat *m, int acc_shift = 3);
// group buckets and the inverses of the pivot groups' diagonal blocks
int prep_blocked(gbgpu_mat *m, uint32_t **tinv);
// the kernel arguments of m (xs, queue, work left null)
void sparse_args(const gbgpu_mat *m, SpArgs *a);
// LDS bytes of one wave of the sparse pass (occupancy)
int sparse_wave_lds(int wshift, int form);
// one persistent launch over the to-reduce rows of nmat matrices
int batch_sparse_launch(gbgpu_ctx *ctx, const BlockedArgs *mats, const int32_t *row_off,
                        int32_t nmat, int32_t *queue, uint32_t *xs, int64_t xs_stride, int wshift,
                        bool big, int form, int grid);
// work (optional, 2 values): pivots applied and pivot entries streamed,
// counted by an instrumented launch
// form: GBGPU_FORM_PLAIN = ballot chain, GBGPU_FORM_MFMA = blocked (group
// inverses); the work counters always use the plain chain
int run_sparse_pass(gbgpu_mat *m, uint64_t *work = nullptr, int form = GBGPU_FORM_PLAIN);

// dense.hip: reduced echelon form of d (in place); form = GBGPU_FORM_*
int dense_rref(gbgpu_dense *d, int form);

struct Panel;
struct ListState;
// the device scratch of one echelon form
struct DenseWs {
    int32_t *rowpiv, *colpiv, *active, *cand;  // m, n, m, m
    uint32_t *L, *rp;                          // m x 64, 64 x roundup(n, 64)
    int8_t *dl, *dr;                           // digit planes (MFMA form)
    Panel *pan;                                // two panel records
    ListState *st;
    int32_t *alive;                            // m: the prune's flags
};
// bytes of that scratch for an m x n block (pieces 256-byte aligned)
size_t dense_ws_bytes(int32_t m, int32_t n, int form, uint32_t p);
// lay the pieces out from base (dense_ws_bytes of room)
void dense_ws_carve(uint8_t *base, int32_t m, int32_t n, int form, uint32_t p, DenseWs *ws);
// the panel loop on scratch ws, no allocation: leaves ws.colpiv[c] = the
// pivot row of column c (-1: none); times itself with ctx->ev[2], ev[3]
int dense_rref_ws(gbgpu_dense *d, int form, const DenseWs &ws);
// after a wait for the main stream: GBGPU_EDEVICE if a wait between the dense
// panel loop's streams ran out of its bound since the loop began
int loop_sync_check(gbgpu_ctx *ctx);
// d->rank, d->orow, d->olead from the column -> pivot row map of an
// eliminated block (decreasing lead order)
int dense_output(gbgpu_dense *d, const int32_t *colpiv);

// Signed base-256 digits for the int8 matrix cores.  With nd digits in
// [-128, 127] the integers 127 S .. -128 S are reachable (S = 1 + 256 + ...
// + 256^(nd-1)), a window of 256^nd consecutive integers; so for
// p <= 256^nd every residue x has a representative in it: x itself when
// x <= 127 S, else x - p.  nd = 1 for p <= 255, 2 for p <= 65521, 3 for
// p <= 2^24, 4 below 2^31.  A plane holds one digit of the left operand L
// (rows) or of the right operand transposed (columns: Rp in the dense echelon
// form, the raw pivot rows P in the sparse pass's dense route) in chunks of
// 64 pivot slots, chunk-major:
// [chunk][row or column][64 bytes], so that one lane's 16 consecutive
// bytes are an MFMA operand and 16 rows or columns of a chunk are 1 KiB
// of consecutive memory.
struct Digits {
    int8_t *L = nullptr;   // nd planes of nchunks x m x 64
    int8_t *R = nullptr;   // nd planes of nchunks x ldr x 64
    int64_t lplane = 0;    // bytes per L plane
    int64_t rplane = 0;    // bytes per R plane
    int64_t lchunk = 0;    // bytes per chunk of an L plane (m x 64)
    int64_t rchunk = 0;    // bytes per chunk of an R plane (ldr x 64)
    int nd = 0;            // 0: no digit planes (plain form)
    uint32_t hi = 0;       // 127 S
    int32_t k0 = 0;        // first chunk an update reads
    uint64_t off = 0;      // multiple of p above 2^57: makes a combined sum positive
    uint64_t off2 = 0;     // multiple of p above 2^41 (the high part for nd = 4)
    uint32_t w32 = 0;      // 2^32 mod p
    uint32_t w32s = 0;     // floor(w32 2^32 / p): Shoup's quotient for products by w32
    float pinvf = 0;       // 1 / p in single precision (digits_balanced2, nd = 2)
    double pinvd = 0;      // 1 / p in double precision (mod_f64, nd >= 3)
};

// nd = 2 stops at 65521, the largest prime below 2^16, where
// digits_balanced2 is exact; the composite moduli above it take nd = 3
__host__ inline int digits_for(uint32_t p) {
    return p <= 255u ? 1 : p <= 65521u ? 2 : p <= (1u << 24) ? 3 : 4;
}

__host__ inline void digits_setup(Digits &g, uint32_t p, int nd) {
    g.nd = nd;
    uint64_t S = 0, pw = 1;
    for (int i = 0; i < nd; ++i, pw *= 256) S += pw;
    g.hi = (uint32_t)(127 * S);
    g.w32 = (uint32_t)((1ull << 32) % p);
    g.off = ((1ull << 57) / p + 1) * p;
    g.off2 = ((1ull << 41) / p + 1) * p;
    g.w32s = (uint32_t)(((uint64_t)g.w32 << 32) / p);
    g.pinvf = 1.0f / (float)p;
    g.pinvd = 1.0 / (double)p;
}

__device__ __forceinline__ uint32_t digit_word(uint32_t x, uint32_t p, uint32_t hi, int i) {
    // digit i of x's representative (as an unsigned byte)
    int64_t v = x <= hi ? (int64_t)x : (int64_t)x - (int64_t)p;
    int8_t d = 0;
    for (int k = 0; k <= i; ++k) {
        d = (int8_t)(uint8_t)(v & 0xFF);
        v = (v - d) >> 8;
    }
    return (uint8_t)d;
}

typedef int v4i __attribute__((ext_vector_type(4)));

// x mod p for |x| < 2^52: x is exact in double precision, the quotient
// floor(x / p) from the double product x (1/p) is off by at most one (the
// product's relative error is below 2^-52, so its absolute error is below
// one), and the remainder x - q p, exact in 64 bits, is brought into [0, p)
// by one correction either way.  Full-rate double operations and one 64-bit
// product instead of a 64-bit Barrett reduction.
__device__ __forceinline__ uint32_t mod_f64(int64_t x, uint32_t p, double pinv) {
    const double q = __builtin_floor((double)x * pinv);
    int64_t r = x - (int64_t)q * (int64_t)p;
    r += r < 0 ? (int64_t)p : 0;
    r -= r >= (int64_t)p ? (int64_t)p : 0;
    return (uint32_t)r;
}

// sum_s Cs 256^s mod p for result element e of the accumulators.  Over up
// to 256 pivot slots |Cs| <= min(s + 1, 2 nd - 1 - s) 256 128^2 < 2^24.
//   nd <= 2: the shifted sum is below 2^41, one reduction (dg.off, a
//     multiple of p above 2^57, makes it positive);
//   nd = 3: hi = C3 + 2^8 C4 (below 2^33) is reduced first, then
//     C0 + 2^8 C1 + 2^16 C2 + (hi mod p) 2^24 (below 2^49, p < 2^24);
//   nd = 4: hi = C4 + 2^8 C5 + 2^16 C6 (below 2^41) is reduced first and
//     multiplied by 2^32 mod p with Shoup's quotient (a result below 2p),
//     then C0 + 2^8 C1 + 2^16 C2 + 2^24 C3 (below 2^49) plus that.
// The reductions of nd >= 3 run in double precision (mod_f64).
template <int ND>
__device__ __forceinline__ uint32_t digits_combine(const v4i *acc, int e, const Mod &md,
                                                   const Digits &dg) {
    if constexpr (ND <= 2) {
        int64_t lo = 0;
#pragma unroll
        for (int q = 0; q < 2 * ND - 1; ++q) lo += (int64_t)acc[q][e] << (8 * q);
        return md.reduce((uint64_t)(lo + (int64_t)dg.off));
    } else if constexpr (ND == 3) {
        const int64_t hi = (int64_t)acc[3][e] + ((int64_t)acc[4][e] << 8);
        const uint32_t hr = mod_f64(hi, md.p, dg.pinvd);
        const int64_t lo = (int64_t)acc[0][e] + ((int64_t)acc[1][e] << 8) +
                           ((int64_t)acc[2][e] << 16) + ((int64_t)hr << 24);
        return mod_f64(lo, md.p, dg.pinvd);
    } else {
        const int64_t hi = (int64_t)acc[4][e] + ((int64_t)acc[5][e] << 8) + ((int64_t)acc[6][e] << 16);
        const uint32_t hr = mod_f64(hi, md.p, dg.pinvd);
        // hr (2^32 mod p) mod p, below 2p, by Shoup's precomputed quotient
        const uint32_t q = __umulhi(hr, dg.w32s);
        const uint32_t t = hr * dg.w32 - q * md.p;
        const int64_t lo = (int64_t)acc[0][e] + ((int64_t)acc[1][e] << 8) +
                           ((int64_t)acc[2][e] << 16) + ((int64_t)acc[3][e] << 24) + (int64_t)t;
        return mod_f64(lo, md.p, dg.pinvd);
    }
}

// C0 + 2^8 C1 + 2^16 C2 (nd = 2, up to 256 pivot slots: |C0|, |C2| <= 2^22,
// |C1| <= 2^23) modulo p for 256 <= p <= 65521, as a residue in
// [-(p/2 + 1), p/2 + 1], with single-precision quotients and 24-bit
// multiplies (all full rate) instead of a 64-bit reduction.  Horner in base
// 256: a = C2, then a = 256 a' + C1, then a = 256 a' + C0, where a' is a
// reduced by q = rint(a / p).  Every a is below 2^24 in size, so it is exact
// in single precision: |a'| <= p/2 + 1 <= 32761 gives |256 a'| <= 8386816.
// The quotient is a (1/p) rounded to the nearest integer in one step: a
// fused multiply-add with 1.5 2^23, whose sum has a unit in the last place
// of one, leaves q in the low bits of its representation.  1/p is within
// 2^-24 of its value relative, so a (1/p) is within |a / p| 2^-24 < 1 / p
// of a / p and |a - q p| <= p/2 + 1.  The 24-bit products q p are exact
// (|q| < 2^17).
__device__ __forceinline__ int32_t rint_quot(int32_t a, float pinv) {
    return __float_as_int(__builtin_fmaf((float)a, pinv, 12582912.0f)) - 0x4B400000;
}

__device__ __forceinline__ int32_t digits_balanced2(const v4i *acc, int e, int32_t p, float pinv) {
    int32_t a = acc[2][e];
    a -= __mul24(rint_quot(a, pinv), p);
    a = a * 256 + acc[1][e];
    a -= __mul24(rint_quot(a, pinv), p);
    a = a * 256 + acc[0][e];
    return a - __mul24(rint_quot(a, pinv), p);
}

// (x - combined digit products) mod p for x in [0, p): the single-precision
// form for nd = 2 (256 <= p <= 65521), the 64-bit combination otherwise
template <int ND>
__device__ __forceinline__ uint32_t digits_sub(uint32_t x, const v4i *acc, int e, const Mod &md,
                                               const Digits &dg) {
    if constexpr (ND == 2) {
        // x - b lies in [-(p/2 + 1), 3p/2 + 1]: of it, it + p and it - p
        // as unsigned words the smallest is the residue
        const uint32_t u = x - (uint32_t)digits_balanced2(acc, e, (int32_t)md.p, dg.pinvf);
        return min(min(u, u + md.p), u - md.p);
    } else {
        const uint32_t u = digits_combine<ND>(acc, e, md, dg);
        return x >= u ? x - u : x + md.p - u;
    }
}

// the int8-digit MFMA update A[r][cb, ce) -= L[r] R (dense.hip) over kc
// chunks of 64 pivot slots (kc = 1 or 4) from chunk dg.k0: L and R^T as
// the digit planes of dg (R^T indexed from column c0), rows with
// active[r].  kc = 4 takes the kernel that walks column tiles with the
// block's R digits staged in LDS.  A is uint32_t, or uint16_t for
// p <= 65521 (nd <= 2).
template <typename T>
void launch_update_i8(int nd, int kc, hipStream_t s, T *a, int64_t ld, int32_t m, int32_t cb,
                      int32_t ce, int32_t c0, const Panel *pan, const int32_t *active, Mod md,
                      const Digits &dg);

// dsparse.hip: the sparse pass as blocked elimination on the matrix cores
bool dense_route_pays(const gbgpu_mat *m);
int run_dense_route(gbgpu_mat *m, const uint32_t *tinv);

// dense.hip, batches: one dense block of a batch and its elimination state
struct DenseDesc {
    uint32_t *a;
    int64_t ld;
    int32_t m, n;
    Mod md;
    uint32_t r64;              // 2^64 mod p
    int32_t *rowpiv, *colpiv;  // m, n
    int32_t *active, *cand;    // m, m
    uint32_t *L, *rp;          // m x 64, 64 x ldr
    int64_t ldr;
    Panel *pan;                // one panel record
    ListState *st;
    const int32_t *dnz;        // m: rows that take part (nullptr: all)
    Digits dg;                 // digit planes (MFMA form)
};
int batch_dense_init(gbgpu_ctx *ctx, const DenseDesc *D, int32_t nmat);
// block counts of step `step` for the four grids (rp, gather, update,
// prune), prefix sums over the matrices: out[4 * (nmat + 1)]; the MFMA
// form's update blocks cover ucols x urows of a block
void batch_dense_counts(const DenseDesc *hd, int32_t nmat, int32_t step, int form, int ucols,
                        int urows, int32_t *out);
// nd: digits of the MFMA form's update (the largest of the batch); its
// blocks cover ucols x urows (64 x 64 tiles, or strips of 256 or 512
// columns of 64 rows, or of 128 rows for nd <= 2)
int batch_dense_step(gbgpu_ctx *ctx, const DenseDesc *D, int32_t nmat, const int32_t *tab,
                     const int32_t *htab, int32_t step, int form, bool big, int nd, int ucols,
                     int urows);
bool batch_update_shape_ok(int nd, int cols, int rows);
size_t dense_panel_bytes();      // sizeof(Panel)
// a launch's grid, block and static LDS against the device's limits
int check_launch(gbgpu_ctx *ctx, const void *fn, int64_t grid, int block, const char *what);
size_t dense_liststate_bytes();  // sizeof(ListState)

// rows.hip: rows of a dense RREF to CSR (columns shifted by col_shift);
// with reverse the output order is increasing lead
int dense_to_rows(gbgpu_dense *d, int32_t col_shift, int reverse, int64_t **orp, int32_t **oci,
                  uint32_t **ocf, int64_t *onnz);

// rows i with keep[i] of m scattered densely into row dst_row[i] of a
// fresh nkeep x nc block
int scatter_rows(gbgpu_mat *m, const int32_t *keep, const int32_t *dst_row, int32_t nkeep,
                 gbgpu_dense **out);

// gather rows `list` (n of them) of src (row stride ld_src, ncols wide) into
// a fresh dense block
int gather_dense(gbgpu_ctx *ctx, const uint32_t *src, int64_t ld_src, const int32_t *list,
                 int32_t n, int32_t ncols, uint32_t p, gbgpu_dense **out);

}  // namespace gbgpu
