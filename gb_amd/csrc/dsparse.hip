// dsparse.hip -- the sparse pass on the matrix cores, for F4 matrices whose
// pivot rows are dense enough.
//
// Same result as the sparse pass (la_ff_32.c:1081-1179): each to-reduce row
// reduced by every known pivot, its right-hand part kept.  When the pivot
// rows carry a few per cent of the columns (dense quadratic systems reach
// ~5 %), nearly every row meets nearly every pivot, and the work is the
// dense product D - C A^-1 B.  This route does it as blocked elimination
// on a dense copy M of the to-reduce rows, one panel of 256 pivot columns
// (four groups of 64) after the other:
//
//   P     = the panel's pivot rows, dense from the panel on
//   Rp_a  = T_a^-1 P_a per group a (T_a the group's unit upper-triangular
//           block, the inverses of the blocked pass), then block back
//           substitution Rp_a -= Rp_a[:, group b] Rp_b (b > a), so that Rp
//           is the identity on all the panel's pivot columns
//   M    -= M[:, panel] Rp      over the columns right of the panel
//
// The update is the int8-digit MFMA kernel of the dense echelon form with
// K = 256 (four chunks of 64 pivot slots), so that each pass over M covers
// four groups.
// Columns without a pivot get identity rows in T_g and zero rows in P_g,
// so a row's entry there is left alone, as the reference skips them
// (la_ff_32.c:517-519); only the right-hand part is kept.
#include "internal.h"

namespace gbgpu {

constexpr int GB = 64;  // pivot columns per group

// M[q][c] = entry of to-reduce row upiv[q] at column c (M zeroed before)
__global__ void k_ds_scatter(int32_t nrl, const int32_t *upiv, const int64_t *rp,
                             const int32_t *ci, const int32_t *rcf, const int64_t *cp,
                             const uint32_t *cf, uint32_t *M, int64_t ldm) {
    const int lane = threadIdx.x & 63;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nrl; q += step) {
        const int32_t row = upiv[q];
        const int64_t b = rp[row], e = rp[row + 1];
        const uint32_t *c = cf + cp[rcf[row]] - b;
        uint32_t *out = M + (int64_t)q * ldm;
        for (int64_t k = b + lane; k < e; k += 64) out[ci[k]] = c[k];
    }
}

// P[s][c - g0] = entry of the pivot of column g0 + s at column c >= g0, for
// s < rows (zero rows for columns without a pivot); P zeroed before
__global__ void k_ds_group_rows(int32_t g0, int32_t ncl, int rows, const int32_t *first,
                                const int64_t *rp, const int32_t *ci, const int32_t *rcf,
                                const int64_t *cp, const uint32_t *cf, uint32_t *P, int64_t ldp) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (s >= rows) return;
    const int32_t col = g0 + s;
    if (col >= ncl) return;
    const int32_t row = first[col];
    if (row == INT32_MAX) return;
    const int64_t b = rp[row], e = rp[row + 1];
    const uint32_t *c = cf + cp[rcf[row]] - b;
    uint32_t *out = P + (int64_t)s * ldp - g0;
    for (int64_t k = b + lane; k < e; k += 64) out[ci[k]] = c[k];
}

// Rp[t][c] = sum_s Tinv[t][s] P[s][c] for the columns c of [g0, n) (c
// relative to g0): a block covers 64 columns, each wave 16 rows t; sums in
// 128 bits, reduced once
__global__ __launch_bounds__(256) void k_ds_rp(int32_t g0, int32_t n, const uint32_t *tinv,
                                               const uint32_t *P, int64_t ldp, uint32_t *Rp,
                                               int64_t ldr, Mod md, uint32_t r64) {
    __shared__ uint32_t T[GB][GB + 1];
    __shared__ uint32_t V[GB][GB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t jb = blockIdx.x * 64;
    for (int i = tid; i < GB * GB; i += 256) T[i >> 6][i & 63] = tinv[i];
    for (int i = tid; i < GB * GB; i += 256) {
        const int s_ = i >> 6, c = i & 63;
        V[s_][c] = (g0 + jb + c < n) ? P[(int64_t)s_ * ldp + jb + c] : 0u;
    }
    __syncthreads();
    const int32_t j = jb + lane;
    if (g0 + j >= n) return;
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
        const int t = 16 * wv + q;
        uint64_t lo = 0;
        uint32_t hi = 0;
        for (int s_ = 0; s_ < GB; ++s_) {
            const uint64_t pr = (uint64_t)T[t][s_] * (uint64_t)V[s_][lane];
            const uint64_t o = lo;
            lo += pr;
            hi += lo < o;
        }
        uint32_t r = md.reduce(lo);
        if (hi) r = (uint32_t)((r + (uint64_t)md.mul(hi, r64)) % md.p);
        Rp[(int64_t)t * ldr + j] = r;
    }
}

// block back substitution inside a panel: X[t][c] -= sum_k C[t][k] Y[k][c]
// for the 64 rows t of one group, K <= 192 rows k of the later groups
// (already final), over the columns c < w
constexpr int KB = 192;
__global__ __launch_bounds__(256) void k_ds_backsub(uint32_t *X, int64_t ldr, const uint32_t *C,
                                                    int K, const uint32_t *Y, int32_t w, Mod md,
                                                    uint32_t r64) {
    __shared__ uint32_t Cs[GB][KB + 1];
    __shared__ uint32_t Ys[KB][GB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t jb = blockIdx.x * 64;
    for (int i = tid; i < GB * K; i += 256) Cs[i / K][i % K] = C[i];
    for (int i = tid; i < K * 64; i += 256) {
        const int k = i >> 6, c = i & 63;
        Ys[k][c] = (jb + c < w) ? Y[(int64_t)k * ldr + jb + c] : 0u;
    }
    __syncthreads();
    const int32_t j = jb + lane;
    if (j >= w) return;
    for (int q = 0; q < 16; ++q) {
        const int t = 16 * wv + q;
        uint64_t lo = 0;
        uint32_t hi = 0;
        for (int k = 0; k < K; ++k) {
            const uint64_t pr = (uint64_t)Cs[t][k] * (uint64_t)Ys[k][lane];
            const uint64_t o = lo;
            lo += pr;
            hi += lo < o;
        }
        uint32_t r = md.reduce(lo);
        if (hi) r = (uint32_t)((r + (uint64_t)md.mul(hi, r64)) % md.p);
        uint32_t *x = X + (int64_t)t * ldr + j;
        *x = *x >= r ? *x - r : *x + md.p - r;
    }
}

// C[t][k] = X[t][c0 + k] for the 64 rows of one group (the coefficients of
// the back substitution, copied before X changes)
__global__ void k_ds_coef(const uint32_t *X, int64_t ldr, int32_t c0, int K, uint32_t *C) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < GB * K; i += gridDim.x * blockDim.x)
        C[i] = X[(int64_t)(i / K) * ldr + c0 + i % K];
}

// digit planes of Rp^T: column c holds 64 kc bytes per digit, byte t = row t
__global__ void k_ds_rpdigits(const uint32_t *Rp, int64_t ldr, int rows, int32_t w, int kc,
                              uint32_t p, Digits dg) {
    const int lane = threadIdx.x & 63;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < w; c += step) {
        for (int t = lane; t < 64 * kc; t += 64) {
            const uint32_t v = t < rows ? Rp[(int64_t)t * ldr + c] : 0u;
            for (int i = 0; i < dg.nd; ++i)
                dg.R[i * dg.rplane + (int64_t)c * (64 * kc) + t] = (int8_t)digit_word(v, p, dg.hi, i);
        }
    }
}

// L = M[q][g0 .. g0 + bw) as digit planes of 64 kc bytes per row;
// active[q] = any nonzero
__global__ void k_ds_gather(int32_t nrl, int32_t g0, int32_t bw, int kc, const uint32_t *M,
                            int64_t ldm, int32_t *active, uint32_t p, Digits dg) {
    const int lane = threadIdx.x & 63;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nrl; q += step) {
        bool nz = false;
        for (int t = lane; t < 64 * kc; t += 64) {
            const uint32_t v = t < bw ? M[(int64_t)q * ldm + g0 + t] : 0u;
            nz = nz || v != 0u;
            for (int i = 0; i < dg.nd; ++i)
                dg.L[i * dg.lplane + (int64_t)q * (64 * kc) + t] = (int8_t)digit_word(v, p, dg.hi, i);
        }
        active[q] = __any(nz) ? 1 : 0;  // every lane stores the same flag
    }
}

// D'[q] = M[q][ncl .. nc), dnz[q] = row nonzero
__global__ void k_ds_extract(int32_t nrl, int32_t ncl, int32_t nc, const uint32_t *M,
                             int64_t ldm, uint32_t *dprime, int32_t *dnz) {
    const int lane = threadIdx.x & 63;
    const int32_t ncr = nc - ncl;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nrl; q += step) {
        const uint32_t *src = M + (int64_t)q * ldm + ncl;
        uint32_t *dst = dprime + (int64_t)q * ncr;
        bool nz = false;
        for (int32_t c = lane; c < ncr; c += 64) {
            const uint32_t v = src[c];
            dst[c] = v;
            nz = nz || v != 0u;
        }
        dnz[q] = __any(nz) ? 1 : 0;  // every lane stores the same flag
    }
}

static int grid_rows(int64_t rows, int per_block) {
    int64_t g = (rows + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

bool dense_route_pays(const gbgpu_mat *m) {
    // pivot rows' mean length against the width: dense work ~ nc / 2 per
    // pivot and row, sparse work ~ the pivot's length; the matrix cores
    // are worth it from ~1/40 density on
    if (m->ncl < GB || m->nrl < 64) return false;
    const double mean_len = (double)m->nbucket / (double)m->ncl;
    return mean_len * 40.0 > (double)m->nc;
}

int run_dense_route(gbgpu_mat *m, const uint32_t *tinv) {
    gbgpu_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    constexpr int KCH = 4;               // groups per panel
    constexpr int PW = GB * KCH;         // pivot columns per panel
    const int32_t nrl = m->nrl, ncl = m->ncl, nc = m->nc;
    const int64_t ldm = ((int64_t)nc + 63) & ~63ll;
    const int64_t ldp = ldm, ldr = ldm;
    const Mod md = Mod::make(m->p);
    const uint32_t r64 = (uint32_t)((((unsigned __int128)1) << 64) % m->p);
    Digits dg;
    digits_setup(dg, m->p, digits_for(m->p));
    uint32_t *M = nullptr, *P = nullptr, *Rp = nullptr, *C = nullptr;
    int8_t *dl = nullptr, *dr = nullptr;
    int32_t *active = nullptr, *pk = nullptr;  // pk: a panel record whose k is 64
    int rc;
    auto cleanup = [&]() {
        dfree(M); dfree(P); dfree(Rp); dfree(C); dfree(dl); dfree(dr); dfree(active); dfree(pk);
    };
    if ((rc = dalloc(ctx, &M, (size_t)nrl * ldm)) || (rc = dalloc(ctx, &P, (size_t)PW * ldp)) ||
        (rc = dalloc(ctx, &Rp, (size_t)PW * ldr)) || (rc = dalloc(ctx, &C, (size_t)GB * KB)) ||
        (rc = dalloc(ctx, &dl, (size_t)dg.nd * nrl * PW)) ||
        (rc = dalloc(ctx, &dr, (size_t)dg.nd * ldr * PW)) || (rc = dalloc(ctx, &active, (size_t)nrl)) ||
        (rc = dalloc(ctx, &pk, (dense_panel_bytes() + 3) / 4))) {
        cleanup();
        return rc;
    }
    dg.L = dl;
    dg.R = dr;
    dg.lplane = (int64_t)nrl * PW;
    dg.rplane = ldr * PW;
    const int32_t k64 = GB;
    GBGPU_HIP(ctx, hipMemcpyAsync(pk, &k64, sizeof(int32_t), hipMemcpyHostToDevice, s));
    GBGPU_HIP(ctx, hipMemsetAsync(M, 0, (size_t)nrl * ldm * sizeof(uint32_t), s));
    k_ds_scatter<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, m->upiv, m->rp, m->ci, m->rcf, m->cp, m->cf,
                                                    M, ldm);
    const Panel *pan = (const Panel *)pk;
    for (int32_t g0 = 0; g0 < ncl; g0 += PW) {
        const int32_t bw = (ncl - g0) < PW ? (ncl - g0) : PW;  // pivot columns of the panel
        const int na = (bw + GB - 1) / GB;                     // its groups
        const int32_t w = nc - g0;                             // columns from the panel on
        const int nb = (w + 63) / 64;
        GBGPU_HIP(ctx, hipMemsetAsync(P, 0, (size_t)PW * ldp * sizeof(uint32_t), s));
        k_ds_group_rows<<<PW / 4, 256, 0, s>>>(g0, ncl, PW, m->first, m->rp, m->ci, m->rcf, m->cp,
                                               m->cf, P, ldp);
        // each group's rows made unit on its own pivot columns ...
        for (int a = 0; a < na; ++a)
            k_ds_rp<<<nb, 256, 0, s>>>(g0, nc, tinv + (int64_t)(g0 / GB + a) * 4096,
                                       P + (int64_t)a * GB * ldp, ldp, Rp + (int64_t)a * GB * ldr, ldr,
                                       md, r64);
        // ... then zero on the later groups' pivot columns (back substitution)
        for (int a = na - 2; a >= 0; --a) {
            const int K = GB * (na - 1 - a);
            uint32_t *X = Rp + (int64_t)a * GB * ldr;
            k_ds_coef<<<(GB * K + 255) / 256, 256, 0, s>>>(X, ldr, (a + 1) * GB, K, C);
            k_ds_backsub<<<nb, 256, 0, s>>>(X, ldr, C, K, Rp + (int64_t)(a + 1) * GB * ldr, w, md, r64);
        }
        k_ds_rpdigits<<<grid_rows(w, 4), 256, 0, s>>>(Rp, ldr, na * GB, w, KCH, m->p, dg);
        k_ds_gather<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, g0, bw, KCH, M, ldm, active, m->p, dg);
        // columns right of the panel's pivot columns (the last panel may be
        // partial: its right-hand columns are updated too)
        const int32_t cb = g0 + bw, ce = nc;
        if (cb < ce) {
            dim3 gr((unsigned)((ce - cb + 63) / 64), (unsigned)((nrl + 63) / 64));
            launch_update_i8(dg.nd, KCH, gr, s, M, ldm, nrl, cb, ce, g0, pan, active, md, dg);
        }
        GBGPU_HIP(ctx, hipGetLastError());
    }
    k_ds_extract<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, ncl, nc, M, ldm, m->dprime, m->dnz);
    GBGPU_HIP(ctx, hipGetLastError());
    cleanup();
    return GBGPU_OK;
}

}  // namespace gbgpu
