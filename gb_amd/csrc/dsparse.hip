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
//           block, the inverses of the blocked pass): identity on the
//           group's pivot columns
//   for a = 0..3:  L_a = M[:, group a] (current values)
//                  M[:, later groups of the panel] -= L_a Rp_a   (K = 64)
//   M[:, right of the panel] -= [L_0 .. L_3] [Rp_0; ..; Rp_3]    (K = 256)
//
// The updates are the int8-digit MFMA kernel of the dense echelon form; the
// wide one covers four groups per pass over M.
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

// digit planes of Rp^T: slot t of column c (zero from row `rows` on) at
// byte t % 64 of column c in chunk t / 64
__global__ void k_ds_rpdigits(const uint32_t *Rp, int64_t ldr, int rows, int32_t w, int kc,
                              uint32_t p, Digits dg) {
    const int lane = threadIdx.x & 63;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < w; c += step) {
        for (int t = lane; t < 64 * kc; t += 64) {
            const uint32_t v = t < rows ? Rp[(int64_t)t * ldr + c] : 0u;
            for (int i = 0; i < dg.nd; ++i)
                dg.R[i * dg.rplane + (t >> 6) * dg.rchunk + (int64_t)c * 64 + (t & 63)] =
                    (int8_t)digit_word(v, p, dg.hi, i);
        }
    }
}

// chunk a of L: M[q][c0 .. c0 + bw) (bw <= 64) as the digits of row q in
// chunk a; act[q] = any nonzero, any[q] = OR over the chunks
__global__ void k_ds_gather(int32_t nrl, int32_t c0, int32_t bw, int a, const uint32_t *M,
                            int64_t ldm, int32_t *act, int32_t *any, uint32_t p, Digits dg) {
    const int lane = threadIdx.x & 63;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nrl; q += step) {
        const uint32_t v = lane < bw ? M[(int64_t)q * ldm + c0 + lane] : 0u;
        for (int i = 0; i < dg.nd; ++i)
            dg.L[i * dg.lplane + a * dg.lchunk + (int64_t)q * 64 + lane] =
                (int8_t)digit_word(v, p, dg.hi, i);
        const int nz = __any(v != 0u) ? 1 : 0;
        act[q] = nz;  // every lane stores the same flags
        any[q] = a == 0 ? nz : (any[q] | nz);
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
    uint32_t *M = nullptr, *P = nullptr, *Rp = nullptr;
    int8_t *dl = nullptr, *dr = nullptr;
    int32_t *active = nullptr, *act = nullptr, *pk = nullptr;  // pk: a panel record whose k is 64
    int rc;
    auto cleanup = [&]() {
        dfree(M); dfree(P); dfree(Rp); dfree(dl); dfree(dr); dfree(active); dfree(act); dfree(pk);
    };
    if ((rc = dalloc(ctx, &M, (size_t)nrl * ldm)) || (rc = dalloc(ctx, &P, (size_t)PW * ldp)) ||
        (rc = dalloc(ctx, &Rp, (size_t)PW * ldr)) || (rc = dalloc(ctx, &act, (size_t)nrl)) ||
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
    dg.lchunk = (int64_t)nrl * 64;
    dg.rchunk = ldr * 64;
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
        // each group's rows made unit on its own pivot columns
        for (int a = 0; a < na; ++a)
            k_ds_rp<<<nb, 256, 0, s>>>(g0, nc, tinv + (int64_t)(g0 / GB + a) * 4096,
                                       P + (int64_t)a * GB * ldp, ldp, Rp + (int64_t)a * GB * ldr, ldr,
                                       md, r64);
        k_ds_rpdigits<<<grid_rows(w, 4), 256, 0, s>>>(Rp, ldr, na * GB, w, KCH, m->p, dg);
        // the panel's groups in order: the multipliers of group a are M's
        // current values on its columns; its rows update the panel's later
        // columns at once (K = 64) ...
        for (int a = 0; a < na; ++a) {
            const int32_t ca = g0 + a * GB;
            const int32_t wa = (bw - a * GB) < GB ? (bw - a * GB) : GB;
            k_ds_gather<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, ca, wa, a, M, ldm, act, active, m->p, dg);
            const int32_t nb0 = ca + GB, nb1 = g0 + bw;
            if (nb0 < nb1) {
                Digits d1 = dg;
                d1.k0 = a;
                launch_update_i8(dg.nd, 1, s, M, ldm, nrl, nb0, nb1, g0, pan, act, md, d1);
            }
        }
        // ... and the columns right of the panel all four together (K = 256;
        // the last panel may be partial: its right-hand columns are updated too)
        const int32_t cb = g0 + bw, ce = nc;
        launch_update_i8(dg.nd, KCH, s, M, ldm, nrl, cb, ce, g0, pan, active, md, dg);
        GBGPU_HIP(ctx, hipGetLastError());
    }
    k_ds_extract<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, ncl, nc, M, ldm, m->dprime, m->dnz);
    GBGPU_HIP(ctx, hipGetLastError());
    cleanup();
    return GBGPU_OK;
}

}  // namespace gbgpu
