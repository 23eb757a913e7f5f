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
//   for a = 0..3:  x_a = M[:, group a] T_a^-1   (T_a the group's unit
//                  upper-triangular block, inverted for the blocked pass):
//                  the multipliers that clear the group's columns
//                  M[:, later groups of the panel] -= x_a P_a   (K = 64)
//   M[:, right of the panel] -= [x_0 .. x_3] [P_0; ..; P_3]   (K = 256)
//
// All products run on the int8 matrix cores in signed base-256 digits: the
// multipliers (nrl x 64 x 64 per group), and the updates with the dense
// echelon form's kernels (the wide one covers four groups per pass over M).
// Columns without a pivot get identity rows in T_g and zero rows in P_g:
// their multiplier meets a zero row, so a row's entry there is left alone,
// as the reference skips them (la_ff_32.c:517-519); only the right-hand part
// is kept.
#include "internal.h"

namespace gbgpu {

constexpr int GB = 64;  // pivot columns per group

// M[q][c] = entry of to-reduce row upiv[q] at column c (M zeroed before)
template <typename T>
__global__ void k_ds_scatter(int32_t nrl, const int32_t *upiv, const int64_t *rp,
                             const int32_t *ci, const int32_t *rcf, const int64_t *cp,
                             const uint32_t *cf, T *M, int64_t ldm) {
    const int lane = threadIdx.x & 63;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nrl; q += step) {
        const int32_t row = upiv[q];
        const int64_t b = rp[row], e = rp[row + 1];
        const uint32_t *c = cf + cp[rcf[row]] - b;
        T *out = M + (int64_t)q * ldm;
        for (int64_t k = b + lane; k < e; k += 64) out[ci[k]] = (T)c[k];
    }
}

// P[s][c - g0] = entry of the pivot of column g0 + s at column c >= g0, for
// s < rows (zero rows for columns without a pivot); P zeroed before.  Row s
// is shared by the gridDim.y waves with that blockIdx.x lane group, each
// taking every gridDim.y-th run of 64 entries.
template <typename T>
__global__ void k_ds_group_rows(int32_t g0, int32_t ncl, int rows, const int32_t *first,
                                const int64_t *rp, const int32_t *ci, const int32_t *rcf,
                                const int64_t *cp, const uint32_t *cf, T *P, int64_t ldp) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (s >= rows) return;
    const int32_t col = g0 + s;
    if (col >= ncl) return;
    const int32_t row = first[col];
    if (row == INT32_MAX) return;
    const int64_t b = rp[row], e = rp[row + 1];
    const uint32_t *c = cf + cp[rcf[row]] - b;
    T *out = P + (int64_t)s * ldp - g0;
    const int64_t step = 64 * (int64_t)gridDim.y;
    for (int64_t k = b + 64 * (int64_t)blockIdx.y + lane; k < e; k += step) out[ci[k]] = (T)c[k];
}

// the nd signed digits of x's representative (x or x - p), packed four to a
// word: byte u of out[i] word w = digit i of value 4 w + u
__device__ __forceinline__ void digits4(const uint32_t (&x)[4], uint32_t p, uint32_t hi, int nd,
                                        uint32_t *out) {
    for (int i = 0; i < nd; ++i) out[i] = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int64_t v = x[u] <= hi ? (int64_t)x[u] : (int64_t)x[u] - (int64_t)p;
        for (int i = 0; i < nd; ++i) {
            const int8_t d = (int8_t)(uint8_t)(v & 0xFF);
            v = (v - d) >> 8;
            out[i] |= (uint32_t)(uint8_t)d << (8 * u);
        }
    }
}

// digit planes of the panel's pivot rows P (rows >= `rows` as zero): slot t
// of column c at byte t % 64 of column c in chunk t / 64.  A block turns 64
// columns x 256 slots: each thread reads four slots of one column
// (coalesced across the wave), and the digits go through LDS (17 words per
// column and chunk, so neither the writes nor the reads collide) to leave as
// 4 KiB of consecutive bytes per digit and chunk.
template <typename T, int NDL>
__global__ __launch_bounds__(256) void k_ds_pdigits(const T *P, int64_t ldp, int rows,
                                                    int32_t w, uint32_t p, Digits dg) {
    __shared__ uint32_t t4[NDL * 4 * 64 * 17];  // [digit][chunk][column][17 words], nd <= NDL
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t cb = (int32_t)blockIdx.x * 64;
    const int32_t c = cb + lane;
    for (int q = wv; q < 64; q += 4) {  // slots 4 q .. 4 q + 3
        uint32_t x[4], d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = 4 * q + u;
            x[u] = (t < rows && c < w) ? (uint32_t)P[(int64_t)t * ldp + c] : 0u;
        }
        digits4(x, p, dg.hi, dg.nd, d);
        const int kc = q >> 4, wd = q & 15;
        for (int i = 0; i < dg.nd; ++i) t4[((i * 4 + kc) * 64 + lane) * 17 + wd] = d[i];
    }
    __syncthreads();
    const int col = tid >> 2, j = tid & 3;
    for (int i = 0; i < dg.nd; ++i)
        for (int kc = 0; kc < 4; ++kc) {
            const uint32_t *src = t4 + ((i * 4 + kc) * 64 + col) * 17 + 4 * j;
            const v4i v = v4i{(int)src[0], (int)src[1], (int)src[2], (int)src[3]};
            *(v4i *)(dg.R + i * dg.rplane + kc * dg.rchunk + (int64_t)(cb + col) * 64 + 16 * j) = v;
        }
}

// T_g^-1 as B operands of x = v T_g^-1: digit i of T_g^-1[k][s] at byte k of
// column s; group g's nd planes of 64 x 64 bytes at td + g nd 4096
__global__ void k_ds_tdigits(int32_t ngroups, const uint32_t *tinv, int8_t *td, uint32_t p,
                             Digits dg) {
    const int64_t n = (int64_t)ngroups * GB * GB;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = q >> 12;
        const int s = (int)((q >> 6) & 63), k = (int)(q & 63);
        const uint32_t v = tinv[(g << 12) + k * GB + s];
        for (int i = 0; i < dg.nd; ++i)
            td[(g * dg.nd + i) * 4096 + s * 64 + k] = (int8_t)digit_word(v, p, dg.hi, i);
    }
}

// Multipliers of group a: x = v T^-1 with v = M[q][ca .. ca + wa) (zero from
// wa on), on the matrix cores (a wave: 16 rows x 64 columns, K = 64), written
// as the digits of chunk a of L; act[q] = row has a nonzero multiplier,
// any[q] = OR over the panel's groups.  The digits of v are formed in
// registers from M.
template <int ND, typename T>
__global__ __launch_bounds__(64) void k_ds_mult(int32_t nrl, int32_t ca, int32_t wa, int a,
                                                 const T *M, int64_t ldm, const int8_t *td,
                                                 int32_t *act, int32_t *any, Mod md, Digits dg) {
    const uint32_t p = md.p;
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int32_t r0 = (int32_t)blockIdx.x * 16;  // one wave per block, 16 rows
    if (r0 >= nrl) return;  // wave-uniform
    // A operand: row r0 + lr, slots 16 lk .. 16 lk + 15 (columns past ncl,
    // from wa on, count as zero; they lie inside the row: ca + 64 <= ldm)
    const int32_t row = r0 + lr;
    const T *src = M + (int64_t)(row < nrl ? row : nrl - 1) * ldm + ca + 16 * lk;
    v4i A[ND];
    bool nz = false;
#pragma unroll
    for (int w4 = 0; w4 < 4; ++w4) {
        uint32_t x[4], d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = 16 * lk + 4 * w4 + u;
            x[u] = (row < nrl && k < wa) ? (uint32_t)src[4 * w4 + u] : 0u;
            nz = nz || x[u] != 0u;
        }
        digits4(x, p, dg.hi, ND, d);
#pragma unroll
        for (int i = 0; i < ND; ++i) A[i][w4] = (int)d[i];
    }
    const uint64_t ball = __ballot(nz);
    v4i acc[4][2 * ND - 1];
#pragma unroll
    for (int cs = 0; cs < 4; ++cs)
#pragma unroll
        for (int q = 0; q < 2 * ND - 1; ++q) acc[cs][q] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int cs = 0; cs < 4; ++cs) {
        v4i B[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) B[j] = *(const v4i *)(td + j * 4096 + (16 * cs + lr) * 64 + 16 * lk);
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j)
                acc[cs][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i], B[j], acc[cs][i + j], 0, 0, 0);
    }
    // result element e of lane l: row r0 + 4 (l / 16) + e, column 16 cs + l % 16
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int rl = 4 * lk + e;
        const int32_t ro = r0 + rl;
        if (ro >= nrl) continue;
        const int f = ((ball >> rl) & 0x0001000100010001ull) != 0 ? 1 : 0;
#pragma unroll
        for (int cs = 0; cs < 4; ++cs) {
            const uint32_t r = digits_sub<ND>(0u, acc[cs], e, md, dg);  // -x mod p
            const uint32_t x = r == 0u ? 0u : p - r;
            int64_t v = x <= dg.hi ? (int64_t)x : (int64_t)x - (int64_t)p;
#pragma unroll
            for (int i = 0; i < ND; ++i) {
                const int8_t d = (int8_t)(uint8_t)(v & 0xFF);
                v = (v - d) >> 8;
                dg.L[i * dg.lplane + a * dg.lchunk + (int64_t)ro * 64 + 16 * cs + lr] = d;
            }
        }
        // the sixteen lanes of a row store the same flags
        act[ro] = f;
        any[ro] = a == 0 ? f : (any[ro] | f);
    }
}

// D'[q] = M[q][ncl .. nc), dnz[q] = row nonzero
template <typename T>
__global__ void k_ds_extract(int32_t nrl, int32_t ncl, int32_t nc, const T *M,
                             int64_t ldm, uint32_t *dprime, int32_t *dnz) {
    const int lane = threadIdx.x & 63;
    const int32_t ncr = nc - ncl;
    const int32_t step = gridDim.x * (blockDim.x >> 6);
    for (int32_t q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nrl; q += step) {
        const T *src = M + (int64_t)q * ldm + ncl;
        uint32_t *dst = dprime + (int64_t)q * ncr;
        bool nz = false;
        for (int32_t c = lane; c < ncr; c += 64) {
            const uint32_t v = src[c];  // widened from T
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
    if (m->ncl < GB || m->nrl < 1) return false;
    const double mean_len = (double)m->nbucket / (double)m->ncl;
    return mean_len * 40.0 > (double)m->nc;
}

// M and the panel rows P hold residues in T: uint16_t for p <= 65521 (one or
// two digits), which halves the bytes every update moves, else uint32_t
template <typename T>
static int dense_route(gbgpu_mat *m, const uint32_t *tinv) {
    gbgpu_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    constexpr int KCH = 4;               // groups per panel
    constexpr int PW = GB * KCH;         // pivot columns per panel
    const int32_t nrl = m->nrl, ncl = m->ncl, nc = m->nc;
    const int32_t ngroups = (ncl + GB - 1) / GB;
    const int64_t ldm = ((int64_t)nc + 63) & ~63ll;
    const int64_t ldp = ldm, ldr = ldm;
    const Mod md = Mod::make(m->p);
    Digits dg;
    digits_setup(dg, m->p, digits_for(m->p));
    const int nd = dg.nd;
    T *M = nullptr, *P = nullptr;
    int8_t *dl = nullptr, *dr = nullptr, *td = nullptr;
    int32_t *active = nullptr, *act = nullptr, *pk = nullptr;  // pk: a panel record whose k is 64
    int rc;
    auto cleanup = [&]() {
        dfree(M); dfree(P); dfree(dl); dfree(dr); dfree(td); dfree(active); dfree(act); dfree(pk);
    };
    if ((rc = dalloc(ctx, &M, (size_t)nrl * ldm)) || (rc = dalloc(ctx, &P, (size_t)PW * ldp)) ||
        (rc = dalloc(ctx, &act, (size_t)nrl)) || (rc = dalloc(ctx, &dl, (size_t)nd * nrl * PW)) ||
        (rc = dalloc(ctx, &dr, (size_t)nd * ldr * PW)) ||
        (rc = dalloc(ctx, &td, (size_t)nd * ngroups * GB * GB)) ||
        (rc = dalloc(ctx, &active, (size_t)nrl)) ||
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
    GBGPU_HIP(ctx, hipMemsetAsync(M, 0, (size_t)nrl * ldm * sizeof(T), s));
    k_ds_scatter<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, m->upiv, m->rp, m->ci, m->rcf, m->cp, m->cf,
                                                    M, ldm);
    k_ds_tdigits<<<grid_rows((int64_t)ngroups * GB * GB, 256), 256, 0, s>>>(ngroups, tinv, td, m->p, dg);
    const Panel *pan = (const Panel *)pk;
    const int mult_blocks = (nrl + 15) / 16;
    for (int32_t g0 = 0; g0 < ncl; g0 += PW) {
        const int32_t bw = (ncl - g0) < PW ? (ncl - g0) : PW;  // pivot columns of the panel
        const int na = (bw + GB - 1) / GB;                     // its groups
        const int32_t w = nc - g0;                             // columns from the panel on
        GBGPU_HIP(ctx, hipMemsetAsync(P, 0, (size_t)PW * ldp * sizeof(T), s));
        k_ds_group_rows<<<dim3(PW / 4, 8), 256, 0, s>>>(g0, ncl, PW, m->first, m->rp, m->ci, m->rcf, m->cp,
                                               m->cf, P, ldp);
        if (nd <= 2) k_ds_pdigits<T, 2><<<(w + 63) / 64, 256, 0, s>>>(P, ldp, na * GB, w, m->p, dg);
        else k_ds_pdigits<T, 4><<<(w + 63) / 64, 256, 0, s>>>(P, ldp, na * GB, w, m->p, dg);
        // the panel's groups in order: group a's multipliers clear its
        // columns; its rows update the panel's later columns at once
        // (K = 64) ...
        for (int a = 0; a < na; ++a) {
            const int32_t ca = g0 + a * GB;
            const int32_t wa = (bw - a * GB) < GB ? (bw - a * GB) : GB;
            const int8_t *tg = td + (int64_t)(g0 / GB + a) * nd * GB * GB;
            if (nd == 1) k_ds_mult<1, T><<<mult_blocks, 64, 0, s>>>(nrl, ca, wa, a, M, ldm, tg, act, active, md, dg);
            else if (nd == 2 || sizeof(T) == 2) k_ds_mult<2, T><<<mult_blocks, 64, 0, s>>>(nrl, ca, wa, a, M, ldm, tg, act, active, md, dg);
            else if constexpr (sizeof(T) == 4) {
                if (nd == 3) k_ds_mult<3, T><<<mult_blocks, 64, 0, s>>>(nrl, ca, wa, a, M, ldm, tg, act, active, md, dg);
                else k_ds_mult<4, T><<<mult_blocks, 64, 0, s>>>(nrl, ca, wa, a, M, ldm, tg, act, active, md, dg);
            }
            const int32_t nb0 = ca + GB, nb1 = g0 + bw;
            if (nb0 < nb1) {
                Digits d1 = dg;
                d1.k0 = a;
                launch_update_i8(nd, 1, s, M, ldm, nrl, nb0, nb1, g0, pan, act, md, d1);
            }
        }
        // ... and the columns right of the panel all four together (K = 256;
        // the digits of a partial panel's missing groups meet zero rows)
        const int32_t cb = g0 + bw, ce = nc;
        launch_update_i8(nd, KCH, s, M, ldm, nrl, cb, ce, g0, pan, active, md, dg);
        GBGPU_HIP(ctx, hipGetLastError());
    }
    k_ds_extract<<<grid_rows(nrl, 4), 256, 0, s>>>(nrl, ncl, nc, M, ldm, m->dprime, m->dnz);
    GBGPU_HIP(ctx, hipGetLastError());
    cleanup();
    return GBGPU_OK;
}

int run_dense_route(gbgpu_mat *m, const uint32_t *tinv) {
    // one or two digits: p <= 65521, residues fit 16 bits
    return digits_for(m->p) <= 2 ? dense_route<uint16_t>(m, tinv) : dense_route<uint32_t>(m, tinv);
}

}  // namespace gbgpu
