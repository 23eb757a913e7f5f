// dense.hip -- exact reduced row echelon form of a dense block mod p.
//
// Replaces exact_dense_linear_algebra_ff_{8,16,32} (la_ff_32.c:1217-1333),
// interreduce_dense_matrix_ff_{8,16,32} (la_ff_32.c:1181-1215) and their
// row kernels reduce_dense_row_by_dense_new_pivots_{17,31}_bit
// (la_ff_32.c:632-753).  The reference reduces rows one at a time against
// published pivots and back-substitutes at the end; the result is the
// unique reduced echelon form of the row space, so any elimination order
// gives the same bits.  This file uses blocked Gauss-Jordan elimination,
// which turns almost all of the work into one product per 64-column panel:
//
//   for each panel of 64 columns [c0, c0 + 64):
//     select  (1 workgroup): scan the non-pivot rows in order, keep the
//             panel parts of the rows that add rank, in reduced form E, and
//             the transform T with E = T * A[sel, panel];  k <= 64 pivots
//     rp      Rp = T * A[sel, c0:n]   (rows normalised, identity on the
//             panel's pivot columns); written back into the selected rows
//     gather  L[r] = A[r, pivot columns] for every other row
//     update  A[r, c0:n] -= L[r] * Rp  for every other row with L[r] != 0
//             -- previous pivot rows too, so the result is fully reduced
//
// Pivot rows stay where they are; rowpiv[r] / colpiv[c] record the pairing.
// The update is the hot kernel: plain form on the integer ALU (64-bit
// multiply-add), MFMA form on the FP64 matrix cores with the exactness
// bound k (p-1)^2 < 2^53 (p < 2^17) or, for larger p, 16-bit operand
// splitting (four exact products recombined in integers).
#include <climits>

#include "internal.h"

namespace gbgpu {

constexpr int PB = 64;  // panel width

struct Panel {
    int32_t k;           // pivots found in this panel
    int32_t done;        // set once no non-pivot row is left
    int32_t erow[PB];    // selected rows
    int32_t ecol[PB];    // their pivot columns (absolute)
    uint32_t T[PB * PB]; // Rp[t] = sum_s T[t][s] * A[erow[s]]
};

__device__ uint32_t inv_mod(uint32_t a, uint32_t p) {
    // extended Euclid (tools.c:83-110 computes the same inverse)
    int64_t r0 = p, r1 = a % p, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const int64_t q = r0 / r1;
        int64_t t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    if (s0 < 0) s0 += p;
    return (uint32_t)s0;
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    return a >= b ? a - b : a + p - b;
}

// ------------------------------------------------------------ select

__global__ __launch_bounds__(256) void k_panel_select(const uint32_t *a, int64_t ld, int32_t m,
                                                      int32_t n, int32_t c0, int32_t *rowpiv,
                                                      int32_t *colpiv, Panel *pan, Mod md) {
    __shared__ uint32_t E[PB][PB];
    __shared__ uint32_t T[PB][PB];
    __shared__ int32_t ecol[PB], erow[PB];
    __shared__ uint32_t F[PB];
    __shared__ int32_t sh_best, sh_any, sh_k, sh_cstar;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t p = md.p;
    if (pan->done) {
        if (tid == 0) pan->k = 0;
        return;
    }
    const int bw = min(PB, n - c0);
    for (int i = tid; i < PB * PB; i += 256) { (&E[0][0])[i] = 0; (&T[0][0])[i] = 0; }
    // shared flags are written by every thread with the same value: no
    // thread-specific branches inside the loops below (see sparse.hip)
    sh_k = 0;
    sh_any = 0;
    __syncthreads();
    int32_t cursor = 0;
    for (;;) {
        sh_best = INT_MAX;
        __syncthreads();
        // values every lane shares are made scalar: loops that contain
        // cross-lane operations must have wave-uniform trip counts and exits
        const int kk = __builtin_amdgcn_readfirstlane(sh_k);
        // parallel scan: first non-pivot row >= cursor not in span(E)
        for (int32_t r = cursor + wv; r < m; r += 4) {
            if (r > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)&sh_best)) break;
            if (__builtin_amdgcn_readfirstlane(rowpiv[r]) >= 0) continue;
            sh_any = 1;
            const uint32_t v = lane < bw ? a[(int64_t)r * ld + c0 + lane] : 0u;
            uint64_t s = 0;
            for (int t = 0; t < kk; ++t) {
                const uint32_t ct = __shfl(v, ecol[t] - c0, 64);
                s += (uint64_t)md.mul(ct, E[t][lane]);
            }
            const uint32_t vr = sub_mod(v, md.reduce(s), p);
            if (__any(vr != 0)) {
                atomicMin(&sh_best, r);
                break;
            }
        }
        __syncthreads();
        const int32_t rs = __builtin_amdgcn_readfirstlane(sh_best);
        if (rs == INT_MAX) break;
        if (wv == 0) {
            const uint32_t v = lane < bw ? a[(int64_t)rs * ld + c0 + lane] : 0u;
            uint64_t s = 0, ts = 0;
            for (int t = 0; t < kk; ++t) {
                const uint32_t ct = __shfl(v, ecol[t] - c0, 64);
                s += (uint64_t)md.mul(ct, E[t][lane]);
                ts += (uint64_t)md.mul(ct, T[t][lane]);
            }
            const uint32_t vr = sub_mod(v, md.reduce(s), p);
            const uint64_t nzm = __ballot(vr != 0);
            const int cs = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)nzm) - 1);
            const uint32_t inv = inv_mod(__shfl(vr, cs, 64), p);
            E[kk][lane] = md.mul(vr, inv);
            // T row of the new pivot: (e_new - sum_t ct T[t]) * inv
            uint32_t tn = sub_mod(0, md.reduce(ts), p);
            if (lane == kk) tn = 1;
            T[kk][lane] = md.mul(tn, inv);
            ecol[kk] = c0 + cs;
            erow[kk] = rs;
            sh_cstar = cs;
        }
        __syncthreads();
        // keep E reduced: eliminate the new pivot column from older rows
        // (factors read into F first: the j == cs thread overwrites E[t][cs])
        const int cstar = __builtin_amdgcn_readfirstlane(sh_cstar);
        if (tid < kk) F[tid] = E[tid][cstar];
        __syncthreads();
        for (int idx = tid; idx < kk * PB; idx += 256) {
            const int t = idx >> 6, j = idx & 63;
            const uint32_t f = F[t];
            if (f == 0) continue;
            E[t][j] = sub_mod(E[t][j], md.mul(f, E[kk][j]), p);
            T[t][j] = sub_mod(T[t][j], md.mul(f, T[kk][j]), p);
        }
        __syncthreads();
        sh_k = kk + 1;
        cursor = rs + 1;
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(sh_k) == bw) break;
    }
    const int kk = __builtin_amdgcn_readfirstlane(sh_k);
    if (tid < kk) {
        pan->erow[tid] = erow[tid];
        pan->ecol[tid] = ecol[tid];
        rowpiv[erow[tid]] = ecol[tid];
        colpiv[ecol[tid]] = erow[tid];
    }
    for (int i = tid; i < PB * PB; i += 256) pan->T[i] = (&T[0][0])[i];
    if (tid == 0) {
        pan->k = kk;
        if (kk == 0 && !sh_any) pan->done = 1;
    }
}

// ---------------------------------------------------------------- rp

// one thread per column j >= c0: Rp[t][j] = sum_s T[t][s] A[erow[s]][j]
__global__ __launch_bounds__(256) void k_panel_rp(uint32_t *a, int64_t ld, int32_t n, int32_t c0,
                                                  const Panel *pan, uint32_t *rp, int64_t ldr,
                                                  Mod md) {
    const int kk = pan->k;
    if (kk == 0) return;
    __shared__ uint32_t T[PB][PB];
    __shared__ int32_t erow[PB];
    for (int i = threadIdx.x; i < PB * PB; i += blockDim.x) (&T[0][0])[i] = pan->T[i];
    if (threadIdx.x < PB) erow[threadIdx.x] = pan->erow[threadIdx.x];
    __syncthreads();
    const int32_t j = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t v[PB];
#pragma unroll
    for (int s = 0; s < PB; ++s) v[s] = s < kk ? a[(int64_t)erow[s] * ld + j] : 0u;
    for (int t = 0; t < kk; ++t) {
        uint64_t acc = 0;
#pragma unroll
        for (int s = 0; s < PB; ++s) acc += (uint64_t)md.mul(T[t][s], v[s]);
        const uint32_t r = md.reduce(acc);
        rp[(int64_t)t * ldr + (j - c0)] = r;
        a[(int64_t)erow[t] * ld + j] = r;
    }
}

// ------------------------------------------------------------ gather

// L[r][t] = A[r][ecol[t]] for rows that are not this panel's pivots;
// active[r] = 1 if any of them is nonzero
__global__ void k_panel_gather(const uint32_t *a, int64_t ld, int32_t m, int32_t c0, int32_t bw,
                               const int32_t *rowpiv, const Panel *pan, uint32_t *L,
                               int32_t *active) {
    const int kk = pan->k;
    if (kk == 0) return;
    const int lane = threadIdx.x & 63;
    const int32_t r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= m) return;
    const int32_t rp = rowpiv[r];
    const bool cur = rp >= c0 && rp < c0 + bw;
    uint32_t v = 0;
    if (!cur && lane < kk) v = a[(int64_t)r * ld + pan->ecol[lane]];
    L[(int64_t)r * PB + lane] = v;
    const bool any = __any(v != 0);
    if (lane == 0) active[r] = any ? 1 : 0;
}

// ------------------------------------------------- update (integer ALU)

// block: 256 columns x RT rows; A[r][j] -= sum_t L[r][t] Rp[t][j]
constexpr int RT = 16;

template <bool BIG>
__global__ __launch_bounds__(256) void k_update_plain(uint32_t *a, int64_t ld, int32_t m,
                                                      int32_t n, int32_t c0, const Panel *pan,
                                                      const uint32_t *rp, int64_t ldr,
                                                      const uint32_t *L, const int32_t *active,
                                                      Mod md, uint32_t r64) {
    const int kk = pan->k;
    if (kk == 0) return;
    __shared__ uint32_t Ls[RT][PB];
    __shared__ int32_t act[RT];
    const int32_t r0 = blockIdx.y * RT;
    for (int i = threadIdx.x; i < RT * PB; i += 256) {
        const int rr = i / PB, t = i % PB;
        Ls[rr][t] = (r0 + rr < m) ? L[(int64_t)(r0 + rr) * PB + t] : 0u;
    }
    if (threadIdx.x < RT) act[threadIdx.x] = (r0 + threadIdx.x < m) ? active[r0 + threadIdx.x] : 0;
    __syncthreads();
    int anyact = 0;
#pragma unroll
    for (int rr = 0; rr < RT; ++rr) anyact |= act[rr];
    if (!anyact) return;
    const int32_t j = c0 + blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    uint32_t rv[PB];
#pragma unroll
    for (int t = 0; t < PB; ++t) rv[t] = t < kk ? rp[(int64_t)t * ldr + (j - c0)] : 0u;
    for (int rr = 0; rr < RT; ++rr) {
        if (!act[rr]) continue;
        uint64_t lo = 0;
        uint32_t hi = 0;
#pragma unroll
        for (int t = 0; t < PB; ++t) {
            const uint64_t pr = (uint64_t)Ls[rr][t] * (uint64_t)rv[t];
            if (BIG) {
                const uint64_t o = lo;
                lo += pr;
                hi += lo < o;
            } else {
                lo += pr;
            }
        }
        uint32_t s = md.reduce(lo);
        if (BIG && hi) s = (uint32_t)((s + (uint64_t)md.mul(hi, r64)) % md.p);
        uint32_t *x = a + (int64_t)(r0 + rr) * ld + j;
        *x = sub_mod(*x, s, md.p);
    }
}

// --------------------------------------------------- update (MFMA f64)

// Block: 4 waves, 128 rows x 64 columns; each wave a 32 x 64 tile held as
// 2 x 4 accumulators of v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, C 16x16).
// Operand layout per lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result
// register e of lane l is C[4 e + l / 16][l % 16] (measured on gfx950 with
// an in-bounds probe of the builtin).
// Exactness: operands are integers below 2^17 (p < 2^17) or 16-bit halves
// of residues (p >= 2^17), so every 64-term dot product is an integer below
// 2^40 (2^36 for halves) and FP64 represents it and every partial sum
// exactly.
typedef double dx4 __attribute__((ext_vector_type(4)));

constexpr int MR = 128;  // rows per block

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// part: 0 = whole value, 1 = high 16 bits, 2 = low 16 bits
__device__ __forceinline__ double piece(uint32_t v, int part) {
    return part == 0 ? (double)v : part == 1 ? (double)(v >> 16) : (double)(v & 0xFFFFu);
}

__device__ __forceinline__ void tile_product(const uint32_t (*Ls)[PB + 1], const uint32_t (*Rs)[PB + 1],
                                             int kp, int lpart, int rpart, dx4 acc[2][4]) {
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = dx4{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < kp; t0 += 4) {
        double av[2], bv[4];
#pragma unroll
        for (int i = 0; i < 2; ++i) av[i] = piece(Ls[16 * i + lr][t0 + lk], lpart);
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = piece(Rs[t0 + lk][16 * j + lr], rpart);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(av[i], bv[j], acc[i][j]);
    }
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_update_mfma(uint32_t *a, int64_t ld, int32_t m,
                                                     int32_t n, int32_t c0, const Panel *pan,
                                                     const uint32_t *rp, int64_t ldr,
                                                     const uint32_t *L, const int32_t *active,
                                                     Mod md) {
    const int kk = pan->k;
    if (kk == 0) return;
    __shared__ uint32_t Rs[PB][PB + 1];
    __shared__ uint32_t Ls[MR][PB + 1];
    __shared__ int32_t act[MR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t jc = c0 + blockIdx.x * 64;
    const int32_t rb = blockIdx.y * MR;
    if (tid < MR) act[tid] = (rb + tid < m) ? active[rb + tid] : 0;
    __syncthreads();
    int anyact = 0;
    for (int i = 0; i < MR; ++i) anyact |= act[i];
    if (!anyact) return;  // uniform over the block
    const int kp = (kk + 3) & ~3;
    for (int i = tid; i < PB * 64; i += 256) {
        const int t = i >> 6, c = i & 63;
        Rs[t][c] = (t < kk && jc + c < n) ? rp[(int64_t)t * ldr + (jc + c - c0)] : 0u;
    }
    for (int i = tid; i < MR * PB; i += 256) {
        const int rr = i >> 6, t = i & 63;
        Ls[rr][t] = (t < kk && act[rr]) ? L[(int64_t)(rb + rr) * PB + t] : 0u;
    }
    __syncthreads();
    const uint32_t (*Lw)[PB + 1] = Ls + 32 * wv;
    uint32_t res[2][4][4];
    dx4 acc[2][4];
    if (!SPLIT) {
        tile_product(Lw, Rs, kp, 0, 0, acc);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) res[i][j][e] = md.reduce((uint64_t)acc[i][j][e]);
    } else {
        // L R = (Lh Rh) 2^32 + (Lh Rl + Ll Rh) 2^16 + Ll Rl
        const uint32_t s16 = (uint32_t)((1ull << 16) % md.p);
        const uint32_t s32 = (uint32_t)((1ull << 32) % md.p);
        const int parts[4][2] = {{1, 1}, {1, 2}, {2, 1}, {2, 2}};
        const uint32_t scale[4] = {s32, s16, s16, 1u};
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) res[i][j][e] = 0;
        for (int q = 0; q < 4; ++q) {
            tile_product(Lw, Rs, kp, parts[q][0], parts[q][1], acc);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const uint32_t v = md.mul(md.reduce((uint64_t)acc[i][j][e]), scale[q]);
                        const uint32_t s = res[i][j][e] + v;
                        res[i][j][e] = s >= md.p ? s - md.p : s;
                    }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int rloc = 32 * wv + 16 * i + 4 * e + (lane >> 4);
            const int32_t r = rb + rloc;
            if (r >= m || !act[rloc]) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t c = jc + 16 * j + (lane & 15);
                if (c >= n) continue;
                uint32_t *x = a + (int64_t)r * ld + c;
                *x = sub_mod(*x, res[i][j][e], md.p);
            }
        }
}

// ------------------------------------------------------------ output

__global__ void k_out_flags(int32_t n, const int32_t *colpiv, int32_t *flag) {
    for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        flag[c] = colpiv[n - 1 - c] >= 0;  // decreasing lead order
}

__global__ void k_out_list(int32_t n, const int32_t *colpiv, const int32_t *flag,
                           const int32_t *rank, int32_t *orow, int32_t *olead) {
    for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        if (flag[c]) {
            orow[rank[c]] = colpiv[n - 1 - c];
            olead[rank[c]] = n - 1 - c;
        }
}

__global__ void k_fill_dense(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int grid_of(int64_t n, int bs) {
    int64_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > 65535 * 4) g = 65535 * 4;
    return (int)g;
}

int dense_rref(gbgpu_dense *d, int form) {
    gbgpu_ctx *ctx = d->ctx;
    hipStream_t s = ctx->stream;
    const int32_t m = d->m, n = d->n;
    const Mod md = Mod::make(d->p);
    const bool big = d->p >= (1u << 17);
    int rc;
    dfree(d->orow);
    dfree(d->olead);
    d->rank = 0;
    int32_t *rowpiv = nullptr, *colpiv = nullptr, *active = nullptr;
    uint32_t *L = nullptr, *rp = nullptr;
    Panel *pan = nullptr;
    const int64_t ldr = ((int64_t)n + 63) & ~63ll;
    auto cleanup = [&]() {
        dfree(rowpiv); dfree(colpiv); dfree(active); dfree(L); dfree(rp); dfree(pan);
    };
    if ((rc = dalloc(ctx, &rowpiv, (size_t)m)) || (rc = dalloc(ctx, &colpiv, (size_t)n)) ||
        (rc = dalloc(ctx, &active, (size_t)m)) || (rc = dalloc(ctx, &L, (size_t)m * PB)) ||
        (rc = dalloc(ctx, &rp, (size_t)PB * ldr)) || (rc = dalloc(ctx, &pan, 1))) {
        cleanup();
        return rc;
    }
    if (m > 0) k_fill_dense<<<grid_of(m, 256), 256, 0, s>>>(rowpiv, m, -1);
    if (n > 0) k_fill_dense<<<grid_of(n, 256), 256, 0, s>>>(colpiv, n, -1);
    if (hipMemsetAsync(pan, 0, sizeof(Panel), s) != hipSuccess) {
        cleanup();
        return check_hip(ctx, hipGetLastError(), "hipMemsetAsync");
    }
    const uint32_t r64 = (uint32_t)((((unsigned __int128)1) << 64) % d->p);
    if (hipEventRecord(ctx->ev[2], s) != hipSuccess) { cleanup(); return check_hip(ctx, hipGetLastError(), "event"); }
    int32_t pan_done = 0;
    int npan = 0;
    for (int32_t c0 = 0; c0 < n && m > 0; c0 += PB) {
        const int bw = (n - c0) < PB ? (n - c0) : PB;
        k_panel_select<<<1, 256, 0, s>>>(d->a, d->ld, m, n, c0, rowpiv, colpiv, pan, md);
        k_panel_rp<<<grid_of(n - c0, 256), 256, 0, s>>>(d->a, d->ld, n, c0, pan, rp, ldr, md);
        k_panel_gather<<<grid_of(m, 4), 256, 0, s>>>(d->a, d->ld, m, c0, bw, rowpiv, pan, L, active);
        if (form == GBGPU_FORM_MFMA) {
            dim3 g((unsigned)((n - c0 + 63) / 64), (unsigned)((m + MR - 1) / MR));
            if (big) k_update_mfma<true><<<g, 256, 0, s>>>(d->a, d->ld, m, n, c0, pan, rp, ldr, L, active, md);
            else k_update_mfma<false><<<g, 256, 0, s>>>(d->a, d->ld, m, n, c0, pan, rp, ldr, L, active, md);
        } else {
            dim3 g((unsigned)((n - c0 + 255) / 256), (unsigned)((m + RT - 1) / RT));
            if (big) k_update_plain<true><<<g, 256, 0, s>>>(d->a, d->ld, m, n, c0, pan, rp, ldr, L, active, md, r64);
            else k_update_plain<false><<<g, 256, 0, s>>>(d->a, d->ld, m, n, c0, pan, rp, ldr, L, active, md, r64);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { cleanup(); return check_hip(ctx, e, "dense panel launch"); }
        if (++npan % 8 == 0) {
            if (hipMemcpyAsync(&pan_done, &pan->done, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                cleanup();
                return check_hip(ctx, hipGetLastError(), "dense panel sync");
            }
            if (pan_done) break;
        }
    }
    // output list: pivot rows by decreasing lead column
    int32_t *flag = nullptr, *rank = nullptr;
    if ((rc = dalloc(ctx, &flag, (size_t)n)) || (rc = dalloc(ctx, &rank, (size_t)n))) {
        dfree(flag); cleanup(); return rc;
    }
    if (n > 0) k_out_flags<<<grid_of(n, 256), 256, 0, s>>>(n, colpiv, flag);
    int64_t total = 0;
    if ((rc = exclusive_scan_i32(ctx, flag, rank, n, &total))) { dfree(flag); dfree(rank); cleanup(); return rc; }
    d->rank = (int32_t)total;
    if ((rc = dalloc(ctx, &d->orow, (size_t)total)) || (rc = dalloc(ctx, &d->olead, (size_t)total))) {
        dfree(flag); dfree(rank); cleanup(); return rc;
    }
    if (n > 0) k_out_list<<<grid_of(n, 256), 256, 0, s>>>(n, colpiv, flag, rank, d->orow, d->olead);
    (void)hipEventRecord(ctx->ev[3], s);
    hipError_t e = hipStreamSynchronize(s);
    dfree(flag); dfree(rank);
    cleanup();
    if (e != hipSuccess) return check_hip(ctx, e, "dense rref");
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) ctx->phase_ms[1] = ms;
    return GBGPU_OK;
}

}  // namespace gbgpu
