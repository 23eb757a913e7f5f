// sparse.hip -- the plain sparse pass of F4 linear algebra on gfx950.
//
// Replaces sparse_AB_CD_linear_algebra_ff_{8,16,32} (la_ff_32.c:1081-1179)
// with its row kernel reduce_dense_row_by_old_pivots_{17,31}_bit
// (la_ff_32.c:496-630, la_ff_16.c:289-351).
//
// Layout in HBM:
//   * the uploaded CSR (rp, ci, rcf, cp, cf) is read once per row;
//   * the known pivot rows are re-laid out once per matrix as (column,
//     coefficient) pairs -- 8 bytes, one coalesced load -- bucketed by column
//     window: seg[w * ncl + j] .. seg[(w+1) * ncl + j] are pivot j's entries
//     in window w (window-major so a wave loads 64 consecutive pivots'
//     offsets with one coalesced load);
//   * every to-reduce row q writes its dense remainder D'[q] (ncr uint32)
//     and a nonzero flag.
//
// One wavefront owns one to-reduce row at a time (persistent grid, rows
// taken from an atomic queue, longest rows first).  Its accumulator is a
// window of W uint64 columns in LDS.  Per window:
//   1. zero the window, scatter the row's own entries that fall in it;
//   2. deferred part: add x_j * P_j[window] for every pivot j already
//      applied in an earlier window (multipliers x_j kept in a per-wave
//      scratch row); work is load-balanced over the lanes and summed with
//      LDS atomics;
//   3. chain part: walk the window's pivot columns left to right -- a ballot
//      over 64 columns finds the next column whose value is nonzero mod p,
//      its multiplier x = p - v is stored and the pivot's entries inside the
//      window are added (entries right of the window are deferred);
//   4. the window's right-hand columns are reduced mod p and written out.
// The addends are non-negative, so the uint64 window is exact with lazy
// reduction (see common.h).
#include <hipcub/hipcub.hpp>

#include <climits>

#include "internal.h"

namespace gbgpu {

// ------------------------------------------------------------ split

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// lead of every row; first row per lead; contract errors into *err
__global__ void k_leads(int32_t nr, const int64_t *rp, const int32_t *ci, int32_t lim,
                        int32_t *first, int32_t *err) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
        const int64_t b = rp[i], e = rp[i + 1];
        if (e <= b) { atomicOr(err, 1); continue; }
        const int32_t lead = ci[b];
        if (lead < 0 || lead >= lim) { atomicOr(err, 2); continue; }
        atomicMin(&first[lead], i);
    }
}

// to-reduce flags; a row that breaks the contract (k_leads reports it) is
// not read past its end nor its lead used as an index
__global__ void k_nonpiv(int32_t nr, const int64_t *rp, const int32_t *ci, const int32_t *first,
                         int32_t lim, int32_t *flag) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
        const int64_t b = rp[i], e = rp[i + 1];
        const int32_t lead = e > b ? ci[b] : -1;
        flag[i] = (lead >= 0 && lead < lim) ? first[lead] != i : 0;
    }
}

// the to-reduce list is filled from its end (upivs[nrl - j],
// la_ff_32.c:1106-1116): reversed order of appearance
__global__ void k_upiv(int32_t nr, const int32_t *flag, const int32_t *rank, int32_t nrl,
                       int32_t *upiv) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x)
        if (flag[i]) upiv[nrl - 1 - rank[i]] = i;
}

static int grid_for(int64_t n, int bs) {
    int64_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > 65535 * 4) g = 65535 * 4;
    return (int)g;
}

int prep_split(gbgpu_mat *m, int32_t lim) {
    gbgpu_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    dfree(m->first);
    dfree(m->upiv);
    int rc;
    if ((rc = dalloc(ctx, &m->first, (size_t)m->nc))) return rc;
    int32_t *err = nullptr, *flag = nullptr, *rank = nullptr;
    if ((rc = dalloc(ctx, &err, 1))) return rc;
    if ((rc = dalloc(ctx, &flag, (size_t)m->nr)) || (rc = dalloc(ctx, &rank, (size_t)m->nr))) {
        dfree(err); dfree(flag); return rc;
    }
    k_fill_i32<<<grid_for(m->nc, 256), 256, 0, s>>>(m->first, m->nc, INT32_MAX);
    k_fill_i32<<<1, 64, 0, s>>>(err, 1, 0);
    k_leads<<<grid_for(m->nr, 256), 256, 0, s>>>(m->nr, m->rp, m->ci, lim, m->first, err);
    k_nonpiv<<<grid_for(m->nr, 256), 256, 0, s>>>(m->nr, m->rp, m->ci, m->first, lim, flag);
    // one wait: the count of rows to reduce and the contract flags
    int64_t nrl = 0;
    int32_t herr = 0;
    if ((rc = exclusive_scan_i32(ctx, flag, rank, m->nr, &nrl, err, &herr))) {
        dfree(err); dfree(flag); dfree(rank); return rc;
    }
    if (herr) {
        dfree(err); dfree(flag); dfree(rank);
        return set_error(ctx, GBGPU_ECONTRACT,
                         (herr & 1) ? "empty row in matrix"
                                    : "row lead outside the known-pivot columns");
    }
    m->nrl = (int32_t)nrl;
    if ((rc = dalloc(ctx, &m->upiv, (size_t)nrl))) {
        dfree(err); dfree(flag); dfree(rank); return rc;
    }
    k_upiv<<<grid_for(m->nr, 256), 256, 0, s>>>(m->nr, flag, rank, m->nrl, m->upiv);
    GBGPU_HIP(ctx, hipGetLastError());
    dfree_later(ctx, err); dfree_later(ctx, flag); dfree_later(ctx, rank);
    m->split_lim = lim;
    return GBGPU_OK;
}

// ---------------------------------------------------------- buckets

// pivot entry counts (lead excluded) and has-pivot bits per left column
__global__ void k_pivlen(int32_t ncl, const int32_t *first, const int64_t *rp, int64_t *plen,
                         uint32_t *pbits) {
    // one column per thread, no loop (the grid covers ncl): the ballot and
    // the lane-specific stores below must not sit inside a loop
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool has = false;
    if (c < ncl) {
        const int32_t r = first[c];
        has = r != INT32_MAX;
        plen[c] = has ? rp[r + 1] - rp[r] - 1 : 0;
    }
    const uint64_t bal = __ballot(has);
    const int lane = threadIdx.x & 63;
    // 64 consecutive columns per wave (blockDim is a multiple of 64)
    if (c < ncl && lane == 0) pbits[c >> 5] = (uint32_t)bal;
    if (c < ncl && lane == 32) pbits[c >> 5] = (uint32_t)(bal >> 32);
}

// one wave per pivot column: histogram by window (counters 0 .. nwin rounded
// up to whole chunks of 64, in dynamic LDS), scan chunk by chunk, scatter
constexpr int MAX_WIN = 4096;
__global__ __launch_bounds__(64) void k_bucket(int32_t ncl, int32_t nwin, int32_t wshift,
                                               const int32_t *first, const int64_t *rp,
                                               const int32_t *ci, const int32_t *rcf,
                                               const int64_t *cp, const uint32_t *cf,
                                               const int64_t *pbase, int64_t *seg, uint2 *bucket) {
    extern __shared__ int32_t pos[];
    const int lane = threadIdx.x;
    const int32_t nw64 = (nwin + 1 + 63) & ~63;  // counters 0 .. nwin, whole chunks
    for (int32_t c = blockIdx.x; c < ncl; c += gridDim.x) {
        const int32_t r = first[c];
        const int64_t base = pbase[c];
        for (int32_t i = lane; i < nw64; i += 64) pos[i] = 0;
        __syncthreads();
        int64_t b = 0, e = 0;
        const uint32_t *rc = nullptr;
        if (r != INT32_MAX) {
            b = rp[r] + 1;  // skip the lead
            e = rp[r + 1];
            rc = cf + cp[rcf[r]] - rp[r];
        }
        for (int64_t k = b + lane; k < e; k += 64) atomicAdd(&pos[ci[k] >> wshift], 1);
        __syncthreads();
        // exclusive scan of the counts; seg[w] for w = 0 .. nwin
        int32_t carry = 0;
        for (int32_t w0 = 0; w0 < nw64; w0 += 64) {
            const int32_t w = w0 + lane;
            const int v = w < nwin ? pos[w] : 0;
            const int incl = wave_incl_scan(v);
            const int32_t start = carry + incl - v;
            pos[w] = start;
            if (w <= nwin) seg[(int64_t)w * ncl + c] = base + start;
            carry += __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
        }
        __syncthreads();
        for (int64_t k = b + lane; k < e; k += 64) {
            const int32_t col = ci[k];
            const int w = col >> wshift;
            const int at = atomicAdd(&pos[w], 1);
            bucket[base + at] = make_uint2((uint32_t)col, rc[k]);
        }
        __syncthreads();
    }
}

int prep_buckets(gbgpu_mat *m, int force_wshift) {
    gbgpu_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    dfree(m->seg);
    dfree(m->bucket);
    dfree(m->pbits);
    // unless a width is forced: the smallest window (2048 .. 8192 columns)
    // that keeps nwin <= 63, else 8192-column windows, as many as it takes
    int wshift = 11;
    while (((int64_t)m->nc + (1 << wshift) - 1) >> wshift > 63 && wshift < 13) ++wshift;
    if (force_wshift) wshift = force_wshift;
    const int64_t nwin64 = ((int64_t)m->nc + (1 << wshift) - 1) >> wshift;
    if (nwin64 > MAX_WIN) return set_error(ctx, GBGPU_ERANGE, "too many columns for the sparse pass");
    const int nwin = (int)nwin64;
    m->wshift = wshift;
    m->nwin = nwin;
    const int32_t ncl = m->ncl;
    int rc;
    int64_t *plen = nullptr, *pbase = nullptr;
    if ((rc = dalloc(ctx, &plen, (size_t)ncl)) || (rc = dalloc(ctx, &pbase, (size_t)ncl))) {
        dfree(plen); return rc;
    }
    if ((rc = dalloc(ctx, &m->pbits, (size_t)(ncl + 63) / 32 + 2))) { dfree(plen); dfree(pbase); return rc; }
    GBGPU_HIP(ctx, hipMemsetAsync(m->pbits, 0, ((size_t)(ncl + 63) / 32 + 2) * 4, s));
    if (ncl > 0) {
        // round the grid up to whole waves so k_pivlen's ballots cover 64 columns
        const int blocks = (int)((ncl + 255) / 256);
        k_pivlen<<<blocks, 256, 0, s>>>(ncl, m->first, m->rp, plen, m->pbits);
    }
    int64_t total = 0;
    if ((rc = exclusive_scan_i64(ctx, plen, pbase, ncl, &total))) { dfree(plen); dfree(pbase); return rc; }
    m->nbucket = total;
    if ((rc = dalloc(ctx, &m->seg, (size_t)(nwin + 1) * (size_t)(ncl > 0 ? ncl : 1))) ||
        (rc = dalloc(ctx, &m->bucket, (size_t)total))) {
        dfree(plen); dfree(pbase); return rc;
    }
    if (ncl > 0) {
        const int g = ncl < 65536 * 4 ? ncl : 65536 * 4;
        const size_t lds = (size_t)((nwin + 1 + 63) & ~63) * sizeof(int32_t);
        k_bucket<<<g, 64, lds, s>>>(ncl, nwin, wshift, m->first, m->rp, m->ci, m->rcf, m->cp, m->cf,
                                  pbase, m->seg, m->bucket);
    }
    GBGPU_HIP(ctx, hipGetLastError());
    dfree_later(ctx, plen);
    dfree_later(ctx, pbase);
    return GBGPU_OK;
}

// pivot entries per (window, group), derived from the per-pivot window
// segments without atomics: one wave per group, lane t = pivot 64 g + t.
// An entry's top byte is the pivot's index modulo 256, i.e. within its
// aligned chunk of four groups; the low 24 bits are its column.
__global__ __launch_bounds__(64) void k_gbucket_count(int32_t ncl, int32_t ngroups, int32_t nwin,
                                                      const int64_t *seg, int64_t *gcnt) {
    const int lane = threadIdx.x;
    const int32_t g = blockIdx.x;
    const int32_t c = g * 64 + lane;
    for (int w = 0; w < nwin; ++w) {
        int len = 0;
        if (c < ncl) len = (int)(seg[(int64_t)(w + 1) * ncl + c] - seg[(int64_t)w * ncl + c]);
        const int tot = __builtin_amdgcn_readfirstlane(__shfl(wave_incl_scan(len), 63, 64));
        gcnt[(int64_t)w * ngroups + g] = tot;  // same value from every lane
    }
}

__global__ __launch_bounds__(64) void k_gbucket_fill(int32_t ncl, int32_t ngroups, int32_t nwin,
                                                     int32_t wshift, int32_t acc_shift,
                                                     const int64_t *seg, const uint2 *bucket,
                                                     const int64_t *gseg, uint2 *gbucket) {
    // one wave per (group, window): block (g, w)
    const int lane = threadIdx.x;
    const int32_t g = blockIdx.x;
    const int w = blockIdx.y;
    const int32_t c = g * 64 + lane;
    const uint32_t tag = ((uint32_t)(g & 3) * 64u + (uint32_t)lane) << 24;
    int64_t sb = 0;
    int len = 0;
    if (c < ncl) {
        sb = seg[(int64_t)w * ncl + c];
        len = (int)(seg[(int64_t)(w + 1) * ncl + c] - sb);
    }
    const int incl = wave_incl_scan(len);
    const int64_t at = gseg[(int64_t)w * ngroups + g] + (incl - len);
    const uint32_t w0 = (uint32_t)w << wshift;
    for (int k = 0; k < len; ++k) {
        const uint2 e = bucket[sb + k];
        gbucket[at + k] = make_uint2(tag | ((e.x - w0) << acc_shift), e.y);
    }
}

int prep_group_buckets(gbgpu_mat *m, int acc_shift) {
    gbgpu_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    dfree(m->gseg);
    dfree(m->gbucket);
    const int32_t ngroups = (m->ncl + 63) / 64;
    const int64_t nseg = (int64_t)m->nwin * ngroups;
    if (m->nc >= (1 << 24)) return set_error(ctx, GBGPU_ERANGE, "too many columns for the blocked pass");
    int rc;
    int64_t *gcnt = nullptr;
    if ((rc = dalloc(ctx, &gcnt, (size_t)nseg + 1))) return rc;
    if ((rc = dalloc(ctx, &m->gseg, (size_t)nseg + 1))) { dfree(gcnt); return rc; }
    GBGPU_HIP(ctx, hipMemsetAsync(gcnt, 0, ((size_t)nseg + 1) * sizeof(int64_t), s));
    if (ngroups > 0) k_gbucket_count<<<ngroups, 64, 0, s>>>(m->ncl, ngroups, m->nwin, m->seg, gcnt);
    // the (window, group) layout holds the per-pivot layout's entries in
    // another order: as many (no wait for the scan's total)
    rc = exclusive_scan_i64(ctx, gcnt, m->gseg, nseg + 1, nullptr);
    dfree_later(ctx, gcnt);
    if (rc) return rc;
    if ((rc = dalloc(ctx, &m->gbucket, (size_t)m->nbucket))) return rc;
    if (ngroups > 0)
        k_gbucket_fill<<<dim3((unsigned)ngroups, (unsigned)m->nwin), 64, 0, s>>>(
            m->ncl, ngroups, m->nwin, m->wshift, acc_shift, m->seg, m->bucket, m->gseg, m->gbucket);
    GBGPU_HIP(ctx, hipGetLastError());
    return GBGPU_OK;
}

// ------------------------------------------------------ sparse pass


// pivot entries staged per 64-column group (6 KB of LDS)
constexpr int STAGE = 768;

// LDS of one wave of the plain pass
template <int WSHIFT>
struct PlainLds {
    uint64_t acc[1 << WSHIFT];
    uint32_t pb[(1 << WSHIFT) / 32];
    int64_t sh_base[64];
    int32_t sh_start[64];
    uint32_t sh_x[64];
    uint2 stage[STAGE];
};

// reduce to-reduce row number q of a (a wave-uniform value) into
// dprime[q] and dnz[q]; xs is this wave's multiplier scratch (ncl entries)
template <int WSHIFT, bool BIG, bool COUNT>
__device__ __forceinline__ void plain_row(const SpArgs &a, int q, uint32_t *xs,
                                          PlainLds<WSHIFT> &L) {
    constexpr int W = 1 << WSHIFT;
    uint64_t *acc = L.acc;
    uint32_t *pb = L.pb;
    int64_t *sh_base = L.sh_base;
    int32_t *sh_start = L.sh_start;
    uint32_t *sh_x = L.sh_x;
    uint2 *stage = L.stage;
    const int lane = threadIdx.x & 63;
    const uint32_t p = a.md.p;
    {
        // per-row values are the same in every lane: keep them scalar so that
        // the loops below (which hold ballots and shuffles) stay uniform
        const int32_t row = __builtin_amdgcn_readfirstlane(a.upiv[q]);
        const int64_t rb = a.rp[row], re = a.rp[row + 1];
        const uint32_t *rcf = a.cf + a.cp[a.rcf[row]] - rb;
        const int32_t lead = __builtin_amdgcn_readfirstlane(a.ci[rb]);
        for (int32_t j = lead + lane; j < a.ncl; j += 64) xs[j] = 0;
        bool nzrow = false;

        for (int w = lead >> WSHIFT; w < a.nwin; ++w) {
            const int32_t w0 = w << WSHIFT;
            const int32_t w1 = min(w0 + W, a.nc);
            for (int c = lane; c < W; c += 64) acc[c] = 0;
            // has-pivot bits of the window's left columns
            const int32_t lend = min(w1, a.ncl);
            for (int k = lane; k < W / 32; k += 64)
                pb[k] = (w0 + 32 * k < lend) ? a.pbits[(w0 >> 5) + k] : 0u;
            __syncthreads();
            // 1. the row's own entries in this window
            for (int64_t k = rb + lane; k < re; k += 64) {
                const int32_t c = a.ci[k];
                if (c >= w0 && c < w1) acc[c - w0] = rcf[k];
            }
            __syncthreads();
            // 2. deferred contributions of pivots applied in earlier windows
            const int32_t jl = min(w0, a.ncl);
            for (int32_t jb = lead; jb < jl; jb += 64) {
                const int32_t j = jb + lane;
                uint32_t x = 0;
                int64_t s = 0;
                int len = 0;
                if (j < jl) {
                    x = xs[j];
                    if (x) {
                        s = a.seg[(int64_t)w * a.ncl + j];
                        len = (int)(a.seg[(int64_t)(w + 1) * a.ncl + j] - s);
                    }
                }
                const int incl = wave_incl_scan(len);
                const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
                if (total == 0) continue;
                sh_start[lane] = incl - len;
                sh_base[lane] = s;
                sh_x[lane] = x;
                __syncthreads();
                for (int f = lane; f < total; f += 64) {
                    int lo = 0, hi = 63;  // last t with sh_start[t] <= f
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (sh_start[mid] <= f) lo = mid; else hi = mid - 1;
                    }
                    const uint2 e = a.bucket[sh_base[lo] + (f - sh_start[lo])];
                    atomicAdd(&acc[e.x - w0], addend<BIG>(sh_x[lo], e.y, a.md));
                }
                __syncthreads();
            }
            // 3. chain over the window's pivot columns, one 64-column group
            //    at a time; the group's in-window pivot entries are staged
            //    in LDS first (one coalesced sweep) so that a chain step
            //    touches LDS only
            for (int32_t g = max(w0, lead) & ~63; g < lend; g += 64) {
                const int32_t c = g + lane;
                int64_t s = 0;
                int len = 0;
                if (c < lend && c >= lead) {
                    s = a.seg[(int64_t)w * a.ncl + c];
                    len = (int)(a.seg[(int64_t)(w + 1) * a.ncl + c] - s);
                }
                const int incl = wave_incl_scan(len);
                const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
                const bool staged = total <= STAGE;
                sh_start[lane] = incl - len;
                sh_base[lane] = s;
                __syncthreads();
                if (staged) {
                    for (int f = lane; f < total; f += 64) {
                        int lo = 0, hi = 63;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (sh_start[mid] <= f) lo = mid; else hi = mid - 1;
                        }
                        stage[f] = a.bucket[sh_base[lo] + (f - sh_start[lo])];
                    }
                    __syncthreads();
                }
                int32_t i = max(g, lead);
                const int32_t gend = min(g + 64, lend);
                while (i < gend) {
                    const int32_t cc = i + lane;
                    uint32_t r = 0;
                    if (cc < gend) {
                        const uint64_t v = acc[cc - w0];
                        if (v) r = a.md.reduce(v);
                        if (r && !((pb[(cc - w0) >> 5] >> ((cc - w0) & 31)) & 1u)) r = 0;
                    }
                    const uint64_t mask = __ballot(r != 0);
                    if (mask == 0) break;
                    const int t = __ffsll((unsigned long long)mask) - 1;
                    const int32_t col = i + t;
                    const uint32_t x = p - __shfl(r, t, 64);
                    const int tg = col - g;
                    const int st = sh_start[tg];
                    const int ln = (tg < 63 ? sh_start[tg + 1] : total) - st;
                    // every lane stores the same value (no lane-0-only branch)
                    xs[col] = x;
                    acc[col - w0] = 0;
                    if (COUNT) {
                        // algorithmic work: the whole pivot row (lead included);
                        // lane 0 contributes, the others add zero (no branch)
                        const uint64_t len = (uint64_t)(a.seg[(int64_t)a.nwin * a.ncl + col] - a.seg[col]) + 1;
                        atomicAdd(&a.work[0], lane == 0 ? 1ull : 0ull);
                        atomicAdd(&a.work[1], lane == 0 ? (unsigned long long)len : 0ull);
                    }
                    if (staged) {
                        for (int k = lane; k < ln; k += 64) {
                            const uint2 ent = stage[st + k];
                            acc[ent.x - w0] += addend<BIG>(x, ent.y, a.md);
                        }
                    } else {
                        const int64_t sb = sh_base[tg];
                        for (int k = lane; k < ln; k += 64) {
                            const uint2 ent = a.bucket[sb + k];
                            acc[ent.x - w0] += addend<BIG>(x, ent.y, a.md);
                        }
                    }
                    __syncthreads();
                    i = col + 1;
                }
                __syncthreads();
            }
            // 4. right-hand columns of this window
            if (w1 > a.ncl) {
                const int32_t cs = max(w0, a.ncl);
                uint32_t *out = a.dprime + (size_t)q * (size_t)a.ncr - a.ncl;
                for (int32_t c = cs + lane; c < w1; c += 64) {
                    const uint32_t r = a.md.reduce(acc[c - w0]);
                    out[c] = r;
                    nzrow |= r != 0;
                }
            }
            __syncthreads();
        }
        const bool anynz = __any(nzrow);
        a.dnz[q] = anynz ? 1 : 0;  // same value from every lane
    }
}

template <int WSHIFT, bool BIG, bool COUNT>
__global__ __launch_bounds__(64) void k_sparse_plain(SpArgs a) {
    __shared__ PlainLds<WSHIFT> L;
    const int lane = threadIdx.x;
    uint32_t *xs = a.xs + (size_t)blockIdx.x * (size_t)(a.ncl > 0 ? a.ncl : 1);
    for (;;) {
        // No lane-specific code may sit in this loop: with a lane-0-only
        // atomic here the compiler gives lanes 1..63 a back edge that skips
        // it (in a one-wave workgroup __syncthreads() is only a wave
        // barrier), and they re-read a stale row index forever.  Every lane
        // issues the atomic (lane 0 adds 1, the others 0) and lane 0's
        // result is broadcast as a scalar.
        const int q0 = atomicAdd(a.queue, lane == 0 ? 1 : 0);
        const int q = __builtin_amdgcn_readfirstlane(q0);
        if (q >= a.nrl) break;
        // upiv ends with the smallest leads: longest rows first
        plain_row<WSHIFT, BIG, COUNT>(a, a.nrl - 1 - q, xs, L);
    }
}

// ------------------------------------------------ blocked sparse pass
//
// The chain of the plain pass (one pivot column after the other) is what
// limits it: ~5400 dependent steps per row on an F4 matrix of 30000
// columns.  Inside one 64-column group the dependencies only run through
// the pivots' entries that fall in the same group: with T the group's unit
// upper-triangular block (T[i][j] = entry of pivot g+i in column g+j, both
// pivot columns), the group's multipliers are x = -v T^-1 mod p, where v is
// the group's accumulator before any of its pivots is applied.  T^-1 is
// computed once per matrix, so a group costs one 64 x 64 vector-matrix
// product plus a fully parallel scatter of the group pivots' remaining
// entries.  Columns without a pivot get an identity row/column in T and
// v = 0, hence x = 0 -- the reference skips them too (la_ff_32.c:517-519).

// one wave per 64-column group: T from the pivots' in-group entries, then
// T^-1 row by row (lane i owns row i: Y[i][j] = -sum_{i<=k<j} Y[i][k] T[k][j])
template <bool BIG>
__global__ __launch_bounds__(64) void k_group_inverse(int32_t ncl, int32_t wshift,
                                                      const uint32_t *pbits, const int64_t *seg,
                                                      const uint2 *bucket, Mod md, uint32_t *tinv) {
    __shared__ uint32_t T[64][65];
    __shared__ uint32_t Y[64][65];
    const int lane = threadIdx.x;
    const int32_t g0 = blockIdx.x * 64;
    const int w = g0 >> wshift;  // a group never straddles a window
    for (int k = 0; k < 64; ++k) {
        T[k][lane] = (k == lane) ? 1u : 0u;
        Y[k][lane] = (k == lane) ? 1u : 0u;
    }
    const int32_t c = g0 + lane;
    const bool has = c < ncl && ((pbits[c >> 5] >> (c & 31)) & 1u);
    const uint64_t hmask = __ballot(has);
    __syncthreads();
    if (has) {
        const int64_t sb = seg[(int64_t)w * ncl + c], se = seg[(int64_t)(w + 1) * ncl + c];
        for (int64_t k = sb; k < se; ++k) {
            const uint2 ent = bucket[k];
            const int j = (int)ent.x - g0;
            if (j > lane && j < 64 && ((hmask >> j) & 1ull)) T[lane][j] = ent.y;
        }
    }
    __syncthreads();
    // row lane of the inverse, left to right; p < 2^17: the products (below
    // 2^34, at most 63 of them) are summed unreduced
    for (int j = lane + 1; j < 64; ++j) {
        uint64_t acc = 0;
        for (int k = lane; k < j; ++k)
            acc += BIG ? (uint64_t)md.mul(Y[lane][k], T[k][j]) : (uint64_t)Y[lane][k] * T[k][j];
        const uint32_t r = md.reduce(acc);
        Y[lane][j] = r ? md.p - r : 0u;
    }
    __syncthreads();
    uint32_t *out = tinv + (int64_t)blockIdx.x * 4096;
    for (int k = 0; k < 64; ++k) out[k * 64 + lane] = Y[k][lane];
}

// The entries of the (window, group) layout: x = (t << 24) | the byte
// offset of the entry's column in the window's accumulator (column - w0
// times the accumulator's size), y = the coefficient; t is the pivot's
// index within its aligned chunk of four groups, its multiplier x_t in
// sh_x[t].
//
// Accumulators: uint64 with the addends of common.h, or, for p < 2^16 when
// (ncl + 1) (2p - 2) < 2^32 (the host checks it), uint32 with addends
// below 2p: x c - floor(x' c / 2^32) p with x' = floor(x 2^32 / p) (Shoup's
// precomputed quotient, kept beside x in sh_xp), which lies in [0, 2p) for
// c < 2^32.  A column of a window receives at most one addend per pivot
// plus the row's own entry.
template <bool BIG, typename ACC>
__device__ __forceinline__ void apply_entry(ACC *acc, uint2 en, const uint32_t *sh_x,
                                            const uint32_t *sh_xp, const Mod &md) {
    const uint32_t t = en.x >> 24;
    ACC *dst = (ACC *)((char *)acc + (en.x & 0xFFFFFFu));
    if constexpr (sizeof(ACC) == 4) {
        const uint32_t x = sh_x[t], xp = sh_xp[t];
        const uint32_t q = __umulhi(xp, en.y);
        atomicAdd(dst, __umul24(x, en.y) - __umul24(q, md.p));
    } else {
        atomicAdd(dst, addend<BIG>(sh_x[t], en.y, md));
    }
}

// x' = floor(x 2^32 / p) for x < p < 2^16, in two exact 32-bit divisions
__device__ __forceinline__ uint32_t shoup16(uint32_t x, uint32_t p) {
    const uint32_t hi = (x << 16) / p, r = (x << 16) - hi * p;
    return (hi << 16) + (r << 16) / p;
}

__device__ __forceinline__ void load_batch(uint2 (&r)[8], const uint2 *src) {
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = src[64 * u];
}

template <bool BIG, typename ACC>
__device__ __forceinline__ void apply_batch(ACC *acc, const uint2 (&r)[8], const uint32_t *sh_x,
                                            const uint32_t *sh_xp, const Mod &md) {
#pragma unroll
    for (int u = 0; u < 8; ++u) apply_entry<BIG>(acc, r[u], sh_x, sh_xp, md);
}

// the entries [b, e): a coalesced stream in batches of 512 entries (eight
// per lane); two register sets in turn, so that while one batch is applied
// the next one's loads are in flight (every load of the steady loop is
// unconditional)
template <bool BIG, typename ACC>
__device__ __forceinline__ void stream_range(ACC *acc, int64_t b, int64_t e, const uint2 *gbucket,
                                             const uint32_t *sh_x, const uint32_t *sh_xp,
                                             const Mod &md) {
    const int lane = threadIdx.x & 63;
    const int64_t nb = (e - b) >> 9;  // full batches (wave-uniform)
    const uint2 *src = gbucket + b + lane;
    uint2 ra[8], rb[8];
    if (nb > 0) load_batch(ra, src);
    int64_t i = 0;
    for (; i + 2 < nb; i += 2) {
        load_batch(rb, src + 512 * (i + 1));
        apply_batch<BIG>(acc, ra, sh_x, sh_xp, md);
        load_batch(ra, src + 512 * (i + 2));
        apply_batch<BIG>(acc, rb, sh_x, sh_xp, md);
    }
    if (nb - i == 2) {
        load_batch(rb, src + 512 * (i + 1));
        apply_batch<BIG>(acc, ra, sh_x, sh_xp, md);
        apply_batch<BIG>(acc, rb, sh_x, sh_xp, md);
    } else if (nb - i == 1) {
        apply_batch<BIG>(acc, ra, sh_x, sh_xp, md);
    }
    for (int64_t k = b + (nb << 9) + lane; k < e; k += 64)
        apply_entry<BIG>(acc, gbucket[k], sh_x, sh_xp, md);
}

// one group's multipliers x (lane t: pivot 64 g + t; xp their Shoup
// quotients for 32-bit accumulators) into their slot, then its entries
// [b, e)
template <bool BIG, typename ACC>
__device__ __forceinline__ void stream_group(ACC *acc, uint32_t x, uint32_t xp, int slot, int64_t b,
                                             int64_t e, const uint2 *gbucket, uint32_t *sh_x,
                                             uint32_t *sh_xp, const Mod &md) {
    const int lane = threadIdx.x & 63;
    if (!__any(x != 0) || e <= b) return;
    sh_x[slot * 64 + lane] = x;
    if constexpr (sizeof(ACC) == 4) sh_xp[slot * 64 + lane] = xp;
    __syncthreads();
    stream_range<BIG>(acc, b, e, gbucket, sh_x, sh_xp, md);
    __syncthreads();
}

// LDS of one wave of the blocked pass
template <int WSHIFT, typename ACC>
struct BlockedLds {
    ACC acc[1 << WSHIFT];
    uint32_t pb[(1 << WSHIFT) / 32];
    uint32_t sh_x[256];                           // multipliers of one chunk of four groups
    uint32_t sh_xp[sizeof(ACC) == 4 ? 256 : 1];   // their Shoup quotients (32-bit accumulators)
};

template <int WSHIFT, bool BIG, typename ACC>
__device__ __forceinline__ void blocked_row(const SpArgs &a, int q, uint32_t *xs,
                                            BlockedLds<WSHIFT, ACC> &L, const uint32_t *tinv,
                                            const int64_t *gseg, const uint2 *gbucket,
                                            int32_t ngroups) {
    constexpr int W = 1 << WSHIFT;
    constexpr bool A32 = sizeof(ACC) == 4;
    ACC *acc = L.acc;
    uint32_t *pb = L.pb;
    uint32_t *sh_x = L.sh_x;
    uint32_t *sh_xp = L.sh_xp;
    uint32_t *xs2 = A32 ? a.xs2 + (xs - a.xs) : nullptr;  // this wave's Shoup quotients
    const int lane = threadIdx.x & 63;
    const uint32_t p = a.md.p;
    {
        const int32_t row = __builtin_amdgcn_readfirstlane(a.upiv[q]);
        const int64_t rb = a.rp[row], re = a.rp[row + 1];
        const uint32_t *rcf = a.cf + a.cp[a.rcf[row]] - rb;
        const int32_t lead = __builtin_amdgcn_readfirstlane(a.ci[rb]);
        const int32_t g_lead = lead >> 6;
        bool nzrow = false;

        for (int w = lead >> WSHIFT; w < a.nwin; ++w) {
            const int32_t w0 = w << WSHIFT;
            const int32_t w1 = min(w0 + W, a.nc);
            const int32_t lend = min(w1, a.ncl);
            const int64_t *gs = gseg + (int64_t)w * ngroups;
            for (int c = lane; c < W; c += 64) acc[c] = 0;
            for (int k = lane; k < W / 32; k += 64)
                pb[k] = (w0 + 32 * k < lend) ? a.pbits[(w0 >> 5) + k] : 0u;
            __syncthreads();
            for (int64_t k = rb + lane; k < re; k += 64) {
                const int32_t c = a.ci[k];
                if (c >= w0 && c < w1) acc[c - w0] = rcf[k];
            }
            __syncthreads();
            // groups of earlier windows (their multipliers are in xs; the
            // last group of pivot columns may be partial), in aligned chunks
            // of four groups: the entries of consecutive groups of one
            // window are consecutive in the layout, so the chunk's groups
            // with a nonzero multiplier form a few contiguous streams.  The
            // next chunk's multipliers are loaded before this chunk streams.
            const int32_t jl = min(w0, a.ncl);
            const int32_t g_end = (jl + 63) >> 6;
            const int32_t jlo = g_lead << 6;
            uint32_t xn[4], xpn[4];
            int32_t c0 = g_lead & ~3;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t j = (c0 + u) * 64 + lane;
                const bool in = c0 < g_end && j >= jlo && j < jl;
                xn[u] = in ? xs[j] : 0u;
                xpn[u] = A32 && in ? xs2[j] : 0u;
            }
            for (; c0 < g_end; c0 += 4) {
                uint32_t xv[4], xpv[4];
                uint32_t nz = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    xv[u] = xn[u];
                    xpv[u] = xpn[u];
                    nz |= (__any(xv[u] != 0) ? 1u : 0u) << u;
                }
                const int32_t c1 = c0 + 4;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int32_t j = (c1 + u) * 64 + lane;
                    const bool in = c1 < g_end && j < jl;
                    xn[u] = in ? xs[j] : 0u;
                    xpn[u] = A32 && in ? xs2[j] : 0u;
                }
                if (nz == 0) continue;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    sh_x[u * 64 + lane] = xv[u];
                    if (A32) sh_xp[u * 64 + lane] = xpv[u];
                }
                __syncthreads();
                // runs of consecutive groups with a nonzero multiplier
                while (nz) {
                    const int ga = __builtin_ctz(nz);
                    int gb = ga;
                    while (gb < 4 && ((nz >> gb) & 1u)) ++gb;
                    nz &= ~(((1u << gb) - 1u) & ~((1u << ga) - 1u));
                    stream_range<BIG>(acc, gs[c0 + ga], gs[c0 + gb], gbucket, sh_x, sh_xp, a.md);
                }
                __syncthreads();
            }
            // this window's groups
            for (int32_t g = max(w0, g_lead << 6); g < lend; g += 64) {
                const int32_t c = g + lane;
                uint32_t u = 0;
                if (c < lend) {
                    const uint64_t v = acc[c - w0];
                    const uint32_t r = v ? a.md.reduce(v) : 0u;
                    if (r && ((pb[(c - w0) >> 5] >> ((c - w0) & 31)) & 1u)) u = p - r;
                }
                // x = u T^-1 (u = -v mod p); lane j owns column j
                const uint32_t *tp = tinv + (int64_t)(g >> 6) * 4096 + lane;
                uint32_t tcol[64];
#pragma unroll
                for (int i = 0; i < 64; ++i) tcol[i] = tp[i * 64];
                uint64_t sx = 0;
#pragma unroll
                for (int i = 0; i < 64; ++i) {
                    const uint32_t ui = __builtin_amdgcn_readlane(u, i);
                    if (BIG) sx += (uint64_t)a.md.mul(ui, tcol[i]);
                    else sx += (uint64_t)ui * (uint64_t)tcol[i];
                }
                uint32_t x = a.md.reduce(sx);
                if (c >= lend) x = 0;
                const uint32_t xp = A32 ? shoup16(x, p) : 0u;
                if (c < lend) {
                    xs[c] = x;
                    if (A32) xs2[c] = xp;
                }
                // also adds the group's own in-group entries: harmless (its
                // accumulator values are consumed) and needed for right-hand
                // columns sharing the last partial group
                stream_group<BIG>(acc, x, xp, (g >> 6) & 3, gs[g >> 6], gs[(g >> 6) + 1], gbucket,
                                  sh_x, sh_xp, a.md);
            }
            if (w1 > a.ncl) {
                const int32_t cs = max(w0, a.ncl);
                uint32_t *out = a.dprime + (size_t)q * (size_t)a.ncr - a.ncl;
                for (int32_t c = cs + lane; c < w1; c += 64) {
                    const uint32_t r = a.md.reduce(acc[c - w0]);
                    out[c] = r;
                    nzrow |= r != 0;
                }
            }
            __syncthreads();
        }
        const bool anynz = __any(nzrow);
        a.dnz[q] = anynz ? 1 : 0;
    }
}

template <int WSHIFT, bool BIG, typename ACC>
__global__ __launch_bounds__(64) void k_sparse_blocked(SpArgs a, const uint32_t *tinv,
                                                       const int64_t *gseg, const uint2 *gbucket,
                                                       int32_t ngroups) {
    __shared__ BlockedLds<WSHIFT, ACC> L;
    const int lane = threadIdx.x;
    uint32_t *xs = a.xs + (size_t)blockIdx.x * (size_t)(a.ncl > 0 ? a.ncl : 1);
    for (;;) {
        // row index broadcast as a scalar (see k_sparse_plain)
        const int q0 = atomicAdd(a.queue, lane == 0 ? 1 : 0);
        const int q = __builtin_amdgcn_readfirstlane(q0);
        if (q >= a.nrl) break;
        blocked_row<WSHIFT, BIG, ACC>(a, a.nrl - 1 - q, xs, L, tinv, gseg, gbucket, ngroups);
    }
}

// ------------------------------------------------------ batched passes
//
// One persistent launch over the to-reduce rows of many matrices: the
// queue runs over the concatenation of the matrices' rows (row_off holds
// the prefix sums, matrices in the caller's order), each wave takes the
// next row, finds its matrix by a scalar binary search and reduces it with
// that matrix's arguments (its own prime among them).  xs holds xs_stride
// multipliers per wave (the largest ncl of the batch).

__device__ __forceinline__ int batch_find(const int32_t *row_off, int32_t nmat, int32_t q) {
    int lo = 0, hi = nmat - 1;  // last i with row_off[i] <= q
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_off[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <int WSHIFT, bool BIG>
__global__ __launch_bounds__(64) void k_sparse_blocked_batch(const BlockedArgs *mats,
                                                             const int32_t *row_off,
                                                             int32_t nmat, int32_t *queue,
                                                             uint32_t *xs_all,
                                                             int64_t xs_stride) {
    __shared__ BlockedLds<WSHIFT, uint64_t> L;
    const int lane = threadIdx.x;
    uint32_t *xs = xs_all + (int64_t)blockIdx.x * xs_stride;
    const int32_t total = __builtin_amdgcn_readfirstlane(row_off[nmat]);
    for (;;) {
        const int q0 = atomicAdd(queue, lane == 0 ? 1 : 0);
        const int q = __builtin_amdgcn_readfirstlane(q0);
        if (q >= total) break;
        const int i = batch_find(row_off, nmat, q);
        const BlockedArgs &b = mats[i];
        const int32_t nrl = __builtin_amdgcn_readfirstlane(b.a.nrl);
        const int32_t r = nrl - 1 - (q - __builtin_amdgcn_readfirstlane(row_off[i]));
        blocked_row<WSHIFT, BIG, uint64_t>(b.a, r, xs, L, b.tinv, b.gseg, b.gbucket, b.ngroups);
    }
}

template <int WSHIFT, bool BIG>
__global__ __launch_bounds__(64) void k_sparse_plain_batch(const BlockedArgs *mats,
                                                           const int32_t *row_off, int32_t nmat,
                                                           int32_t *queue, uint32_t *xs_all,
                                                           int64_t xs_stride) {
    __shared__ PlainLds<WSHIFT> L;
    const int lane = threadIdx.x;
    uint32_t *xs = xs_all + (int64_t)blockIdx.x * xs_stride;
    const int32_t total = __builtin_amdgcn_readfirstlane(row_off[nmat]);
    for (;;) {
        const int q0 = atomicAdd(queue, lane == 0 ? 1 : 0);
        const int q = __builtin_amdgcn_readfirstlane(q0);
        if (q >= total) break;
        const int i = batch_find(row_off, nmat, q);
        const SpArgs &a = mats[i].a;
        const int32_t nrl = __builtin_amdgcn_readfirstlane(a.nrl);
        const int32_t r = nrl - 1 - (q - __builtin_amdgcn_readfirstlane(row_off[i]));
        plain_row<WSHIFT, BIG, false>(a, r, xs, L);
    }
}

__global__ void k_compact(int32_t n, const int32_t *flag, const int32_t *rank, int32_t *list) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (flag[i]) list[rank[i]] = i;
}

template <int WS>
static void launch_sparse(const SpArgs &a, bool big, bool count, int grid, hipStream_t s) {
    if (count) {
        if (big) k_sparse_plain<WS, true, true><<<grid, 64, 0, s>>>(a);
        else k_sparse_plain<WS, false, true><<<grid, 64, 0, s>>>(a);
    } else {
        if (big) k_sparse_plain<WS, true, false><<<grid, 64, 0, s>>>(a);
        else k_sparse_plain<WS, false, false><<<grid, 64, 0, s>>>(a);
    }
}

// LDS bytes of one wave of the blocked pass
template <typename ACC>
static size_t blocked_lds_bytes(int wshift) {
    switch (wshift) {
        case 10: return sizeof(BlockedLds<10, ACC>);
        case 11: return sizeof(BlockedLds<11, ACC>);
        case 12: return sizeof(BlockedLds<12, ACC>);
        default: return sizeof(BlockedLds<13, ACC>);
    }
}

template <int WS>
static void launch_blocked(const SpArgs &a, const uint32_t *tinv, const int64_t *gseg,
                           const uint2 *gbucket, int32_t ngroups, bool big, bool acc32, int grid,
                           hipStream_t s) {
    if (big) k_sparse_blocked<WS, true, uint64_t><<<grid, 64, 0, s>>>(a, tinv, gseg, gbucket, ngroups);
    else if (acc32) k_sparse_blocked<WS, false, uint32_t><<<grid, 64, 0, s>>>(a, tinv, gseg, gbucket, ngroups);
    else k_sparse_blocked<WS, false, uint64_t><<<grid, 64, 0, s>>>(a, tinv, gseg, gbucket, ngroups);
}

void sparse_args(const gbgpu_mat *m, SpArgs *a) {
    a->nc = m->nc; a->ncl = m->ncl; a->ncr = m->nc - m->ncl; a->nrl = m->nrl; a->nwin = m->nwin;
    a->rp = m->rp; a->ci = m->ci; a->rcf = m->rcf; a->cp = m->cp; a->cf = m->cf;
    a->upiv = m->upiv; a->pbits = m->pbits; a->seg = m->seg; a->bucket = m->bucket;
    a->xs = nullptr; a->xs2 = nullptr; a->dprime = m->dprime; a->dnz = m->dnz;
    a->queue = nullptr; a->work = nullptr;
    a->md = Mod::make(m->p);
}

// the inverses of the groups' pivot blocks (they need the per-pivot layout
// only) on stream sx
static void launch_group_inverse(gbgpu_mat *m, uint32_t *tinv, hipStream_t sx) {
    const int32_t ngroups = (m->ncl + 63) / 64;
    const Mod md = Mod::make(m->p);
    if (m->p >= (1u << 17))
        k_group_inverse<true><<<ngroups, 64, 0, sx>>>(m->ncl, m->wshift, m->pbits, m->seg, m->bucket, md, tinv);
    else
        k_group_inverse<false><<<ngroups, 64, 0, sx>>>(m->ncl, m->wshift, m->pbits, m->seg, m->bucket, md, tinv);
}

int prep_blocked(gbgpu_mat *m, uint32_t **tinv) {
    gbgpu_ctx *ctx = m->ctx;
    int rc;
    *tinv = nullptr;
    if ((rc = prep_group_buckets(m))) return rc;
    const int32_t ngroups = (m->ncl + 63) / 64;
    if (ngroups > 0) {
        if ((rc = dalloc(ctx, tinv, (size_t)ngroups * 4096))) return rc;
        launch_group_inverse(m, *tinv, ctx->stream);
        GBGPU_HIP(ctx, hipGetLastError());
    }
    return GBGPU_OK;
}

int batch_sparse_launch(gbgpu_ctx *ctx, const BlockedArgs *mats, const int32_t *row_off,
                        int32_t nmat, int32_t *queue, uint32_t *xs, int64_t xs_stride, int wshift,
                        bool big, int form, int grid) {
    hipStream_t s = ctx->stream;
    int rc;
#define GBGPU_BATCH_SP(WS)                                                                       \
    do {                                                                                         \
        const void *fn = form == GBGPU_FORM_MFMA                                                 \
            ? (big ? (const void *)k_sparse_blocked_batch<WS, true> : (const void *)k_sparse_blocked_batch<WS, false>) \
            : (big ? (const void *)k_sparse_plain_batch<WS, true> : (const void *)k_sparse_plain_batch<WS, false>);   \
        if ((rc = check_launch(ctx, fn, grid, 64, "batch sparse pass"))) return rc;            \
        if (form == GBGPU_FORM_MFMA) {                                                           \
            if (big) k_sparse_blocked_batch<WS, true><<<grid, 64, 0, s>>>(mats, row_off, nmat, queue, xs, xs_stride); \
            else k_sparse_blocked_batch<WS, false><<<grid, 64, 0, s>>>(mats, row_off, nmat, queue, xs, xs_stride); \
        } else {                                                                                 \
            if (big) k_sparse_plain_batch<WS, true><<<grid, 64, 0, s>>>(mats, row_off, nmat, queue, xs, xs_stride); \
            else k_sparse_plain_batch<WS, false><<<grid, 64, 0, s>>>(mats, row_off, nmat, queue, xs, xs_stride); \
        }                                                                                        \
    } while (0)
    switch (wshift) {
        case 10: GBGPU_BATCH_SP(10); break;
        case 11: GBGPU_BATCH_SP(11); break;
        case 12: GBGPU_BATCH_SP(12); break;
        default: GBGPU_BATCH_SP(13); break;
    }
#undef GBGPU_BATCH_SP
    GBGPU_HIP(ctx, hipGetLastError());
    return GBGPU_OK;
}

int sparse_wave_lds(int wshift, int form) {
    const int W = 1 << wshift;
    return W * 8 + (form == GBGPU_FORM_MFMA ? 0 : STAGE * 8) + 2048;
}

int run_sparse_pass(gbgpu_mat *m, uint64_t *work, int form) {
    gbgpu_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = prep_split(m, m->ncl))) return rc;
    // the kernels keep one uint64 per window column in LDS: windows of 1024
    // columns (when at most 63 of them cover the matrix) fit more waves per
    // CU than 2048-column ones (blocked: 16 against 9, 12 % faster on a
    // 20000 x 30000 matrix)
    // the blocked form with 32-bit accumulators (p < 2^16 and at most
    // (2^32 - 1) / (2p - 2) addends per column: ncl pivots and the row's own
    // entry) keeps 2048 columns in the space of 1024: half the windows, so
    // half the passes over the earlier groups' entries, at the same
    // occupancy
    const bool acc32 = form == GBGPU_FORM_MFMA && work == nullptr && m->p < (1u << 16) &&
                       ((uint64_t)m->ncl + 1) * (2ull * m->p - 2) < (1ull << 32);
    int wshift = 0;
    if (((int64_t)m->nc + 1023) / 1024 <= 63) wshift = 10;
    if (acc32 && ((int64_t)m->nc + 2047) / 2048 <= 63) wshift = 11;
    if ((rc = prep_buckets(m, wshift))) return rc;
    const int32_t ncr = m->nc - m->ncl;
    dfree(m->dprime); dfree(m->dnz); dfree(m->dlist);
    m->nd = 0;
    if ((rc = dalloc(ctx, &m->dprime, (size_t)m->nrl * (size_t)ncr)) ||
        (rc = dalloc(ctx, &m->dnz, (size_t)m->nrl)))
        return rc;
    // the blocked form needs the inverses of the groups' pivot blocks; on
    // matrices with dense enough pivot rows it runs on the matrix cores
    const bool blocked = form == GBGPU_FORM_MFMA && work == nullptr;
    const bool dense = blocked && dense_route_pays(m);
    uint32_t *tinv = nullptr;
    const int32_t ngroups = (m->ncl + 63) / 64;
    if (blocked && ngroups > 0) {
        if ((rc = dalloc(ctx, &tinv, (size_t)ngroups * 4096))) return rc;
        if (!dense) {
            // the inverses run on the side stream while the (window, group)
            // layout is built on the main one
            GBGPU_HIP(ctx, hipEventRecord(ctx->sync_ev[0], s));
            GBGPU_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->sync_ev[0], 0));
            launch_group_inverse(m, tinv, ctx->side);
            GBGPU_HIP(ctx, hipGetLastError());
            GBGPU_HIP(ctx, hipEventRecord(ctx->lag_ev[0], ctx->side));
        } else {
            launch_group_inverse(m, tinv, s);
            GBGPU_HIP(ctx, hipGetLastError());
        }
    }
    if (blocked && !dense) {
        rc = prep_group_buckets(m, acc32 ? 2 : 3);
        if (rc) {
            (void)hipStreamSynchronize(ctx->side);  // the inverses still write tinv
            dfree(tinv);
            return rc;
        }
        if (ngroups > 0) GBGPU_HIP(ctx, hipStreamWaitEvent(s, ctx->lag_ev[0], 0));
    }
    // waves in flight: LDS of one wave is W*8 bytes + staging, 160 KB per CU
    const int W = 1 << m->wshift;
    const int lds = blocked ? (int)(acc32 ? blocked_lds_bytes<uint32_t>(m->wshift) : blocked_lds_bytes<uint64_t>(m->wshift))
                            : W * 8 + STAGE * 8 + 2048;
    const int per_cu = (160 * 1024) / lds;
    int grid = ctx->num_cu * (per_cu > 0 ? per_cu : 1);
    if (grid > m->nrl) grid = m->nrl > 0 ? m->nrl : 1;
    uint32_t *xs = nullptr;
    int32_t *queue = nullptr;
    // per-wave multipliers of the row-parallel kernels (the dense route has none)
    // (with 32-bit accumulators followed by as many Shoup quotients)
    const size_t nxs = dense ? 1 : (size_t)grid * (size_t)(m->ncl > 0 ? m->ncl : 1) * (acc32 ? 2 : 1);
    if ((rc = dalloc(ctx, &xs, nxs))) { dfree(tinv); return rc; }
    if ((rc = dalloc(ctx, &queue, 6))) { dfree(xs); dfree(tinv); return rc; }
    GBGPU_HIP(ctx, hipMemsetAsync(queue, 0, 6 * sizeof(int32_t), s));
    SpArgs a;
    a.nc = m->nc; a.ncl = m->ncl; a.ncr = ncr; a.nrl = m->nrl; a.nwin = m->nwin;
    a.rp = m->rp; a.ci = m->ci; a.rcf = m->rcf; a.cp = m->cp; a.cf = m->cf;
    a.upiv = m->upiv; a.pbits = m->pbits; a.seg = m->seg; a.bucket = m->bucket;
    a.xs = xs; a.dprime = m->dprime; a.dnz = m->dnz; a.queue = queue;
    a.xs2 = acc32 ? xs + (size_t)grid * (size_t)(m->ncl > 0 ? m->ncl : 1) : nullptr;
    a.work = (unsigned long long *)(queue + 2);  // 8-byte aligned inside the 24-byte block
    a.md = Mod::make(m->p);
    const bool big = m->p >= (1u << 17);
    GBGPU_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    if (m->nrl > 0 && dense) {
        if ((rc = run_dense_route(m, tinv))) { dfree(xs); dfree(queue); dfree(tinv); return rc; }
    } else if (m->nrl > 0 && blocked) {
        switch (m->wshift) {
            case 10: launch_blocked<10>(a, tinv, m->gseg, m->gbucket, ngroups, big, acc32, grid, s); break;
            case 11: launch_blocked<11>(a, tinv, m->gseg, m->gbucket, ngroups, big, acc32, grid, s); break;
            case 12: launch_blocked<12>(a, tinv, m->gseg, m->gbucket, ngroups, big, acc32, grid, s); break;
            default: launch_blocked<13>(a, tinv, m->gseg, m->gbucket, ngroups, big, acc32, grid, s); break;
        }
    } else if (m->nrl > 0) {
        switch (m->wshift) {
            case 10: launch_sparse<10>(a, big, work != nullptr, grid, s); break;
            case 11: launch_sparse<11>(a, big, work != nullptr, grid, s); break;
            case 12: launch_sparse<12>(a, big, work != nullptr, grid, s); break;
            default: launch_sparse<13>(a, big, work != nullptr, grid, s); break;
        }
    }
    GBGPU_HIP(ctx, hipGetLastError());
    GBGPU_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    // compact the nonzero rows, keeping the order (la_ff_32.c:1166-1171)
    int32_t *rank = nullptr;
    if ((rc = dalloc(ctx, &rank, (size_t)m->nrl))) { dfree(xs); dfree(queue); return rc; }
    // (the list sized for every row to reduce, so that the count is read in
    // the call's one final wait)
    if ((rc = exclusive_scan_i32(ctx, m->dnz, rank, m->nrl, nullptr)) ||
        (rc = dalloc(ctx, &m->dlist, (size_t)m->nrl))) {
        dfree(xs); dfree(queue); dfree(rank); return rc;
    }
    if (m->nrl > 0) {
        k_compact<<<grid_for(m->nrl, 256), 256, 0, s>>>(m->nrl, m->dnz, rank, m->dlist);
        k_scan_total<int32_t><<<1, 1, 0, s>>>(m->dnz, rank, m->nrl, nullptr, ctx->dscan);
    }
    GBGPU_HIP(ctx, hipGetLastError());
    GBGPU_HIP(ctx, hipStreamSynchronize(s));
    flush_later(ctx);
    m->nd = m->nrl > 0 ? (int32_t)ctx->hscan[0] : 0;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->phase_ms[0] = ms;
    if (work) {
        GBGPU_HIP(ctx, hipMemcpy(work, a.work, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    dfree(xs);
    dfree(queue);
    dfree(rank);
    dfree(tinv);
    return GBGPU_OK;
}

// ----------------------------------------------------------- scans

template <typename I>
static int exclusive_scan(gbgpu_ctx *ctx, const I *in, I *out, int64_t n, int64_t *total,
                          const int32_t *dflag, int32_t *hflag) {
    hipStream_t s = ctx->stream;
    if (total) *total = 0;
    if (hflag) *hflag = 0;
    if (n > 0) {
        size_t tmp = 0;
        GBGPU_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, s));
        void *d_tmp = nullptr;
        int rc;
        if ((rc = dalloc(ctx, (char **)&d_tmp, tmp))) return rc;
        GBGPU_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp, in, out, (int)n, s));
        dfree_later(ctx, d_tmp);
    }
    if (!total) return GBGPU_OK;
    if (n > 0 || dflag) {
        k_scan_total<I><<<1, 1, 0, s>>>(n > 0 ? in : nullptr, n > 0 ? out : nullptr, n, dflag, ctx->dscan);
        GBGPU_HIP(ctx, hipGetLastError());
        GBGPU_HIP(ctx, hipStreamSynchronize(s));
        flush_later(ctx);
        if (n > 0) *total = ctx->hscan[0];
        if (hflag) *hflag = (int32_t)ctx->hscan[1];
    }
    return GBGPU_OK;
}

int exclusive_scan_i64(gbgpu_ctx *ctx, const int64_t *in, int64_t *out, int64_t n,
                       int64_t *total, const int32_t *dflag, int32_t *hflag) {
    return exclusive_scan<int64_t>(ctx, in, out, n, total, dflag, hflag);
}

int exclusive_scan_i32(gbgpu_ctx *ctx, const int32_t *in, int32_t *out, int64_t n,
                       int64_t *total, const int32_t *dflag, int32_t *hflag) {
    return exclusive_scan<int32_t>(ctx, in, out, n, total, dflag, hflag);
}

}  // namespace gbgpu
