// rows.hip -- dense <-> sparse row conversion on the device.
//
// dense_to_rows replaces convert_to_sparse_matrix_rows_ff_{8,16,32}
// (la_ff_32.c:1666-1736): every pivot row of a reduced echelon form becomes
// a sparse row (strictly ascending columns, shifted by ncl), in decreasing
// lead column order.  The reference mallocs and reallocs one array per
// row; here one wave per row counts, a scan places the rows, and one wave
// per row compacts with ballots into a single CSR.
#include "internal.h"

namespace gbgpu {

__global__ void k_row_count(const uint32_t *a, int64_t ld, int32_t n, int32_t rank,
                            const int32_t *orow, const int32_t *olead, int reverse,
                            int64_t *cnt) {
    const int lane = threadIdx.x & 63;
    const int32_t k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= rank) return;
    const int32_t src = reverse ? rank - 1 - k : k;
    const uint32_t *row = a + (int64_t)orow[src] * ld;
    int64_t c = 0;
    for (int32_t j = olead[src] + lane; j < n; j += 64) c += row[j] != 0;
    for (int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
    if (lane == 0) cnt[k] = c;
}

__global__ void k_row_write(const uint32_t *a, int64_t ld, int32_t n, int32_t rank,
                            const int32_t *orow, const int32_t *olead, int reverse,
                            int32_t col_shift, const int64_t *off, int32_t *oci, uint32_t *ocf) {
    const int lane = threadIdx.x & 63;
    const int32_t k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= rank) return;
    const int32_t src = reverse ? rank - 1 - k : k;
    const uint32_t *row = a + (int64_t)orow[src] * ld;
    int64_t w = off[k];
    for (int32_t j0 = olead[src]; j0 < n; j0 += 64) {
        const int32_t j = j0 + lane;
        const uint32_t v = j < n ? row[j] : 0u;
        const uint64_t mask = __ballot(v != 0);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (v != 0) {
            oci[w + before] = j + col_shift;
            ocf[w + before] = v;
        }
        w += __popcll(mask);
    }
}

int dense_to_rows(gbgpu_dense *d, int32_t col_shift, int reverse, int64_t **orp, int32_t **oci,
                  uint32_t **ocf, int64_t *onnz) {
    gbgpu_ctx *ctx = d->ctx;
    hipStream_t s = ctx->stream;
    int rc;
    *orp = nullptr; *oci = nullptr; *ocf = nullptr; *onnz = 0;
    const int32_t rank = d->rank;
    int64_t *cnt = nullptr;
    if ((rc = dalloc(ctx, &cnt, (size_t)rank + 1))) return rc;
    if ((rc = dalloc(ctx, orp, (size_t)rank + 1))) { dfree(cnt); return rc; }
    const int blocks = (rank + 3) / 4;
    if (rank > 0)
        k_row_count<<<blocks, 256, 0, s>>>(d->a, d->ld, d->n, rank, d->orow, d->olead, reverse, cnt);
    GBGPU_HIP(ctx, hipMemsetAsync(cnt + rank, 0, sizeof(int64_t), s));
    int64_t total = 0;
    // rp[k] = exclusive prefix, rp[rank] = total
    if ((rc = exclusive_scan_i64(ctx, cnt, *orp, (int64_t)rank + 1, &total))) { dfree(cnt); return rc; }
    *onnz = total;
    if ((rc = dalloc(ctx, oci, (size_t)total)) || (rc = dalloc(ctx, ocf, (size_t)total))) {
        dfree(cnt); return rc;
    }
    if (rank > 0)
        k_row_write<<<blocks, 256, 0, s>>>(d->a, d->ld, d->n, rank, d->orow, d->olead, reverse,
                                           col_shift, *orp, *oci, *ocf);
    GBGPU_HIP(ctx, hipGetLastError());
    GBGPU_HIP(ctx, hipStreamSynchronize(s));
    dfree(cnt);
    return GBGPU_OK;
}

__global__ void k_gather(const uint32_t *src, int64_t ld_src, const int32_t *list, int32_t n,
                         int32_t ncols, uint32_t *dst, int64_t ld) {
    for (int32_t r = blockIdx.y; r < n; r += gridDim.y) {
        const uint32_t *s = src + (int64_t)list[r] * ld_src;
        uint32_t *d = dst + (int64_t)r * ld;
        for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < ncols; j += gridDim.x * blockDim.x)
            d[j] = s[j];
    }
}

int gather_dense(gbgpu_ctx *ctx, const uint32_t *src, int64_t ld_src, const int32_t *list,
                 int32_t n, int32_t ncols, uint32_t p, gbgpu_dense **out) {
    *out = nullptr;
    gbgpu_dense *d = new gbgpu_dense();
    d->ctx = ctx;
    d->m = n;
    d->n = ncols;
    d->ld = ncols;
    d->p = p;
    int rc;
    if ((rc = dalloc(ctx, &d->a, (size_t)n * (size_t)(ncols > 0 ? ncols : 1)))) { delete d; return rc; }
    if (n > 0 && ncols > 0) {
        dim3 g((unsigned)((ncols + 255) / 256 < 64 ? (ncols + 255) / 256 : 64),
               (unsigned)(n < 65535 ? n : 65535));
        k_gather<<<g, 256, 0, ctx->stream>>>(src, ld_src, list, n, ncols, d->a, d->ld);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { delete d; return check_hip(ctx, e, "gather"); }
    *out = d;
    return GBGPU_OK;
}

// rows of a CSR matrix scattered into a dense block (for interreduction)
__global__ void k_scatter_rows(int32_t nr, const int64_t *rp, const int32_t *ci, const int32_t *rcf,
                               const int64_t *cp, const uint32_t *cf, const int32_t *keep,
                               const int32_t *dst_row, uint32_t *a, int64_t ld) {
    const int lane = threadIdx.x & 63;
    const int32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= nr || !keep[i]) return;
    const int64_t b = rp[i], e = rp[i + 1];
    const uint32_t *c = cf + cp[rcf[i]] - b;
    uint32_t *row = a + (int64_t)dst_row[i] * ld;
    for (int64_t k = b + lane; k < e; k += 64) row[ci[k]] = c[k];
}

int scatter_rows(gbgpu_mat *m, const int32_t *keep, const int32_t *dst_row, int32_t nkeep,
                 gbgpu_dense **out) {
    gbgpu_ctx *ctx = m->ctx;
    *out = nullptr;
    gbgpu_dense *d = new gbgpu_dense();
    d->ctx = ctx;
    d->m = nkeep;
    d->n = m->nc;
    d->ld = m->nc;
    d->p = m->p;
    int rc;
    if ((rc = dalloc(ctx, &d->a, (size_t)nkeep * (size_t)(m->nc > 0 ? m->nc : 1)))) { delete d; return rc; }
    if (hipMemsetAsync(d->a, 0, (size_t)nkeep * (size_t)m->nc * sizeof(uint32_t), ctx->stream) != hipSuccess) {
        delete d;
        return check_hip(ctx, hipGetLastError(), "memset");
    }
    if (m->nr > 0)
        k_scatter_rows<<<(m->nr + 3) / 4, 256, 0, ctx->stream>>>(m->nr, m->rp, m->ci, m->rcf, m->cp,
                                                                m->cf, keep, dst_row, d->a, d->ld);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { delete d; return check_hip(ctx, e, "scatter rows"); }
    *out = d;
    return GBGPU_OK;
}

}  // namespace gbgpu
