// noalloc.hip -- the dense echelon form for callers that cannot allocate.
//
// gbgpu_dense_rref_noalloc runs the same kernels as gbgpu_dense_rref
// (dense.hip) on a block the caller already holds in device memory, with
// every scratch array carved from one file-scope device array instead of
// the caching pool.  Nothing here calls hipMalloc, hipFree or a library
// routine that allocates.  The array lives on each device the library runs
// on until the process ends; its address on a device comes from
// hipGetSymbolAddress, looked up once per device on the first call there
// (the array's name in host code is not a device address).  Every call on a
// device shares that device's array and finds the previous call's contents
// in it: the kernels initialise what they read, so nothing is cleared.
// Calls that overlap on one device race on the array.
#include <mutex>

#include "internal.h"

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)GBGPU_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t *g_base[MAX_DEVICES] = {};

// the device address of `scratch` on ctx's device (looked up once)
int scratch_on(gbgpu_ctx *ctx, uint8_t **out) {
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
        return gbgpu::set_error(ctx, GBGPU_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_base[ctx->device]) {
        GBGPU_HIP(ctx, hipSetDevice(ctx->device));
        void *ptr = nullptr;
        GBGPU_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return gbgpu::set_error(ctx, GBGPU_EDEVICE, "no address for the scratch array");
        g_base[ctx->device] = (uint8_t *)ptr;
    }
    *out = g_base[ctx->device];
    return GBGPU_OK;
}

// The output list in one workgroup: the k-th pivot column counted from the
// right (decreasing lead, as k_out_flags / k_out_list order it) goes to slot
// k.  1024 columns per step; a step's offsets come from the waves' ballots.
constexpr int OUT_T = 1024;
__global__ void __launch_bounds__(OUT_T)
k_out_one_block(int32_t n, const int32_t *colpiv, int32_t *orow, int32_t *olead,
                int32_t *count) {
    __shared__ int32_t wsum[OUT_T / 64];
    __shared__ int32_t base;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) base = 0;
    __syncthreads();
    for (int32_t c0 = 0; c0 < n; c0 += OUT_T) {
        const int32_t c = c0 + t;
        const int32_t col = n - 1 - c;
        const int32_t row = c < n ? colpiv[col] : -1;
        const uint64_t bal = __ballot(row >= 0);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int v = 0; v < w; ++v) off += wsum[v];
        if (row >= 0) {
            orow[off + before] = row;
            olead[off + before] = col;
        }
        __syncthreads();
        if (t == 0) {
            int tot = 0;
            for (int v = 0; v < OUT_T / 64; ++v) tot += wsum[v];
            base += tot;
        }
        __syncthreads();
    }
    if (t == 0) *count = base;
}

// rows[k] = the k-th output row, zero left of its lead
__global__ void k_rows_out(const uint32_t *a, int32_t n, const int32_t *orow, const int32_t *olead,
                           const int32_t *count, uint32_t *rows) {
    const int32_t rank = *count;
    for (int32_t k = blockIdx.y; k < rank; k += gridDim.y) {
        const uint32_t *src = a + (int64_t)orow[k] * n;
        const int32_t l = olead[k];
        for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
            rows[(int64_t)k * n + j] = j < l ? 0u : src[j];
    }
}

}  // namespace

using namespace gbgpu;

extern "C" {

size_t gbgpu_dense_rref_noalloc_bytes(int32_t m, int32_t n, uint32_t p, int form) {
    if (m < 0 || n < 0) return 0;
    // scratch of the panel loop, then orow, olead (n each) and the count
    return dense_ws_bytes(m, n, form, p) + 2 * (((size_t)n * 4 + 255) & ~(size_t)255) + 256;
}

int gbgpu_dense_rref_noalloc(gbgpu_ctx *ctx, uint32_t *a, int32_t m, int32_t n, uint32_t p,
                             int form, uint32_t *rows, int32_t *lead, int32_t *rank) {
    if (!ctx || !rank || m < 0 || n < 0) return GBGPU_EINVAL;
    if (form != GBGPU_FORM_PLAIN && form != GBGPU_FORM_MFMA) return set_error(ctx, GBGPU_EINVAL, "bad form");
    if (p < 2 || p >= (1u << 31)) return set_error(ctx, GBGPU_EINVAL, "p must be in [2, 2^31)");
    if ((int64_t)m * n > 0 && (!a || !rows || !lead)) return set_error(ctx, GBGPU_EINVAL, "NULL buffer");
    *rank = 0;
    if ((int64_t)m * n == 0) return GBGPU_OK;
    const size_t need = gbgpu_dense_rref_noalloc_bytes(m, n, p, form);
    if (need > SCRATCH_BYTES)
        return set_error(ctx, GBGPU_ERANGE,
                         "block needs " + std::to_string(need) + " bytes of scratch, more than the " +
                             std::to_string(SCRATCH_BYTES) + " of gbgpu_dense_rref_noalloc");
    uint8_t *base = nullptr;
    int rc;
    if ((rc = scratch_on(ctx, &base))) return rc;
    GBGPU_HIP(ctx, hipSetDevice(ctx->device));
    DenseWs ws;
    dense_ws_carve(base, m, n, form, p, &ws);
    const size_t wsb = dense_ws_bytes(m, n, form, p);
    const size_t ob = ((size_t)n * 4 + 255) & ~(size_t)255;
    int32_t *orow = (int32_t *)(base + wsb);
    int32_t *olead = (int32_t *)(base + wsb + ob);
    int32_t *count = (int32_t *)(base + wsb + 2 * ob);
    // the block is the caller's: a view with no memory of its own
    gbgpu_dense d;
    d.ctx = ctx;
    d.m = m;
    d.n = n;
    d.ld = n;
    d.p = p;
    d.a = a;
    d.own_a = false;
    rc = dense_rref_ws(&d, form, ws);
    d.a = nullptr;
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    k_out_one_block<<<1, OUT_T, 0, s>>>(n, ws.colpiv, orow, olead, count);
    const int32_t mn = m < n ? m : n;
    dim3 g((unsigned)((n + 255) / 256 < 64 ? (n + 255) / 256 : 64), (unsigned)(mn < 65535 ? mn : 65535));
    k_rows_out<<<g, 256, 0, s>>>(a, n, orow, olead, count, rows);
    GBGPU_HIP(ctx, hipGetLastError());
    int32_t hr = 0;
    GBGPU_HIP(ctx, hipMemcpyAsync(lead, olead, (size_t)mn * 4, hipMemcpyDeviceToDevice, s));
    GBGPU_HIP(ctx, hipMemcpyAsync(&hr, count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GBGPU_HIP(ctx, hipStreamSynchronize(s));
    if ((rc = loop_sync_check(ctx))) return rc;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) ctx->phase_ms[1] = ms;
    *rank = hr;
    return GBGPU_OK;
}

}  // extern "C"
