// common.h -- shared device helpers and internal host structures of libgbgpu.
//
// Arithmetic conventions (all exact, all integer):
//   * residues are uint32 in [0, p), 2 <= p < 2^31;
//   * accumulators are uint64 and only ever receive non-negative addends,
//     so they can be summed with LDS atomics and reduced lazily;
//   * for p < 2^17 an addend is the raw product (p - v) * c < 2^34 (the
//     delayed reduction of la_ff_32.c:528-536, good for 2^29 additions);
//   * for larger p the product is reduced first, (p - v) * c mod p < 2^31,
//     which keeps 2^33 additions exact (la_ff_32.c:593-606 reaches the same
//     residues by subtracting and lifting by p^2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gbgpu.h"

namespace gbgpu {

// Barrett reduction of a 64-bit value by p < 2^31 with m = floor((2^64-1)/p):
// q = floor(x m / 2^64) is floor(x/p) or one less, so one correction step.
struct Mod {
    uint32_t p;
    uint64_t m;

    __host__ static Mod make(uint32_t p_) {
        Mod r;
        r.p = p_;
        r.m = ~0ull / p_;
        return r;
    }
    __device__ __forceinline__ uint32_t reduce(uint64_t x) const {
        const uint64_t q = __umul64hi(x, m);
        uint64_t r = x - q * (uint64_t)p;
        return (uint32_t)(r >= p ? r - p : r);
    }
    __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const {
        return reduce((uint64_t)a * (uint64_t)b);
    }
};

// addend for an accumulator: see the header comment
template <bool BIG>
__device__ __forceinline__ uint64_t addend(uint32_t mul, uint32_t c, const Mod &md) {
    if (BIG) return (uint64_t)md.mul(mul, c);
    return (uint64_t)mul * (uint64_t)c;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

}  // namespace gbgpu

// ------------------------------------------------------------- host side

struct gbgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cu = 256;
    std::string err;
    double phase_ms[3] = {-1, -1, -1};
    hipEvent_t ev[6] = {};          // timing events of the phases
    hipStream_t side = nullptr;     // second stream of the dense look-ahead
    hipEvent_t sync_ev[1] = {};     // orders the main stream's work before the side's
    // the side stream's work before the main stream's next step, recorded
    // in turn; the dense panel loop also keeps at most as many panels queued
    // ahead of the device as there are lag events and learns that the
    // echelon form is complete from a flag in mapped host memory, which the
    // device sets: no stream is drained while panels remain
    hipEvent_t lag_ev[4] = {};
    volatile int32_t *hdone = nullptr;  // host view of the flag
    int32_t *ddone = nullptr;           // device view of the same word
    // a scan's total, written by the device into the same mapped block
    volatile int64_t *hscan = nullptr;
    int64_t *dscan = nullptr;
    // the dense panel loop's MFMA form orders its two streams through device
    // words instead of events (dsync, 512 bytes, zero at start; values of a
    // sequence that grows by one per panel, sync_seq the next free one); the
    // side stream's last panel done and a wait that ran out of its bound
    // reach the host through the same mapped block
    int32_t *dsync = nullptr;
    int32_t sync_seq = 1;
    volatile int32_t *hprog = nullptr, *herr = nullptr;
    int32_t *dprog = nullptr, *derr = nullptr;
    // buffers freed while queued work of this context may still read them:
    // they go back to the pool at the next wait for the stream
    std::vector<void *> later;
};

namespace gbgpu {

int set_error(gbgpu_ctx *ctx, int code, const std::string &msg);
int check_hip(gbgpu_ctx *ctx, hipError_t e, const char *what);

#define GBGPU_HIP(ctx, call)                                   \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return gbgpu::check_hip((ctx), e_, #call); \
    } while (0)

// Device memory comes from a caching pool (api.hip): buffers freed by one
// call are reused by the next one instead of going back to hipFree, so a
// step of the hot path does not pay for hipMalloc.  Every entry point
// synchronises its stream before it returns, so a freed buffer is idle.
void *pool_alloc(size_t bytes);
void pool_free(void *p);

template <class T>
int dalloc(gbgpu_ctx *ctx, T **p, size_t n) {
    if (n == 0) n = 1;
    *p = (T *)pool_alloc(n * sizeof(T));
    if (!*p)
        return set_error(ctx, GBGPU_ENOMEM,
                         "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    return GBGPU_OK;
}

template <class T>
void dfree(T *&p) {
    if (p) pool_free((void *)p);
    p = nullptr;
}

// free p once the context's stream has passed its last use: a buffer back
// in the pool at once could go to another context's thread while this
// stream's queued work still reads it
template <class T>
void dfree_later(gbgpu_ctx *ctx, T *&p) {
    if (p) ctx->later.push_back((void *)p);
    p = nullptr;
}

// after a wait for the stream: the deferred buffers are idle
inline void flush_later(gbgpu_ctx *ctx) {
    for (void *q : ctx->later) pool_free(q);
    ctx->later.clear();
}

// in[n - 1] + out[n - 1] (the sum of an exclusive scan's input) and a flag
// word into the context's mapped host block
template <typename I>
__global__ void k_scan_total(const I *in, const I *out, int64_t n, const int32_t *flag,
                             int64_t *dst) {
    volatile int64_t *d = dst;
    d[0] = n > 0 ? (int64_t)in[n - 1] + (int64_t)out[n - 1] : 0;
    d[1] = flag ? *flag : 0;
}

// exclusive prefix sum of n values into out (not aliasing in); with total
// given, waits for the stream once and returns the sum of the values in
// *total, and a device word dflag (if given) in *hflag; with total null it
// waits for nothing.  Scratch from hipcub
int exclusive_scan_i64(gbgpu_ctx *ctx, const int64_t *in, int64_t *out, int64_t n,
                       int64_t *total, const int32_t *dflag = nullptr, int32_t *hflag = nullptr);
int exclusive_scan_i32(gbgpu_ctx *ctx, const int32_t *in, int32_t *out, int64_t n,
                       int64_t *total, const int32_t *dflag = nullptr, int32_t *hflag = nullptr);

}  // namespace gbgpu
