// api.hip -- the flat C-ABI of libgbgpu (include/gbgpu.h).
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "internal.h"

namespace gbgpu {

int set_error(gbgpu_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int check_hip(gbgpu_ctx *ctx, hipError_t e, const char *what) {
    return set_error(ctx, GBGPU_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------ caching pool
//
// Size classes: powers of two up to 2 MiB, multiples of 2 MiB above.  A
// cached block is reused for a request of its class up to twice its size.
// Blocks are tied to the device they were allocated on.
namespace {
struct Pool {
    std::mutex mu;
    std::multimap<size_t, std::pair<int, void *>> idle;   // class -> (device, ptr)
    std::unordered_map<void *, std::pair<int, size_t>> live;
    size_t idle_bytes = 0;
};
Pool &pool() {
    static Pool *p = new Pool();  // never destroyed: freed blocks may outlive statics
    return *p;
}
size_t size_class(size_t b) {
    const size_t mb2 = (size_t)2 << 20;
    if (b >= mb2) return (b + mb2 - 1) / mb2 * mb2;
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
}
void release_idle(Pool &P) {
    for (auto &kv : P.idle) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(kv.second.first);
        (void)hipFree(kv.second.second);
        (void)hipSetDevice(cur);
    }
    P.idle.clear();
    P.idle_bytes = 0;
}
}  // namespace

void *pool_alloc(size_t bytes) {
    Pool &P = pool();
    const size_t cls = size_class(bytes);
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto it = P.idle.lower_bound(cls); it != P.idle.end() && it->first <= 2 * cls; ++it) {
        if (it->second.first != dev) continue;
        void *ptr = it->second.second;
        P.live[ptr] = {dev, it->first};
        P.idle_bytes -= it->first;
        P.idle.erase(it);
        return ptr;
    }
    void *ptr = nullptr;
    if (hipMalloc(&ptr, cls) != hipSuccess) {
        (void)hipGetLastError();
        release_idle(P);
        ptr = nullptr;
        if (hipMalloc(&ptr, cls) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    P.live[ptr] = {dev, cls};
    return ptr;
}

void pool_free(void *ptr) {
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(ptr);
    if (it == P.live.end()) return;
    P.idle.insert({it->second.second, {it->second.first, ptr}});
    P.idle_bytes += it->second.second;
    P.live.erase(it);
    // keep at most 64 GiB idle
    if (P.idle_bytes > ((size_t)64 << 30)) release_idle(P);
}

}  // namespace gbgpu

using namespace gbgpu;

gbgpu_mat::~gbgpu_mat() {
    dfree(rp); dfree(ci); dfree(rcf); dfree(cp);
    if (!cf_borrowed) dfree(cf);
    dfree(first); dfree(upiv);
    dfree(seg); dfree(bucket); dfree(pbits); dfree(gseg); dfree(gbucket);
    dfree(dprime); dfree(dnz); dfree(dlist);
    dfree(orp); dfree(oci); dfree(ocf);
}

gbgpu_dense::~gbgpu_dense() {
    if (own_a) dfree(a);
    dfree(orow); dfree(olead);
}

extern "C" {

int gbgpu_arith_bits(uint32_t p) {
    // io.c:815-829 (storage width) and io.c:710-729 (17/31-bit kernels)
    if (p < (1u << 7)) return 8;
    if (p < (1u << 15)) return 16;
    if (p < (1u << 17)) return 17;
    return 31;
}

int gbgpu_init(int device, gbgpu_ctx **out) {
    if (!out) return GBGPU_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return GBGPU_EDEVICE;
    }
    if (device < 0 || device >= n) return GBGPU_EINVAL;
    gbgpu_ctx *ctx = new gbgpu_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return GBGPU_EDEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return GBGPU_EDEVICE;
    }
    ctx->own_stream = true;
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
        gbgpu_finalize(ctx);
        return GBGPU_EDEVICE;
    }
    for (auto &e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            gbgpu_finalize(ctx);
            return GBGPU_EDEVICE;
        }
    }
    // ordering between the context's two streams on one device needs no
    // system-scope fence; the lag events order nothing (the host reads only
    // the mapped flag after them, which the device writes uncached)
    for (auto &e : ctx->sync_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
            gbgpu_finalize(ctx);
            return GBGPU_EDEVICE;
        }
    }
    for (auto &e : ctx->lag_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
            gbgpu_finalize(ctx);
            return GBGPU_EDEVICE;
        }
    }
    void *hflag = nullptr;
    if (hipHostMalloc(&hflag, 256, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        (void)hipGetLastError();
        gbgpu_finalize(ctx);
        return GBGPU_EDEVICE;
    }
    ctx->hdone = (volatile int32_t *)hflag;
    *ctx->hdone = 0;
    ctx->hscan = (volatile int64_t *)((char *)hflag + 64);
    *ctx->hscan = 0;
    ctx->hprog = (volatile int32_t *)((char *)hflag + 128);
    *ctx->hprog = 0;
    ctx->herr = (volatile int32_t *)((char *)hflag + 192);
    *ctx->herr = 0;
    void *dflag = nullptr;
    if (hipHostGetDevicePointer(&dflag, hflag, 0) != hipSuccess) {
        (void)hipGetLastError();
        gbgpu_finalize(ctx);
        return GBGPU_EDEVICE;
    }
    ctx->ddone = (int32_t *)dflag;
    ctx->dscan = (int64_t *)((char *)dflag + 64);
    ctx->dprog = (int32_t *)((char *)dflag + 128);
    ctx->derr = (int32_t *)((char *)dflag + 192);
    // zero before any stream of the context runs (hipMemset may return
    // before it has run, and the streams do not wait for the null stream)
    if (hipMalloc(&ctx->dsync, 512) != hipSuccess ||
        hipMemsetAsync(ctx->dsync, 0, 512, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        gbgpu_finalize(ctx);
        return GBGPU_ENOMEM;
    }
    *out = ctx;
    return GBGPU_OK;
}

void gbgpu_finalize(gbgpu_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    flush_later(ctx);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->sync_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->lag_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->hdone) (void)hipHostFree((void *)ctx->hdone);
    if (ctx->dsync) (void)hipFree(ctx->dsync);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int gbgpu_set_stream(gbgpu_ctx *ctx, void *stream) {
    if (!ctx) return GBGPU_EINVAL;
    if (ctx->own_stream && ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    ctx->stream = (hipStream_t)stream;
    ctx->own_stream = false;
    return GBGPU_OK;
}

const char *gbgpu_last_error(const gbgpu_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "no context";
}

double gbgpu_phase_ms(const gbgpu_ctx *ctx, int phase) {
    if (!ctx || phase < 0 || phase > 2) return -1;
    return ctx->phase_ms[phase];
}

static int check_csr_header(gbgpu_ctx *ctx, const gbgpu_csr_t *m, bool need_ci = true,
                            bool need_cf = true) {
    if (!m) return set_error(ctx, GBGPU_EINVAL, "NULL matrix");
    if (m->nr < 0 || m->nc < 0 || m->ncl < 0 || m->ncl > m->nc || m->ncf < 0)
        return set_error(ctx, GBGPU_EINVAL, "bad matrix dimensions");
    if (m->p < 2 || m->p >= (1u << 31)) return set_error(ctx, GBGPU_EINVAL, "p must be in [2, 2^31)");
    if ((m->nr > 0 && (!m->rp || (need_ci && !m->ci) || !m->rcf)) ||
        (m->ncf > 0 && (!m->cp || (need_cf && !m->cf))))
        return set_error(ctx, GBGPU_EINVAL, "NULL array in matrix");
    return GBGPU_OK;
}

// host-side validation of the row structure the kernels index with (the
// entries themselves are checked on the device right after the copy)
static int check_csr_body(gbgpu_ctx *ctx, const gbgpu_csr_t *m) {
    if (m->nr > 0 && m->rp[0] != 0) return set_error(ctx, GBGPU_EINVAL, "rp[0] must be 0");
    for (int32_t i = 0; i < m->nr; ++i) {
        if (m->rp[i + 1] <= m->rp[i]) return set_error(ctx, GBGPU_ECONTRACT, "empty row in matrix");
        const int32_t id = m->rcf[i];
        if (id < 0 || id >= m->ncf) return set_error(ctx, GBGPU_EINVAL, "coefficient array id out of range");
        if (m->cp[id + 1] - m->cp[id] < m->rp[i + 1] - m->rp[i])
            return set_error(ctx, GBGPU_EINVAL, "coefficient array shorter than its row");
    }
    for (int32_t i = 0; i < m->ncf; ++i)
        if (m->cp[i + 1] < m->cp[i]) return set_error(ctx, GBGPU_EINVAL, "cp must not decrease");
    return GBGPU_OK;
}

// flag |= 1 for a column index outside [0, nc), |= 2 for a coefficient >= p
__global__ void k_check_entries(int64_t nnz, const int32_t *ci, int32_t nc, int64_t npool,
                                const uint32_t *cf, uint32_t p, int32_t *flag) {
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    int bad = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += step)
        if (ci[k] < 0 || ci[k] >= nc) bad |= 1;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npool; k += step)
        if (cf[k] >= p) bad |= 2;
    if (bad) atomicOr(flag, bad);
}

// row q of a matrix given by monomial ids: mon[src[q] .. src[q] + len(q))
// mapped through colmap to the column indices ci[rp[q] ..]; flag |= 1 for an
// id outside the map.  One workgroup per row at a time, coalesced in t.
__global__ void k_map_rows(int32_t nr, const int64_t *rp, const int64_t *src, const int32_t *mon,
                           const int32_t *colmap, int32_t nmap, int32_t *ci, int32_t *flag) {
    int bad = 0;
    for (int32_t q = blockIdx.x; q < nr; q += gridDim.x) {
        const int64_t b = rp[q], len = rp[q + 1] - b;
        const int32_t *in = mon + src[q];
        for (int64_t t = threadIdx.x; t < len; t += blockDim.x) {
            const int32_t id = in[t];
            if ((uint32_t)id >= (uint32_t)nmap) {
                bad = 1;
                ci[b + t] = -1;
            } else {
                ci[b + t] = colmap[id];
            }
        }
    }
    if (bad) atomicOr(flag, 1);
}

// the monomial-id form of a matrix's column indices (gbgpu_mat_upload_rows)
struct RowsById {
    const int32_t *mon;
    int64_t nmon;
    const int64_t *src;
    const int32_t *colmap;
    int32_t nmap;
    const gbgpu_cfstore *store;  // coefficient arrays already on the device, or null
};

// Coefficient arrays kept on the device (entries checked < p on append).
struct gbgpu_cfstore {
    gbgpu_ctx *ctx = nullptr;
    uint32_t p = 0;
    uint32_t *d = nullptr;
    int64_t n = 0, cap = 0;
    ~gbgpu_cfstore() { dfree(d); }
};

int gbgpu_cfstore_create(gbgpu_ctx *ctx, uint32_t p, gbgpu_cfstore **out) {
    if (!ctx || !out) return GBGPU_EINVAL;
    *out = nullptr;
    if (p < 2 || p >= (1u << 31)) return set_error(ctx, GBGPU_EINVAL, "p must be in [2, 2^31)");
    gbgpu_cfstore *c = new gbgpu_cfstore();
    c->ctx = ctx;
    c->p = p;
    *out = c;
    return GBGPU_OK;
}

int gbgpu_cfstore_append(gbgpu_cfstore *c, const uint32_t *cf, int64_t n, int64_t *off) {
    if (!c || n < 0 || (n > 0 && !cf)) return GBGPU_EINVAL;
    gbgpu_ctx *ctx = c->ctx;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] >= c->p) return set_error(ctx, GBGPU_EINVAL, "coefficient not reduced mod p");
    GBGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (c->n + n > c->cap) {
        int64_t nc = c->cap ? c->cap : ((int64_t)1 << 20);
        while (nc < c->n + n) nc += nc / 2;
        uint32_t *nd = nullptr;
        int rc = dalloc(ctx, &nd, (size_t)nc);
        if (rc) return rc;
        hipError_t e = c->n ? hipMemcpyAsync(nd, c->d, (size_t)c->n * 4, hipMemcpyDeviceToDevice, s)
                            : hipSuccess;
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            dfree(nd);
            return check_hip(ctx, e, "coefficient store");
        }
        dfree(c->d);
        c->d = nd;
        c->cap = nc;
    }
    if (n > 0) {
        hipError_t e = hipMemcpyAsync(c->d + c->n, cf, (size_t)n * 4, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return check_hip(ctx, e, "coefficient store");
    }
    if (off) *off = c->n;
    c->n += n;
    return GBGPU_OK;
}

int64_t gbgpu_cfstore_size(const gbgpu_cfstore *c) { return c ? c->n : -1; }

void gbgpu_cfstore_free(gbgpu_cfstore *c) { delete c; }

static int mat_upload(gbgpu_ctx *ctx, const gbgpu_csr_t *h, const RowsById *byid, gbgpu_mat **out);

int gbgpu_mat_upload(gbgpu_ctx *ctx, const gbgpu_csr_t *h, gbgpu_mat **out) {
    return mat_upload(ctx, h, nullptr, out);
}

int gbgpu_mat_upload_rows(gbgpu_ctx *ctx, const gbgpu_csr_t *h, const int32_t *mon, int64_t nmon,
                          const int64_t *src, const int32_t *colmap, int32_t nmap,
                          const gbgpu_cfstore *store, gbgpu_mat **out) {
    if (!ctx || !out || !h) return GBGPU_EINVAL;
    if (nmon < 0 || nmap < 0 || (h->nr > 0 && (!mon || !src || !colmap)))
        return set_error(ctx, GBGPU_EINVAL, "NULL or negative monomial-id arrays");
    if (store && (store->ctx != ctx || store->p != h->p))
        return set_error(ctx, GBGPU_EINVAL, "coefficient store of another context or prime");
    const RowsById b = {mon, nmon, src, colmap, nmap, store};
    return mat_upload(ctx, h, &b, out);
}

static int mat_upload(gbgpu_ctx *ctx, const gbgpu_csr_t *h, const RowsById *byid, gbgpu_mat **out) {
    if (!ctx || !out) return GBGPU_EINVAL;
    *out = nullptr;
    int rc;
    const gbgpu_cfstore *store = byid ? byid->store : nullptr;
    if ((rc = check_csr_header(ctx, h, byid == nullptr, store == nullptr)) ||
        (rc = check_csr_body(ctx, h)))
        return rc;
    if (store && h->ncf > 0 && h->cp[h->ncf] > store->n)
        return set_error(ctx, GBGPU_EINVAL, "coefficient array outside the store");
    if (byid)
        for (int32_t q = 0; q < h->nr; ++q)
            if (byid->src[q] < 0 || byid->src[q] + (h->rp[q + 1] - h->rp[q]) > byid->nmon)
                return set_error(ctx, GBGPU_EINVAL, "row outside the monomial-id array");
    if ((int64_t)h->nr * 8 > INT32_MAX || (int64_t)h->nc * 64 > INT32_MAX)
        return set_error(ctx, GBGPU_ERANGE, "matrix too large for this build");
    GBGPU_HIP(ctx, hipSetDevice(ctx->device));
    gbgpu_mat *m = new gbgpu_mat();
    m->ctx = ctx;
    m->nr = h->nr; m->nc = h->nc; m->ncl = h->ncl; m->ncf = h->ncf; m->p = h->p;
    m->nnz = h->nr > 0 ? h->rp[h->nr] : 0;
    m->npool = h->ncf > 0 ? h->cp[h->ncf] : 0;
    hipStream_t s = ctx->stream;
    if ((rc = dalloc(ctx, &m->rp, (size_t)m->nr + 1)) || (rc = dalloc(ctx, &m->ci, (size_t)m->nnz)) ||
        (rc = dalloc(ctx, &m->rcf, (size_t)m->nr)) || (rc = dalloc(ctx, &m->cp, (size_t)m->ncf + 1)) ||
        (!store && (rc = dalloc(ctx, &m->cf, (size_t)m->npool)))) {
        delete m;
        return rc;
    }
    if (store) {
        // the store's arrays, borrowed: checked when they were appended
        m->cf = store->d;
        m->cf_borrowed = true;
    }
    // monomial ids, row sources and the column map: device copies for the
    // mapping kernel, freed after the synchronisation below
    int32_t *dmon = nullptr, *dmap = nullptr;
    int64_t *dsrc = nullptr;
    if (byid && m->nr > 0 &&
        ((rc = dalloc(ctx, &dmon, (size_t)byid->nmon)) || (rc = dalloc(ctx, &dsrc, (size_t)m->nr)) ||
         (rc = dalloc(ctx, &dmap, (size_t)byid->nmap)))) {
        dfree(dmon); dfree(dsrc); dfree(dmap);
        delete m;
        return rc;
    }
    hipError_t e = hipSuccess;
    if (m->nr > 0) {
        e = hipMemcpyAsync(m->rp, h->rp, ((size_t)m->nr + 1) * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && !byid)
            e = hipMemcpyAsync(m->ci, h->ci, (size_t)m->nnz * 4, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && byid) {
            e = hipMemcpyAsync(dmon, byid->mon, (size_t)byid->nmon * 4, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipMemcpyAsync(dsrc, byid->src, (size_t)m->nr * 8, hipMemcpyHostToDevice, s);
            if (e == hipSuccess && byid->nmap > 0)
                e = hipMemcpyAsync(dmap, byid->colmap, (size_t)byid->nmap * 4, hipMemcpyHostToDevice, s);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(m->rcf, h->rcf, (size_t)m->nr * 4, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess && m->ncf > 0) {
        e = hipMemcpyAsync(m->cp, h->cp, ((size_t)m->ncf + 1) * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && !store)
            e = hipMemcpyAsync(m->cf, h->cf, (size_t)m->npool * 4, hipMemcpyHostToDevice, s);
    }
    // every column index and coefficient checked before any kernel uses them
    int32_t *flag = nullptr;
    int32_t hflag = 0;
    if (e == hipSuccess && (rc = dalloc(ctx, &flag, 1))) {
        dfree(dmon); dfree(dsrc); dfree(dmap);
        delete m;
        return rc;
    }
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(int32_t), s);
    if (e == hipSuccess && byid && m->nr > 0) {
        const int32_t grid = m->nr < 8192 ? m->nr : 8192;
        k_map_rows<<<grid, 256, 0, s>>>(m->nr, m->rp, dsrc, dmon, dmap, byid->nmap, m->ci, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess && (m->nnz > 0 || m->npool > 0)) {
        k_check_entries<<<1024, 256, 0, s>>>(m->nnz, m->ci, m->nc, store ? 0 : m->npool, m->cf,
                                             m->p, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&hflag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(flag);
    dfree(dmon); dfree(dsrc); dfree(dmap);
    if (e != hipSuccess) {
        delete m;
        return check_hip(ctx, e, "upload");
    }
    if (hflag) {
        delete m;
        return set_error(ctx, GBGPU_EINVAL, (hflag & 1) ? "column index out of range"
                                                        : "coefficient not reduced mod p");
    }
    *out = m;
    return GBGPU_OK;
}

void gbgpu_mat_free(gbgpu_mat *m) { delete m; }

int gbgpu_sparse_pass(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *nrows) {
    if (!ctx || !m || m->ctx != ctx) return GBGPU_EINVAL;
    if (form != GBGPU_FORM_PLAIN && form != GBGPU_FORM_MFMA) return set_error(ctx, GBGPU_EINVAL, "bad form");
    int rc = run_sparse_pass(m, nullptr, form);
    if (rc) return rc;
    if (nrows) *nrows = m->nd;
    return GBGPU_OK;
}

int gbgpu_sparse_pass_work(gbgpu_ctx *ctx, gbgpu_mat *m, int32_t *nrows, uint64_t *work) {
    if (!ctx || !m || m->ctx != ctx || !work) return GBGPU_EINVAL;
    int rc = run_sparse_pass(m, work);
    if (rc) return rc;
    if (nrows) *nrows = m->nd;
    return GBGPU_OK;
}

__global__ void k_src_rows(int32_t nd, const int32_t *dlist, const int32_t *upiv, int32_t *src) {
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nd; k += gridDim.x * blockDim.x)
        src[k] = upiv[dlist[k]];
}

int gbgpu_sparse_pass_download(gbgpu_ctx *ctx, gbgpu_mat *m, uint32_t *dense, int32_t *src_row) {
    if (!ctx || !m || m->ctx != ctx) return GBGPU_EINVAL;
    if (!m->dlist && m->nd > 0) return set_error(ctx, GBGPU_EINVAL, "no sparse pass result");
    hipStream_t s = ctx->stream;
    const int32_t ncr = m->nc - m->ncl;
    int rc;
    if (dense && m->nd > 0) {
        gbgpu_dense *d = nullptr;
        if ((rc = gather_dense(ctx, m->dprime, ncr, m->dlist, m->nd, ncr, m->p, &d))) return rc;
        hipError_t e = hipMemcpyAsync(dense, d->a, (size_t)m->nd * (size_t)ncr * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        delete d;
        if (e != hipSuccess) return check_hip(ctx, e, "download");
    }
    if (src_row && m->nd > 0) {
        int32_t *ds = nullptr;
        if ((rc = dalloc(ctx, &ds, (size_t)m->nd))) return rc;
        k_src_rows<<<(m->nd + 255) / 256, 256, 0, s>>>(m->nd, m->dlist, m->upiv, ds);
        hipError_t e = hipMemcpyAsync(src_row, ds, (size_t)m->nd * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        dfree(ds);
        if (e != hipSuccess) return check_hip(ctx, e, "download");
    }
    return GBGPU_OK;
}

static void clear_rows(gbgpu_mat *m) {
    dfree(m->orp); dfree(m->oci); dfree(m->ocf);
    m->onp = 0;
    m->onnz = 0;
}

static int run_la1(gbgpu_ctx *ctx, gbgpu_mat *m, int form) {
    if (form != GBGPU_FORM_PLAIN && form != GBGPU_FORM_MFMA) return set_error(ctx, GBGPU_EINVAL, "bad form");
    GBGPU_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    clear_rows(m);
    int rc = run_sparse_pass(m, nullptr, form);
    if (rc) return rc;
    const int32_t ncr = m->nc - m->ncl;
    if (m->nd > 0) {
        gbgpu_dense *d = nullptr;
        if ((rc = gather_dense(ctx, m->dprime, ncr, m->dlist, m->nd, ncr, m->p, &d))) return rc;
        rc = dense_rref(d, form);
        if (!rc) rc = dense_to_rows(d, m->ncl, 0, &m->orp, &m->oci, &m->ocf, &m->onnz);
        if (!rc) m->onp = d->rank;
        delete d;
        if (rc) return rc;
    } else {
        if ((rc = dalloc(ctx, &m->orp, 1))) return rc;
        GBGPU_HIP(ctx, hipMemsetAsync(m->orp, 0, sizeof(int64_t), ctx->stream));
    }
    GBGPU_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    GBGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]) == hipSuccess) ctx->phase_ms[2] = ms;
    return GBGPU_OK;
}

int gbgpu_la_sparse_dense(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *np) {
    if (!ctx || !m || m->ctx != ctx) return GBGPU_EINVAL;
    int rc = run_la1(ctx, m, form);
    if (rc) return rc;
    if (np) *np = m->onp;
    return GBGPU_OK;
}

int gbgpu_la_sparse(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *np) {
    // la_option 2 returns the same canonical rows as la_option 1 whenever
    // every row's lead is a known-pivot column, which is what F4 produces
    // (symbol.c:129, convert.c:77-79); the same kernels serve both.
    return gbgpu_la_sparse_dense(ctx, m, form, np);
}

// last row per lead survives, as pivs[lead] = row in la_ff_32.c:1968-1976
__global__ void k_last_of_lead(int32_t nr, const int64_t *rp, const int32_t *ci, int32_t *last) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x)
        atomicMax(&last[ci[rp[i]]], i);
}

__global__ void k_keep_last(int32_t nr, const int64_t *rp, const int32_t *ci, const int32_t *last,
                            int32_t *keep) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x)
        keep[i] = last[ci[rp[i]]] == i;
}

__global__ void k_fill(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int gbgpu_interreduce(gbgpu_ctx *ctx, gbgpu_mat *m, int form, int32_t *np) {
    if (!ctx || !m || m->ctx != ctx) return GBGPU_EINVAL;
    if (m->ncl != 0) return set_error(ctx, GBGPU_ECONTRACT, "interreduction needs ncl == 0");
    hipStream_t s = ctx->stream;
    clear_rows(m);
    int rc;
    int32_t *last = nullptr, *keep = nullptr, *dst = nullptr;
    auto cleanup = [&]() { dfree(last); dfree(keep); dfree(dst); };
    if ((rc = dalloc(ctx, &last, (size_t)m->nc)) || (rc = dalloc(ctx, &keep, (size_t)m->nr)) ||
        (rc = dalloc(ctx, &dst, (size_t)m->nr))) {
        cleanup();
        return rc;
    }
    const int g = m->nr > 0 ? (m->nr + 255) / 256 : 1;
    k_fill<<<(m->nc + 255) / 256 + 1, 256, 0, s>>>(last, m->nc, -1);
    if (m->nr > 0) {
        k_last_of_lead<<<g, 256, 0, s>>>(m->nr, m->rp, m->ci, last);
        k_keep_last<<<g, 256, 0, s>>>(m->nr, m->rp, m->ci, last, keep);
    }
    int64_t nkeep = 0;
    if ((rc = exclusive_scan_i32(ctx, keep, dst, m->nr, &nkeep))) { cleanup(); return rc; }
    gbgpu_dense *d = nullptr;
    if ((rc = scatter_rows(m, keep, dst, (int32_t)nkeep, &d))) { cleanup(); return rc; }
    rc = dense_rref(d, form);
    if (!rc) rc = dense_to_rows(d, 0, 1, &m->orp, &m->oci, &m->ocf, &m->onnz);
    if (!rc) m->onp = d->rank;
    delete d;
    cleanup();
    if (rc) return rc;
    if (np) *np = m->onp;
    return GBGPU_OK;
}

int gbgpu_rows_download(gbgpu_ctx *ctx, gbgpu_mat *m, gbgpu_rows_t *out) {
    if (!ctx || !m || !out || m->ctx != ctx) return GBGPU_EINVAL;
    if (!m->orp) return set_error(ctx, GBGPU_EINVAL, "no LA result on this matrix");
    memset(out, 0, sizeof(*out));
    const size_t nnz = (size_t)m->onnz;
    out->rp = (int64_t *)malloc(((size_t)m->onp + 1) * 8);
    out->ci = (int32_t *)malloc((nnz ? nnz : 1) * 4);
    out->cf = (uint32_t *)malloc((nnz ? nnz : 1) * 4);
    if (!out->rp || !out->ci || !out->cf) {
        gbgpu_rows_free(out);
        return set_error(ctx, GBGPU_ENOMEM, "host allocation failed");
    }
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemcpyAsync(out->rp, m->orp, ((size_t)m->onp + 1) * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(out->ci, m->oci, nnz * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(out->cf, m->ocf, nnz * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        gbgpu_rows_free(out);
        return check_hip(ctx, e, "download rows");
    }
    out->np = m->onp;
    return GBGPU_OK;
}

void gbgpu_rows_free(gbgpu_rows_t *r) {
    if (!r) return;
    free(r->rp); free(r->ci); free(r->cf);
    r->rp = nullptr; r->ci = nullptr; r->cf = nullptr; r->np = 0;
}

int gbgpu_la_ff(gbgpu_ctx *ctx, const gbgpu_csr_t *h, int la_option, int form, gbgpu_rows_t *out) {
    if (!ctx || !out) return GBGPU_EINVAL;
    gbgpu_mat *m = nullptr;
    int rc = gbgpu_mat_upload(ctx, h, &m);
    if (rc) return rc;
    // io.c:867-872 maps la_option <= 0 to 1; io.c:608-779 sends 1/42/43 to
    // the sparse-dense routine and 2/44/unknown to the sparse one
    if (la_option <= 0) la_option = 1;
    int32_t np = 0;
    if (la_option == 1 || la_option == 42 || la_option == 43) rc = gbgpu_la_sparse_dense(ctx, m, form, &np);
    else rc = gbgpu_la_sparse(ctx, m, form, &np);
    if (!rc) rc = gbgpu_rows_download(ctx, m, out);
    gbgpu_mat_free(m);
    return rc;
}

int gbgpu_dense_upload(gbgpu_ctx *ctx, int32_t m, int32_t n, uint32_t p, const uint32_t *host,
                       gbgpu_dense **out) {
    if (!ctx || !out || m < 0 || n < 0 || (!host && (int64_t)m * n > 0)) return GBGPU_EINVAL;
    if (p < 2 || p >= (1u << 31)) return set_error(ctx, GBGPU_EINVAL, "p must be in [2, 2^31)");
    *out = nullptr;
    for (int64_t i = 0; i < (int64_t)m * n; ++i)
        if (host[i] >= p) return set_error(ctx, GBGPU_EINVAL, "entry not reduced mod p");
    GBGPU_HIP(ctx, hipSetDevice(ctx->device));
    gbgpu_dense *d = new gbgpu_dense();
    d->ctx = ctx;
    d->m = m;
    d->n = n;
    d->ld = n;
    d->p = p;
    int rc;
    if ((rc = dalloc(ctx, &d->a, (size_t)m * (size_t)(n > 0 ? n : 1)))) { delete d; return rc; }
    hipError_t e = hipSuccess;
    if ((int64_t)m * n > 0) e = hipMemcpyAsync(d->a, host, (size_t)m * (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { delete d; return check_hip(ctx, e, "dense upload"); }
    *out = d;
    return GBGPU_OK;
}

void gbgpu_dense_free(gbgpu_dense *d) { delete d; }

int gbgpu_dense_rref(gbgpu_ctx *ctx, gbgpu_dense *d, int form, int32_t *rank) {
    if (!ctx || !d || d->ctx != ctx) return GBGPU_EINVAL;
    if (form != GBGPU_FORM_PLAIN && form != GBGPU_FORM_MFMA) return set_error(ctx, GBGPU_EINVAL, "bad form");
    int rc = dense_rref(d, form);
    if (rc) return rc;
    if (rank) *rank = d->rank;
    return GBGPU_OK;
}

__global__ void k_copy_rows_full(const uint32_t *a, int64_t ld, int32_t n, const int32_t *orow,
                                 const int32_t *olead, int32_t rank, uint32_t *out) {
    for (int32_t k = blockIdx.y; k < rank; k += gridDim.y) {
        const uint32_t *src = a + (int64_t)orow[k] * ld;
        const int32_t l = olead[k];
        for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
            out[(int64_t)k * n + j] = j < l ? 0u : src[j];
    }
}

int gbgpu_dense_download(gbgpu_ctx *ctx, gbgpu_dense *d, uint32_t *rows, int32_t *lead) {
    if (!ctx || !d || d->ctx != ctx) return GBGPU_EINVAL;
    hipStream_t s = ctx->stream;
    if (d->rank == 0) return GBGPU_OK;
    if (!d->orow) return set_error(ctx, GBGPU_EINVAL, "no echelon form computed");
    int rc;
    if (rows) {
        uint32_t *tmp = nullptr;
        if ((rc = dalloc(ctx, &tmp, (size_t)d->rank * (size_t)d->n))) return rc;
        dim3 g((unsigned)((d->n + 255) / 256 < 64 ? (d->n + 255) / 256 : 64),
               (unsigned)(d->rank < 65535 ? d->rank : 65535));
        k_copy_rows_full<<<g, 256, 0, s>>>(d->a, d->ld, d->n, d->orow, d->olead, d->rank, tmp);
        hipError_t e = hipMemcpyAsync(rows, tmp, (size_t)d->rank * (size_t)d->n * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        dfree(tmp);
        if (e != hipSuccess) return check_hip(ctx, e, "dense download");
    }
    if (lead) {
        hipError_t e = hipMemcpyAsync(lead, d->olead, (size_t)d->rank * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return check_hip(ctx, e, "dense download");
    }
    return GBGPU_OK;
}

}  // extern "C"
