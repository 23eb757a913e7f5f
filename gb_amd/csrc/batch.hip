// batch.hip -- la_option 1 over many independent F4 matrices, each modulo its
// own prime, as one graph launch.
//
// gb has no counterpart: its rational path eliminates fraction-free over
// GMP integers (la_qq.c:271-362).  A multimodular path instead runs the
// finite-field step once per prime; those runs are independent, and most
// of them are too small to fill an MI355X on their own.  A batch puts them
// on the device together:
//
//   * one persistent launch of the sparse pass over the concatenated
//     to-reduce rows of all matrices (largest matrices first), every wave
//     taking its next row from one queue and reducing it with that
//     matrix's prime, pivots and window layout;
//   * the panel loop of the dense echelon form stepping all blocks at once
//     (one select workgroup per matrix, update tiles of all matrices in
//     one grid), with block counts precomputed from the shapes;
//   * the whole sequence (queue reset, sparse pass, dense init, every panel
//     step) captured once in a HIP graph and replayed; no host
//     synchronisation inside.
//
// The launch parameters of the sparse pass (window width 2^10 .. 2^13
// columns and waves per CU) and the MFMA form's dense update tile (64 x 64
// tiles, or strips of 256 or 512 columns of 64 or 128 rows) are chosen by
// timing the candidates on the batch itself the first time a batch of that
// shape runs in a form; the choice is kept per process.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.h"

struct gbgpu_batch {
    gbgpu_ctx *ctx = nullptr;
    std::vector<gbgpu_mat *> mats;  // owned
    std::vector<int32_t> order;     // matrices by decreasing to-reduce rows
    int form = -1;                  // form the device state is laid out for
    int wshift = 0, waves_per_cu = 0;
    bool big = false;               // some prime needs reduced products
    int nd = 0;                     // digits of the MFMA form's dense update
    int ucols = 64, urows = 64;     // its blocks' tile (columns x rows)
    // sparse pass
    gbgpu::BlockedArgs *d_args = nullptr;
    int32_t *d_rowoff = nullptr;
    int32_t *d_queue = nullptr;
    uint32_t *d_xs = nullptr;
    int64_t xs_stride = 1;
    int grid = 0;
    std::vector<uint32_t *> tinv;
    // dense phase
    std::vector<gbgpu::DenseDesc> hd;
    gbgpu::DenseDesc *d_dense = nullptr;
    std::vector<int32_t> htab;
    int32_t *d_tab = nullptr;
    int32_t nsteps = 0;
    std::vector<void *> dense_bufs;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t replays = 0;
    int32_t nodes = 0;              // kernel launches the graph holds
    bool eager = false;             // launched on the stream, no graph
    ~gbgpu_batch();
};

namespace gbgpu {
namespace {

struct TuneKey {
    int form;
    int32_t nmat;
    int64_t rows, nnz;
    int32_t max_nc;
    bool operator<(const TuneKey &o) const {
        return std::tie(form, nmat, rows, nnz, max_nc) <
               std::tie(o.form, o.nmat, o.rows, o.nnz, o.max_nc);
    }
};
// batches with fewer rows to reduce than this are not tuned
constexpr int64_t SMALL_BATCH_ROWS = 20000;
std::mutex g_tune_mu;
struct Tuned {
    int wshift, waves_per_cu;  // sparse pass
    int ucols, urows;          // dense update tile
};
std::map<TuneKey, Tuned> g_tuned;

__global__ void k_zero_i32(int32_t *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

void free_layout(gbgpu_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    dfree(b->d_args); dfree(b->d_rowoff); dfree(b->d_queue); dfree(b->d_xs);
    dfree(b->d_dense); dfree(b->d_tab);
    for (auto &t : b->tinv) dfree(t);
    b->tinv.clear();
    for (auto &p : b->dense_bufs) pool_free(p);
    b->dense_bufs.clear();
    b->hd.clear();
    b->htab.clear();
    b->form = -1;
    b->replays = 0;
}

template <class T>
int balloc(gbgpu_batch *b, T **p, size_t n) {
    int rc = dalloc(b->ctx, p, n);
    if (!rc) b->dense_bufs.push_back((void *)*p);
    return rc;
}

// window layout of every matrix for 2^wshift columns, plus the blocked
// form's group inverses; the sparse pass's argument table
int layout_sparse(gbgpu_batch *b, int form, int wshift) {
    gbgpu_ctx *ctx = b->ctx;
    int rc;
    for (auto &t : b->tinv) dfree(t);
    b->tinv.assign(b->mats.size(), nullptr);
    std::vector<BlockedArgs> args(b->mats.size());
    std::vector<int32_t> rowoff(b->mats.size() + 1, 0);
    int32_t max_ncl = 1;
    for (size_t k = 0; k < b->order.size(); ++k) {
        gbgpu_mat *m = b->mats[b->order[k]];
        if ((rc = prep_buckets(m, wshift))) return rc;
        if (form == GBGPU_FORM_MFMA && (rc = prep_blocked(m, &b->tinv[b->order[k]]))) return rc;
        BlockedArgs &a = args[k];
        sparse_args(m, &a.a);
        a.tinv = b->tinv[b->order[k]];
        a.gseg = m->gseg;
        a.gbucket = m->gbucket;
        a.ngroups = (m->ncl + 63) / 64;
        rowoff[k + 1] = rowoff[k] + m->nrl;
        max_ncl = std::max(max_ncl, m->ncl);
    }
    b->xs_stride = max_ncl;
    dfree(b->d_args);
    dfree(b->d_rowoff);
    if ((rc = dalloc(ctx, &b->d_args, args.size())) ||
        (rc = dalloc(ctx, &b->d_rowoff, rowoff.size())))
        return rc;
    GBGPU_HIP(ctx, hipMemcpyAsync(b->d_args, args.data(), args.size() * sizeof(BlockedArgs),
                                  hipMemcpyHostToDevice, ctx->stream));
    GBGPU_HIP(ctx, hipMemcpyAsync(b->d_rowoff, rowoff.data(), rowoff.size() * sizeof(int32_t),
                                  hipMemcpyHostToDevice, ctx->stream));
    GBGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    flush_later(ctx);
    b->wshift = wshift;
    return GBGPU_OK;
}

int set_grid(gbgpu_batch *b, int waves_per_cu) {
    gbgpu_ctx *ctx = b->ctx;
    int64_t rows = 0;
    for (auto *m : b->mats) rows += m->nrl;
    int grid = ctx->num_cu * waves_per_cu;
    if (grid > rows) grid = rows > 0 ? (int)rows : 1;
    dfree(b->d_xs);
    int rc = dalloc(ctx, &b->d_xs, (size_t)grid * (size_t)b->xs_stride);
    if (rc) return rc;
    b->grid = grid;
    b->waves_per_cu = waves_per_cu;
    return GBGPU_OK;
}

int launch_sparse(gbgpu_batch *b, int form) {
    gbgpu_ctx *ctx = b->ctx;
    k_zero_i32<<<1, 64, 0, ctx->stream>>>(b->d_queue, 1);
    return batch_sparse_launch(ctx, b->d_args, b->d_rowoff, (int32_t)b->mats.size(), b->d_queue,
                               b->d_xs, b->xs_stride, b->wshift, b->big, form, b->grid);
}

int max_waves(int wshift, int form) {
    int w = (160 * 1024) / sparse_wave_lds(wshift, form);
    return w < 1 ? 1 : (w > 16 ? 16 : w);
}

// times the candidate (window, waves per CU) pairs on the batch itself
int tune(gbgpu_batch *b, int form, int *best_ws, int *best_wpc) {
    gbgpu_ctx *ctx = b->ctx;
    int32_t max_nc = 0;
    for (auto *m : b->mats) max_nc = std::max(max_nc, m->nc);
    float best = -1;
    int rc;
    for (int ws = 10; ws <= 13; ++ws) {
        if ((((int64_t)max_nc + (1 << ws) - 1) >> ws) > 63 && ws < 13) continue;
        if ((rc = layout_sparse(b, form, ws))) return rc;
        const int mw = max_waves(ws, form);
        for (int wpc : {mw, std::max(1, mw / 2)}) {
            if ((rc = set_grid(b, wpc))) return rc;
            float t = 0;
            for (int rep = 0; rep < 2; ++rep) {
                GBGPU_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
                if ((rc = launch_sparse(b, form))) return rc;
                GBGPU_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
                GBGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
                float ms = 0;
                GBGPU_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
                if (rep == 1) t = ms;
            }
            if (best < 0 || t < best) {
                best = t;
                *best_ws = ws;
                *best_wpc = wpc;
            }
            if (wpc == 1) break;
        }
    }
    if (best < 0) return set_error(ctx, GBGPU_ERANGE, "too many columns for the sparse pass");
    return GBGPU_OK;
}

// every step's block counts for the update tile b->ucols x b->urows
int layout_counts(gbgpu_batch *b, int form) {
    gbgpu_ctx *ctx = b->ctx;
    const int32_t nmat = (int32_t)b->mats.size();
    int rc;
    b->htab.assign((size_t)b->nsteps * 4 * (nmat + 1), 0);
    for (int32_t k = 0; k < b->nsteps; ++k)
        batch_dense_counts(b->hd.data(), nmat, k, form, b->ucols, b->urows,
                           b->htab.data() + (size_t)k * 4 * (nmat + 1));
    dfree(b->d_tab);
    if ((rc = dalloc(ctx, &b->d_tab, b->htab.size() ? b->htab.size() : 1))) return rc;
    if (!b->htab.empty())
        GBGPU_HIP(ctx, hipMemcpyAsync(b->d_tab, b->htab.data(), b->htab.size() * sizeof(int32_t),
                                      hipMemcpyHostToDevice, ctx->stream));
    GBGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    flush_later(ctx);
    return GBGPU_OK;
}

int layout_dense(gbgpu_batch *b, int form) {
    gbgpu_ctx *ctx = b->ctx;
    int rc;
    const int32_t nmat = (int32_t)b->mats.size();
    b->hd.assign(nmat, DenseDesc{});
    int32_t max_n = 0;
    // one digit count for the whole batch (enough for its largest prime)
    int nd = 0;
    if (form == GBGPU_FORM_MFMA)
        for (auto *m : b->mats) nd = std::max(nd, digits_for(m->p));
    b->nd = nd;
    for (int32_t i = 0; i < nmat; ++i) {
        gbgpu_mat *m = b->mats[b->order[i]];
        DenseDesc &d = b->hd[i];
        const int32_t ncr = m->nc - m->ncl;
        const int32_t mr = m->nrl;
        d.a = m->dprime;
        d.ld = ncr;
        d.m = mr;
        d.n = ncr;
        d.md = Mod::make(m->p);
        d.r64 = (uint32_t)((((unsigned __int128)1) << 64) % m->p);
        d.ldr = ((int64_t)ncr + 63) & ~63ll;
        d.dnz = m->dnz;
        if ((rc = balloc(b, &d.rowpiv, (size_t)mr)) || (rc = balloc(b, &d.colpiv, (size_t)ncr)) ||
            (rc = balloc(b, &d.active, (size_t)mr)) || (rc = balloc(b, &d.cand, (size_t)mr)) ||
            (rc = balloc(b, &d.L, (size_t)mr * 64)) || (rc = balloc(b, &d.rp, (size_t)64 * d.ldr)))
            return rc;
        if (nd) {
            digits_setup(d.dg, m->p, nd);
            d.dg.lplane = (int64_t)mr * 64;
            d.dg.rplane = d.ldr * 64;
            if ((rc = balloc(b, &d.dg.L, (size_t)nd * d.dg.lplane)) ||
                (rc = balloc(b, &d.dg.R, (size_t)nd * d.dg.rplane)))
                return rc;
        }
        void *pan = pool_alloc(dense_panel_bytes());
        void *st = pool_alloc(dense_liststate_bytes());
        if (pan) b->dense_bufs.push_back(pan);
        if (st) b->dense_bufs.push_back(st);
        if (!pan || !st) return set_error(ctx, GBGPU_ENOMEM, "hipMalloc failed");
        d.pan = (Panel *)pan;
        d.st = (ListState *)st;
        if (mr > 0) max_n = std::max(max_n, ncr);
    }
    b->nsteps = (max_n + 63) / 64;
    if ((rc = dalloc(ctx, &b->d_dense, (size_t)nmat))) return rc;
    GBGPU_HIP(ctx, hipMemcpyAsync(b->d_dense, b->hd.data(), nmat * sizeof(DenseDesc),
                                  hipMemcpyHostToDevice, ctx->stream));
    return layout_counts(b, form);
}

// everything one run launches, in stream order
int enqueue_run(gbgpu_batch *b, int form) {
    gbgpu_ctx *ctx = b->ctx;
    const int32_t nmat = (int32_t)b->mats.size();
    int rc;
    if ((rc = launch_sparse(b, form))) return rc;
    if ((rc = batch_dense_init(ctx, b->d_dense, nmat))) return rc;
    for (int32_t k = 0; k < b->nsteps; ++k)
        if ((rc = batch_dense_step(ctx, b->d_dense, nmat, b->d_tab, b->htab.data(), k, form, b->big,
                                   b->nd, b->ucols, b->urows)))
            return rc;
    return GBGPU_OK;
}

// times the MFMA form's dense update tiles on the batch itself: each
// candidate's run of the dense phase after a fresh sparse pass (the phase
// works on D' in place), the second of two; keeps the fastest
int tune_dense(gbgpu_batch *b, int form) {
    gbgpu_ctx *ctx = b->ctx;
    const int32_t nmat = (int32_t)b->mats.size();
    static const int cand[5][2] = {{64, 64}, {256, 64}, {512, 64}, {256, 128}, {512, 128}};
    float best = -1;
    int bc = b->ucols, br = b->urows, rc;
    for (const auto &c : cand) {
        if (!batch_update_shape_ok(b->nd, c[0], c[1])) continue;
        b->ucols = c[0];
        b->urows = c[1];
        if ((rc = layout_counts(b, form))) return rc;
        float t = 0;
        for (int rep = 0; rep < 2; ++rep) {
            if ((rc = launch_sparse(b, form))) return rc;
            GBGPU_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
            if ((rc = batch_dense_init(ctx, b->d_dense, nmat))) return rc;
            for (int32_t k = 0; k < b->nsteps; ++k)
                if ((rc = batch_dense_step(ctx, b->d_dense, nmat, b->d_tab, b->htab.data(), k, form,
                                           b->big, b->nd, b->ucols, b->urows)))
                    return rc;
            GBGPU_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
            GBGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
            float ms = 0;
            GBGPU_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            if (rep == 1) t = ms;
        }
        if (best < 0 || t < best) {
            best = t;
            bc = c[0];
            br = c[1];
        }
    }
    b->ucols = bc;
    b->urows = br;
    return layout_counts(b, form);
}

int prepare(gbgpu_batch *b, int form) {
    gbgpu_ctx *ctx = b->ctx;
    int rc;
    free_layout(b);
    // pivot split and the sparse pass's output buffers (shapes only)
    for (auto *m : b->mats) {
        if ((rc = prep_split(m, m->ncl))) return rc;
        dfree(m->dprime);
        dfree(m->dnz);
        const int32_t ncr = m->nc - m->ncl;
        if ((rc = dalloc(ctx, &m->dprime, (size_t)m->nrl * (size_t)ncr)) ||
            (rc = dalloc(ctx, &m->dnz, (size_t)m->nrl)))
            return rc;
    }
    b->order.resize(b->mats.size());
    for (size_t i = 0; i < b->order.size(); ++i) b->order[i] = (int32_t)i;
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int32_t x, int32_t y) { return b->mats[x]->nrl > b->mats[y]->nrl; });
    if ((rc = dalloc(ctx, &b->d_queue, 1))) return rc;
    TuneKey key{form, (int32_t)b->mats.size(), 0, 0, 0};
    for (auto *m : b->mats) {
        key.rows += m->nrl;
        key.nnz += m->nnz;
        key.max_nc = std::max(key.max_nc, m->nc);
    }
    Tuned tp{0, 0, 0, 0};
    bool have = false;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            tp = it->second;
            have = true;
        }
    }
    const bool small = key.rows < SMALL_BATCH_ROWS;
    if (!have && small) {
        // a small batch (the images of a multimodular round, say) runs in
        // less time than timing the candidates would take: the smallest
        // window that covers the widest matrix in 63, at full occupancy
        tp.wshift = 10;
        while ((((int64_t)key.max_nc + (1 << tp.wshift) - 1) >> tp.wshift) > 63 && tp.wshift < 13)
            ++tp.wshift;
        tp.waves_per_cu = max_waves(tp.wshift, form);
    }
    if (!have && !small && (rc = tune(b, form, &tp.wshift, &tp.waves_per_cu))) return rc;
    if ((rc = layout_sparse(b, form, tp.wshift))) return rc;
    if ((rc = set_grid(b, tp.waves_per_cu))) return rc;
    // the dense update tile: the one kept for this shape, else 512-column
    // strips for three and four digits and 64 x 64 tiles otherwise, timed
    // against the other tiles on a batch large enough
    int nd = 0;
    if (form == GBGPU_FORM_MFMA)
        for (auto *m : b->mats) nd = std::max(nd, digits_for(m->p));
    b->ucols = have ? tp.ucols : (nd >= 3 ? 512 : 64);
    b->urows = have ? tp.urows : 64;
    // GBGPU_BATCH_TILE=cols,rows: that tile, untimed (one of the candidates)
    const char *tl = getenv("GBGPU_BATCH_TILE");
    bool forced = false;
    if (tl && tl[0] && form == GBGPU_FORM_MFMA) {
        int c = 0, r = 0;
        if (sscanf(tl, "%d,%d", &c, &r) != 2 || !batch_update_shape_ok(nd, c, r))
            return set_error(ctx, GBGPU_EINVAL, std::string("GBGPU_BATCH_TILE: no update tile ") + tl);
        b->ucols = c;
        b->urows = r;
        forced = true;
    }
    if ((rc = layout_dense(b, form))) return rc;
    if (!have && !small && !forced && form == GBGPU_FORM_MFMA && (rc = tune_dense(b, form))) return rc;
    if (!have && !forced) {
        tp.ucols = b->ucols;
        tp.urows = b->urows;
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tuned[key] = tp;
    }
    // GBGPU_BATCH_EAGER=1: the same launches straight on the stream, each
    // run (for tracing tools that follow stream launches one by one)
    const char *eg = getenv("GBGPU_BATCH_EAGER");
    b->eager = eg && eg[0] == '1';
    if (b->eager) {
        b->form = form;
        return GBGPU_OK;
    }
    // capture the run
    hipStream_t s = ctx->stream;
    GBGPU_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = enqueue_run(b, form);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return check_hip(ctx, e, "hipStreamEndCapture");
    b->graph = g;
    size_t nn = 0;
    GBGPU_HIP(ctx, hipGraphGetNodes(g, nullptr, &nn));
    b->nodes = (int32_t)nn;
    GBGPU_HIP(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->form = form;
    return GBGPU_OK;
}

}  // namespace
}  // namespace gbgpu

using namespace gbgpu;

gbgpu_batch::~gbgpu_batch() {
    free_layout(this);
    for (auto *m : mats) delete m;
}

extern "C" {

int gbgpu_batch_upload(gbgpu_ctx *ctx, const gbgpu_csr_t *mats, int32_t nmat, gbgpu_batch **out) {
    if (!ctx || !out || nmat < 0 || (nmat > 0 && !mats)) return GBGPU_EINVAL;
    *out = nullptr;
    gbgpu_batch *b = new gbgpu_batch();
    b->ctx = ctx;
    for (int32_t i = 0; i < nmat; ++i) {
        gbgpu_mat *m = nullptr;
        int rc = gbgpu_mat_upload(ctx, &mats[i], &m);
        if (rc) {
            delete b;
            return rc;
        }
        b->mats.push_back(m);
        if (m->p >= (1u << 17)) b->big = true;
    }
    *out = b;
    return GBGPU_OK;
}

int gbgpu_batch_la(gbgpu_ctx *ctx, gbgpu_batch *b, int form, int32_t *np) {
    if (!ctx || !b || b->ctx != ctx) return GBGPU_EINVAL;
    if (form != GBGPU_FORM_PLAIN && form != GBGPU_FORM_MFMA)
        return set_error(ctx, GBGPU_EINVAL, "bad form");
    GBGPU_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    if (b->form != form && (rc = prepare(b, form))) {
        free_layout(b);
        return rc;
    }
    hipStream_t s = ctx->stream;
    GBGPU_HIP(ctx, hipEventRecord(ctx->ev[4], s));
    if (b->eager) {
        if ((rc = enqueue_run(b, form))) return rc;
    } else {
        GBGPU_HIP(ctx, hipGraphLaunch(b->exec, s));
    }
    GBGPU_HIP(ctx, hipEventRecord(ctx->ev[5], s));
    GBGPU_HIP(ctx, hipStreamSynchronize(s));
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]) == hipSuccess) ctx->phase_ms[2] = ms;
    ctx->phase_ms[0] = ctx->phase_ms[1] = -1;
    ++b->replays;
    // rows of every echelon form (decreasing lead, columns shifted by ncl)
    for (size_t k = 0; k < b->order.size(); ++k) {
        gbgpu_mat *m = b->mats[b->order[k]];
        const DenseDesc &hd = b->hd[k];
        dfree(m->orp); dfree(m->oci); dfree(m->ocf);
        m->onp = 0;
        m->onnz = 0;
        gbgpu_dense view;
        view.ctx = ctx;
        view.m = hd.m;
        view.n = hd.n;
        view.ld = hd.ld;
        view.p = m->p;
        view.a = m->dprime;
        view.own_a = false;
        if ((rc = dense_output(&view, hd.colpiv))) return rc;
        if ((rc = dense_to_rows(&view, m->ncl, 0, &m->orp, &m->oci, &m->ocf, &m->onnz))) return rc;
        m->onp = view.rank;
        if (np) np[b->order[k]] = view.rank;
    }
    return GBGPU_OK;
}

int gbgpu_batch_rows_download(gbgpu_ctx *ctx, gbgpu_batch *b, int32_t i, gbgpu_rows_t *out) {
    if (!ctx || !b || b->ctx != ctx || i < 0 || i >= (int32_t)b->mats.size()) return GBGPU_EINVAL;
    return gbgpu_rows_download(ctx, b->mats[i], out);
}

int gbgpu_batch_info(gbgpu_batch *b, int32_t *out, int32_t nout) {
    if (!b || !out) return GBGPU_EINVAL;
    const int32_t v[8] = {(int32_t)b->mats.size(), b->wshift, b->waves_per_cu, b->grid, b->replays,
                          b->nodes, b->ucols, b->urows};
    for (int32_t i = 0; i < nout && i < 8; ++i) out[i] = v[i];
    return GBGPU_OK;
}

void gbgpu_batch_free(gbgpu_batch *b) {
    if (!b) return;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    delete b;
}

}  // extern "C"
