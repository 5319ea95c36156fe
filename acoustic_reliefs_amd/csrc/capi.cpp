// extern "C" entry points of libarh (declared in include/arh.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "arh_internal.h"

namespace arh {
void partition(const double *C, int N, int leaf, double factor, std::vector<int> &perm, std::vector<int> &dir,
               std::vector<int> &app);
void geom_build_index(arh_geom *g);
}  // namespace arh

using namespace arh;

static thread_local std::string tl_err;

namespace arh {
void set_thread_error(const std::string &m) { tl_err = m; }
}  // namespace arh

template <class F>
static int guard(arh_ctx *ctx, F &&f) {
    try {
        f();
        return ARH_OK;
    } catch (const arh::Error &e) {
        tl_err = e.what();
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        tl_err = e.what();
        if (ctx) ctx->err = e.what();
        return ARH_EINVAL;
    }
}

// Every call working for a context: its device is current and its stream owns the buffers the
// call allocates.
static void set_device(arh_ctx *ctx) {
    ARH_HIP(hipSetDevice(ctx->device));
    tl_stream = ctx->stream;
}

void *arh_ctx_s::pinned(size_t bytes) {
    if (bytes > hostpin_bytes) {
        if (hostpin.p) (void)hipHostFree(hostpin.p), hostpin.p = nullptr;
        void *p = nullptr;
        ARH_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        hostpin.p = p;
        hostpin_bytes = bytes;
    }
    return hostpin.p;
}

extern "C" {

int arh_ctx_create(int device, arh_ctx **out) {
    return guard(nullptr, [&] {
        ARH_CHECK(out, ARH_EINVAL, "out is NULL");
        int n = 0;
        ARH_HIP(hipGetDeviceCount(&n));
        ARH_CHECK(device >= 0 && device < n, ARH_EINVAL, "no such HIP device");
        auto c = std::make_unique<arh_ctx>();
        c->device = device;
        ARH_HIP(hipSetDevice(device));
        ARH_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
        c->stream = c->own;
        *out = c.release();
    });
}

}  // extern "C"

// Contexts and geometries are reference counted: a geometry keeps its context alive and an
// H-matrix keeps its geometry alive, so handles may be destroyed in any order.
static void ctx_release(arh_ctx *ctx) {
    if (--ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->hostpin.p) (void)hipHostFree(ctx->hostpin.p), ctx->hostpin.p = nullptr;
    ctx->red.reset();
    ctx->scal.reset();
    if (tl_stream == ctx->stream || tl_stream == ctx->own) tl_stream = nullptr;
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

static void geom_release(arh_geom *g) {
    if (--g->refs > 0) return;
    arh_ctx *ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete g;
    ctx_release(ctx);
}

extern "C" {

void arh_ctx_destroy(arh_ctx *ctx) {
    if (!ctx || ctx->closed) return;
    ctx->closed = true;
    ctx_release(ctx);
}

const char *arh_last_error(const arh_ctx *ctx) { return ctx ? ctx->err.c_str() : tl_err.c_str(); }

int arh_ctx_set_stream(arh_ctx *ctx, void *hip_stream) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx, ARH_EINVAL, "ctx is NULL");
        // buffers belong to the stream they were allocated for: switch before any geometry exists
        ARH_CHECK(ctx->refs == 1, ARH_ESTATE, "the stream can only be set while the context has no geometry");
        set_device(ctx);
        ARH_HIP(hipStreamSynchronize(ctx->stream));
        ctx->red.reset();
        ctx->scal.reset();
        ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own;
        tl_stream = ctx->stream;
    });
}

int arh_ctx_synchronize(arh_ctx *ctx) {
    return guard(ctx, [&] {
        set_device(ctx);
        ARH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int arh_ctx_profile(arh_ctx *ctx, int enable) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx, ARH_EINVAL, "ctx is NULL");
        set_device(ctx);
        ctx->prof.flush(true);
        ctx->prof.on = enable != 0;
        if (enable) ctx->prof.acc = arh_profile{};
    });
}

int arh_ctx_get_profile(arh_ctx *ctx, arh_profile *out) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx && out, ARH_EINVAL, "NULL argument");
        set_device(ctx);
        ctx->prof.flush(true);
        *out = ctx->prof.acc;
    });
}

int arh_device_trim(int device, size_t *freed_bytes) {
    return guard(nullptr, [&] {
        int n = 0;
        ARH_HIP(hipGetDeviceCount(&n));
        ARH_CHECK(device >= 0 && device < n, ARH_EINVAL, "no such device");
        int cur = 0;
        ARH_HIP(hipGetDevice(&cur));
        ARH_HIP(hipSetDevice(device));
        const size_t f = dev_trim(device);
        ARH_HIP(hipSetDevice(cur));
        if (freed_bytes) *freed_bytes = f;
    });
}

int arh_device_cached(int device, size_t *bytes) {
    return guard(nullptr, [&] {
        ARH_CHECK(bytes, ARH_EINVAL, "NULL argument");
        *bytes = dev_cached(device);
    });
}

int arh_malloc(arh_ctx *ctx, size_t bytes, void **dptr) {
    return guard(ctx, [&] {
        set_device(ctx);
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
        ARH_CHECK(e == hipSuccess, ARH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        *dptr = p;
    });
}

int arh_free(arh_ctx *ctx, void *dptr) {
    return guard(ctx, [&] {
        set_device(ctx);
        ARH_HIP(hipStreamSynchronize(ctx->stream));
        if (dptr) ARH_HIP(hipFree(dptr));
    });
}

int arh_upload(arh_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard(ctx, [&] {
        set_device(ctx);
        if (bytes) ARH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        ARH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int arh_download(arh_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard(ctx, [&] {
        set_device(ctx);
        if (bytes) ARH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        ARH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int arh_add(arh_ctx *ctx, int64_t n, const double *x, double *y) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx && (n == 0 || (x && y)) && n >= 0, ARH_EINVAL, "NULL argument or negative length");
        set_device(ctx);
        launch_add(ctx->stream, x, y, n);
    });
}

int arh_partition(const double *C, int N, int leaf_size, double radius_factor, int *perm, int *blocks, int cap,
                  int *n_dense, int *n_lowrank) {
    return guard(nullptr, [&] {
        ARH_CHECK(C && perm && n_dense && n_lowrank, ARH_EINVAL, "NULL argument");
        std::vector<int> p, d, a;
        arh::partition(C, N, leaf_size, radius_factor, p, d, a);
        std::copy(p.begin(), p.end(), perm);
        *n_dense = (int)d.size() / 4;
        *n_lowrank = (int)a.size() / 4;
        if (blocks) {
            ARH_CHECK((int64_t)(d.size() + a.size()) <= 4 * (int64_t)cap, ARH_EINVAL,
                      "arh_partition: block capacity too small");
            std::copy(d.begin(), d.end(), blocks);
            std::copy(a.begin(), a.end(), blocks + d.size());
        }
    });
}

int arh_geom_create(arh_ctx *ctx, const double *Ps, int V, const int *Es, int N, const int *blocks, int n_dense,
                    int n_lowrank, arh_geom **out) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx && Ps && Es && blocks && out, ARH_EINVAL, "NULL argument");
        ARH_CHECK(V > 0 && N > 0 && n_dense > 0 && n_lowrank >= 0, ARH_EINVAL, "empty mesh or partition");
        for (size_t i = 0; i < 3 * (size_t)N; i++)
            ARH_CHECK(Es[i] >= 0 && Es[i] < V, ARH_EINVAL, "element vertex index out of range");
        for (size_t i = 0; i < 3 * (size_t)V; i++)
            ARH_CHECK(std::isfinite(Ps[i]), ARH_EINVAL, "non-finite vertex coordinate");
        for (int b = 0; b < n_dense + n_lowrank; b++) {
            const int *q = blocks + 4 * b;
            ARH_CHECK(q[0] >= 0 && q[0] < q[1] && q[1] <= N && q[2] >= 0 && q[2] < q[3] && q[3] <= N, ARH_EINVAL,
                      "block range out of [0, N)");
        }
        set_device(ctx);
        auto g = std::make_unique<arh_geom>();
        g->ctx = ctx;
        g->V = V, g->N = N;
        g->nd = n_dense, g->na = n_lowrank;
        g->dense.assign(blocks, blocks + 4 * n_dense);
        g->lowrank.assign(blocks + 4 * n_dense, blocks + 4 * (n_dense + n_lowrank));
        std::vector<double> P(Ps, Ps + 3 * (size_t)V);
        std::vector<int> E(Es, Es + 3 * (size_t)N);
        upload(g->P, P, ctx->stream);
        upload(g->E, E, ctx->stream);
        geom_build_index(g.get());
        geom_build_tri(g.get());
        ARH_HIP(hipStreamSynchronize(ctx->stream));
        *out = g.release();
        ctx->refs++;
    });
}

int arh_geom_set_vertices(arh_geom *g, const double *Ps) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && Ps, ARH_EINVAL, "NULL argument");
        for (size_t i = 0; i < 3 * (size_t)g->V; i++)
            ARH_CHECK(std::isfinite(Ps[i]), ARH_EINVAL, "non-finite vertex coordinate");
        set_device(g->ctx);
        ARH_HIP(hipMemcpyAsync(g->P.p, Ps, sizeof(double) * 3 * g->V, hipMemcpyHostToDevice, g->ctx->stream));
        geom_build_tri(g);
        ARH_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int arh_geom_set_heights_device(arh_geom *g, const double *base, const int *vids, int HH, const double *h) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && base && (HH == 0 || (vids && h)), ARH_EINVAL, "NULL argument");
        ARH_CHECK(HH >= 0, ARH_EINVAL, "HH must be >= 0");
        std::vector<char> seen(g->V, 0);
        for (int t = 0; t < HH; t++) {
            if (vids[t] < 0) continue;  // a skipped entry
            ARH_CHECK(vids[t] < g->V, ARH_EINVAL, "vertex id out of range");
            ARH_CHECK(!seen[vids[t]], ARH_EINVAL, "repeated vertex id");
            seen[vids[t]] = 1;
        }
        set_device(g->ctx);
        DBuf vd;
        upload(vd, std::vector<int>(vids, vids + HH), g->ctx->stream);
        geom_set_heights(g, base, vd.as<int>(), HH, h);
    });
}

int arh_geom_set_vertices_device(arh_geom *g, const double *Ps) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && Ps, ARH_EINVAL, "NULL argument");
        set_device(g->ctx);
        ARH_HIP(hipMemcpyAsync(g->P.p, Ps, sizeof(double) * 3 * g->V, hipMemcpyDeviceToDevice, g->ctx->stream));
        geom_build_tri(g);
    });
}

void arh_geom_destroy(arh_geom *g) {
    if (!g || g->closed) return;
    g->closed = true;
    geom_release(g);
}

static void fill_info(arh_hmat *h) {
    arh_geom *g = h->g;
    arh_hmat_info &I = h->info;
    I.n = g->N;
    I.n_dense = g->nd;
    I.n_lowrank = g->na;
    I.n_leaves = g->nleaf;
    I.dense_entries = g->dense_entries;
    I.lowrank_entries = h->u_total + h->v_total;
    I.max_rank = 0;
    I.sum_rank = 0;
    for (auto &B : h->hblocks) I.max_rank = std::max(I.max_rank, B.rank_ref), I.sum_rank += B.rank;
    I.bytes_fp64 = (int64_t)sizeof(double2) * (g->dense_entries + h->u_total + h->v_total);
    I.bytes_fp32 = (int64_t)(h->dense32.bytes + h->denseT32.bytes + h->U32.bytes + h->V32.bytes);
    int64_t fp64_extra = (int64_t)h->denseT.bytes;
    for (auto &L : h->mf) I.bytes_fp32 += (int64_t)(L.panel[0].bytes + L.vs[0].bytes);
    for (auto &L : h->mf) fp64_extra += (int64_t)(L.panel[1].bytes + L.vs[1].bytes);
    I.bytes_total = (int64_t)(h->dense.bytes + h->U.bytes + h->V.bytes + h->blocks.bytes + h->d_active_lr.bytes) +
                    fp64_extra + I.bytes_fp32;
}

int arh_hmat_assemble(arh_geom *g, double k, double aca_tol, int max_rank, int storage, arh_hmat **out) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && out, ARH_EINVAL, "NULL argument");
        ARH_CHECK_K(k);
        ARH_CHECK(max_rank >= 0, ARH_EINVAL, "max_rank must be >= 0");
        ARH_CHECK(storage == ARH_FP64 || storage == ARH_MIXED, ARH_EINVAL, "storage must be ARH_FP64 or ARH_MIXED");
        set_device(g->ctx);
        auto h = std::make_unique<arh_hmat>();
        h->g = g;
        h->k = k;
        h->storage = storage;
        h->max_rank = max_rank;
        h->aca_tol = aca_tol;
        hipEvent_t e0, e1;
        ARH_HIP(hipEventCreate(&e0));
        ARH_HIP(hipEventCreate(&e1));
        ARH_HIP(hipEventRecord(e0, g->ctx->stream));
        assemble_dense(h.get());
        assemble_aca(h.get());
        if (storage == ARH_MIXED) make_fp32_copy(h.get());
        ARH_HIP(hipEventRecord(e1, g->ctx->stream));
        ARH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        ARH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        fill_info(h.get());
        h->info.assemble_ms = ms;
        *out = h.release();
        g->refs++;
    });
}

int arh_hmat_get_info(const arh_hmat *h, arh_hmat_info *info) {
    return guard(nullptr, [&] {
        ARH_CHECK(h && info, ARH_EINVAL, "NULL argument");
        *info = h->info;
    });
}

int arh_hmat_ranks(const arh_hmat *h, int *ranks) {
    return guard(nullptr, [&] {
        ARH_CHECK(h && ranks, ARH_EINVAL, "NULL argument");
        for (size_t a = 0; a < h->hblocks.size(); a++) ranks[a] = h->hblocks[a].rank_ref;
    });
}

int arh_hmat_dense_block(const arh_hmat *h, int b, double *out) {
    return guard(h ? h->g->ctx : nullptr, [&] {
        ARH_CHECK(h && out, ARH_EINVAL, "NULL argument");
        arh_geom *g = h->g;
        ARH_CHECK(b >= 0 && b < g->nd, ARH_EINVAL, "dense block index out of range");
        set_device(g->ctx);
        const int lr = g->dblk_leafr[b];
        const int m = g->leaf_start[lr + 1] - g->leaf_start[lr];
        const int n = g->dense[4 * b + 3] - g->dense[4 * b + 2];
        const int64_t off = g->panel_off[lr] + (int64_t)g->dblk_col[b] * m;
        ARH_HIP(hipMemcpyAsync(out, h->dense.as<double2>() + off, sizeof(double2) * (size_t)m * n,
                               hipMemcpyDeviceToHost, g->ctx->stream));
        ARH_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int arh_hmat_factors(const arh_hmat *h, int a, double *U, double *V, int *rank) {
    return guard(h ? h->g->ctx : nullptr, [&] {
        ARH_CHECK(h && rank, ARH_EINVAL, "NULL argument");
        arh_geom *g = h->g;
        ARH_CHECK(a >= 0 && a < g->na, ARH_EINVAL, "low-rank block index out of range");
        set_device(g->ctx);
        const LRBlock &B = h->hblocks[a];
        const int m = B.re - B.rs, n = B.ce - B.cs, r = B.rank;
        *rank = r;
        if (!U || !V) return;
        std::vector<double2> seg((size_t)m * r);
        if (r > 0) {
            ARH_HIP(hipMemcpyAsync(seg.data(), h->U.as<double2>() + B.u_off, sizeof(double2) * seg.size(),
                                   hipMemcpyDeviceToHost, g->ctx->stream));
            ARH_HIP(hipMemcpyAsync(V, h->V.as<double2>() + B.v_off, sizeof(double2) * (size_t)r * n,
                                   hipMemcpyDeviceToHost, g->ctx->stream));
        }
        ARH_HIP(hipStreamSynchronize(g->ctx->stream));
        // leaf-segmented -> plain m x r column-major
        double2 *Uo = (double2 *)U;
        for (int l = B.l0; l < g->nleaf && g->leaf_start[l] < B.re; l++) {
            const int ls = g->leaf_start[l] - B.rs, ms = g->leaf_start[l + 1] - g->leaf_start[l];
            for (int p = 0; p < r; p++)
                for (int i = 0; i < ms; i++) Uo[(size_t)p * m + ls + i] = seg[(size_t)ls * r + (size_t)p * ms + i];
        }
    });
}

int arh_hmat_compact(arh_hmat *h) {
    return guard(h ? h->g->ctx : nullptr, [&] {
        ARH_CHECK(h, ARH_EINVAL, "h is NULL");
        set_device(h->g->ctx);
        ARH_HIP(hipStreamSynchronize(h->g->ctx->stream));
        for (auto &L : h->mf)
            for (int p = 0; p < 2; p++) {
                L.panel[p].reset(), L.vs[p].reset(), L.xt[p].reset(), L.part[p].reset();
                L.built[p] = false;
            }
        h->dense32.reset(), h->U32.reset(), h->V32.reset(), h->tbuf.reset();
        h->denseT.reset(), h->denseT32.reset();
        h->generic32 = false;
        fill_info(h);
    });
}

void arh_hmat_destroy(arh_hmat *h) {
    if (!h) return;
    arh_geom *g = h->g;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    delete h;
    geom_release(g);
}

int arh_hmat_matvec(arh_hmat *h, int transpose, int precision, int nrhs, const double *x, double *y) {
    return guard(h ? h->g->ctx : nullptr, [&] {
        ARH_CHECK(h && x && y, ARH_EINVAL, "NULL argument");
        ARH_CHECK(x != y, ARH_EINVAL, "x and y must not alias");
        set_device(h->g->ctx);
        hmat_apply(h, transpose, precision, nrhs, (const double2 *)x, (double2 *)y, h->g->ctx->stream);
        ARH_HIP(hipStreamSynchronize(h->g->ctx->stream));
    });
}

int arh_solve(arh_hmat *h, int transpose, int nrhs, const double *b, double *x, const arh_solve_opts *opts,
              arh_solve_stats *stats) {
    return guard(h ? h->g->ctx : nullptr, [&] {
        ARH_CHECK(h && b && x && opts, ARH_EINVAL, "NULL argument");
        ARH_CHECK(b != x, ARH_EINVAL, "b and x must not alias");
        set_device(h->g->ctx);
        solve(h, transpose, nrhs, (const double2 *)b, (double2 *)x, *opts, stats);
    });
}

int arh_rhs(arh_geom *g, double k, int kind, int nrhs, const double *vecs, double *b) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && vecs && b, ARH_EINVAL, "NULL argument");
        ARH_CHECK(kind == ARH_RHS_POINT || kind == ARH_RHS_PLANE, ARH_EINVAL, "unknown rhs kind");
        ARH_CHECK(nrhs >= 1, ARH_EINVAL, "nrhs must be >= 1");
        ARH_CHECK_K(k);
        set_device(g->ctx);
        DBuf v;
        std::vector<double> hv(vecs, vecs + 3 * (size_t)nrhs);
        upload(v, hv, g->ctx->stream);
        launch_rhs(g, k, kind, nrhs, v.as<double>(), (double2 *)b);
        ARH_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int arh_receivers(arh_geom *g, double k, const double *L, int LL, int transpose, int nrhs, const double *x,
                  double *y) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && L && x && y, ARH_EINVAL, "NULL argument");
        ARH_CHECK(LL >= 1 && nrhs >= 1, ARH_EINVAL, "LL and nrhs must be >= 1");
        ARH_CHECK_K(k);
        set_device(g->ctx);
        DBuf Ld;
        std::vector<double> hL(L, L + 3 * (size_t)LL);
        upload(Ld, hL, g->ctx->stream);
        launch_receivers(g, k, Ld.as<double>(), LL, transpose, nrhs, (const double2 *)x, (double2 *)y);
    });
}

int arh_receivers_noalloc(arh_geom *g, double k, const double *L, int LL, int transpose, int nrhs, const double *x,
                          double *y) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && L && x && y, ARH_EINVAL, "NULL argument");
        ARH_CHECK(LL >= 1 && nrhs >= 1, ARH_EINVAL, "LL and nrhs must be >= 1");
        ARH_CHECK_K(k);
        set_device(g->ctx);
        launch_receivers_noalloc(g, k, L, LL, transpose, nrhs, (const double2 *)x, (double2 *)y);
    });
}

int arh_qaca_create(arh_geom *g, double k, const double *L, int LL, double tol, int max_rank, arh_qaca **out) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && L && out, ARH_EINVAL, "NULL argument");
        ARH_CHECK(LL >= 1, ARH_EINVAL, "LL must be >= 1");
        ARH_CHECK(max_rank >= 0, ARH_EINVAL, "max_rank must be >= 0");
        ARH_CHECK_K(k);
        ARH_CHECK(k > 0, ARH_EINVAL, "k must be > 0");
        set_device(g->ctx);
        auto q = std::make_unique<arh_qaca>();
        q->g = g, q->k = k, q->tol = tol, q->LL = LL, q->N = g->N;
        q->lim = std::min(max_rank, std::min(LL - 1, g->N - 1));
        qaca_build(q.get(), L);
        *out = q.release();
        g->refs++;
    });
}

int arh_qaca_rank(const arh_qaca *q, int *rank) {
    return guard(nullptr, [&] {
        ARH_CHECK(q && rank, ARH_EINVAL, "NULL argument");
        *rank = q->rank;
    });
}

int arh_qaca_apply(arh_qaca *q, int transpose, int nrhs, const double *x, double *y) {
    return guard(q ? q->g->ctx : nullptr, [&] {
        ARH_CHECK(q && x && y, ARH_EINVAL, "NULL argument");
        ARH_CHECK(x != y, ARH_EINVAL, "x and y must not alias");
        ARH_CHECK(nrhs >= 1, ARH_EINVAL, "nrhs must be >= 1");
        set_device(q->g->ctx);
        qaca_apply(q, transpose, nrhs, (const double2 *)x, (double2 *)y);
        ARH_HIP(hipStreamSynchronize(q->g->ctx->stream));
    });
}

int arh_qaca_factors(const arh_qaca *q, double *U, double *V) {
    return guard(q ? q->g->ctx : nullptr, [&] {
        ARH_CHECK(q, ARH_EINVAL, "NULL argument");
        set_device(q->g->ctx);
        hipStream_t s = q->g->ctx->stream;
        const size_t r = (size_t)q->rank;
        if (U && r) ARH_HIP(hipMemcpyAsync(U, q->U.p, sizeof(double2) * r * q->LL, hipMemcpyDeviceToHost, s));
        if (V && r) ARH_HIP(hipMemcpyAsync(V, q->V.p, sizeof(double2) * r * q->N, hipMemcpyDeviceToHost, s));
        ARH_HIP(hipStreamSynchronize(s));
    });
}

void arh_qaca_destroy(arh_qaca *q) {
    if (!q) return;
    arh_geom *g = q->g;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    delete q;
    geom_release(g);
}

int arh_relief_create(arh_ctx *ctx, const double *uvs, int n, int tex_h, int tex_w, arh_relief **out) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx && uvs && out, ARH_EINVAL, "NULL argument");
        ARH_CHECK(n >= 1 && tex_h >= 1 && tex_w >= 1, ARH_EINVAL, "n, tex_h and tex_w must be >= 1");
        ARH_CHECK((int64_t)tex_h * tex_w <= (int64_t)1 << 28 && (int64_t)n <= (int64_t)1 << 28, ARH_EINVAL,
                  "texture or vertex count too large");
        set_device(ctx);
        auto r = std::make_unique<arh_relief>();
        r->ctx = ctx, r->n = n, r->h = tex_h, r->w = tex_w;
        relief_build(r.get(), uvs);
        *out = r.release();
        ctx->refs++;
    });
}

int arh_relief_eval(arh_relief *r, const double *tex, double *heights) {
    return guard(r ? r->ctx : nullptr, [&] {
        ARH_CHECK(r && tex && heights, ARH_EINVAL, "NULL argument");
        set_device(r->ctx);
        relief_eval(r, tex, heights);
        ARH_HIP(hipStreamSynchronize(r->ctx->stream));
    });
}

int arh_relief_adjoint(arh_relief *r, const double *grad, double *tex_grad) {
    return guard(r ? r->ctx : nullptr, [&] {
        ARH_CHECK(r && grad && tex_grad, ARH_EINVAL, "NULL argument");
        set_device(r->ctx);
        relief_adjoint(r, grad, tex_grad);
        ARH_HIP(hipStreamSynchronize(r->ctx->stream));
    });
}

void arh_relief_destroy(arh_relief *r) {
    if (!r) return;
    arh_ctx *ctx = r->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete r;
    ctx_release(ctx);
}

int arh_diffusion(arh_ctx *ctx, int LL, int nrhs, const double *y, double *d, double *u, double *vf) {
    return guard(ctx, [&] {
        ARH_CHECK(ctx && y && d, ARH_EINVAL, "NULL argument");
        ARH_CHECK(LL >= 2 && nrhs >= 1, ARH_EINVAL, "LL must be >= 2");
        set_device(ctx);
        DBuf dd;
        dd.alloc(sizeof(double) * nrhs);
        launch_diffusion(ctx->stream, LL, nrhs, (const double2 *)y, dd.as<double>(), u, (double2 *)vf);
        ARH_HIP(hipMemcpyAsync(d, dd.p, sizeof(double) * nrhs, hipMemcpyDeviceToHost, ctx->stream));
        ARH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int arh_shape_grad(arh_geom *g, double k, const double *L, int LL, const int *diff, int HH, int nrhs, const double *x,
                   const double *lam, const double *vf, int src_kind, const double *src, double *A, double *B,
                   double *C) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && L && diff && x && lam && vf && src && A && B && C, ARH_EINVAL, "NULL argument");
        ARH_CHECK(HH >= 0 && nrhs >= 1 && LL >= 1, ARH_EINVAL, "bad sizes");
        ARH_CHECK_K(k);
        set_device(g->ctx);
        DBuf Ld, Sd;
        upload(Ld, std::vector<double>(L, L + 3 * (size_t)LL), g->ctx->stream);
        upload(Sd, std::vector<double>(src, src + 3 * (size_t)nrhs), g->ctx->stream);
        launch_shape_grad(g, k, Ld.as<double>(), LL, diff, HH, nrhs, (const double2 *)x, (const double2 *)lam,
                          (const double2 *)vf, src_kind, Sd.as<double>(), A, B, C);
    });
}

int arh_shape_grad_accumulate(arh_geom *g, double k, const double *L, int LL, const int *diff, int HH, int nrhs,
                              const double *x, const double *lam, const double *vf, int src_kind, const double *src,
                              double *grad) {
    return guard(g ? g->ctx : nullptr, [&] {
        ARH_CHECK(g && L && diff && x && lam && vf && src && grad, ARH_EINVAL, "NULL argument");
        ARH_CHECK(HH >= 0 && nrhs >= 1 && LL >= 1, ARH_EINVAL, "bad sizes");
        ARH_CHECK_K(k);
        set_device(g->ctx);
        DBuf Ld, Sd;
        upload(Ld, std::vector<double>(L, L + 3 * (size_t)LL), g->ctx->stream);
        upload(Sd, std::vector<double>(src, src + 3 * (size_t)nrhs), g->ctx->stream);
        launch_shape_grad(g, k, Ld.as<double>(), LL, diff, HH, nrhs, (const double2 *)x, (const double2 *)lam,
                          (const double2 *)vf, src_kind, Sd.as<double>(), nullptr, nullptr, nullptr, grad);
    });
}

int arh_band_solve(arh_geom *g, double k, double aca_tol, int max_rank, int storage, int kind, int nrhs,
                   const double *vecs, const double *L, int LL, const arh_solve_opts *opts, double *d, double *y,
                   double *x, arh_solve_stats *stats, arh_hmat_info *info) {
    arh_hmat *h = nullptr;
    int rc = arh_hmat_assemble(g, k, aca_tol, max_rank, storage, &h);
    if (rc != ARH_OK) return rc;
    rc = guard(g->ctx, [&] {
        ARH_CHECK(vecs && L && opts && d, ARH_EINVAL, "NULL argument");
        ARH_CHECK(LL >= 2, ARH_EINVAL, "LL must be >= 2");
        set_device(g->ctx);
        hipStream_t s = g->ctx->stream;
        const int N = g->N;
        DBuf vb, bb, xb, yb, Ld, dd;
        upload(vb, std::vector<double>(vecs, vecs + 3 * (size_t)nrhs), s);
        upload(Ld, std::vector<double>(L, L + 3 * (size_t)LL), s);
        bb.alloc(sizeof(double2) * (size_t)N * nrhs);
        double2 *xo = (double2 *)x;
        if (!xo) xb.alloc(sizeof(double2) * (size_t)N * nrhs), xo = xb.as<double2>();
        double2 *yo = (double2 *)y;
        if (!yo) yb.alloc(sizeof(double2) * (size_t)LL * nrhs), yo = yb.as<double2>();
        dd.alloc(sizeof(double) * nrhs);
        launch_rhs(g, k, kind, nrhs, vb.as<double>(), bb.as<double2>());
        // solve (M - I/2) x = -b (adjoint.cpp:95)
        launch_scale(s, bb.as<double2>(), (int64_t)N * nrhs, -1.0);
        solve(h, 0, nrhs, bb.as<double2>(), xo, *opts, stats);
        const auto t1 = std::chrono::steady_clock::now();
        launch_receivers(g, k, Ld.as<double>(), LL, 0, nrhs, xo, yo);
        launch_diffusion(s, LL, nrhs, yo, dd.as<double>(), nullptr, nullptr);
        ARH_HIP(hipMemcpyAsync(d, dd.p, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
        ARH_HIP(hipStreamSynchronize(s));
        if (stats)
            stats->post_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        if (info) *info = h->info;
    });
    arh_hmat_destroy(h);
    return rc;
}

}  // extern "C"
