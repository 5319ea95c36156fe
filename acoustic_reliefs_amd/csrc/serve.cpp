// Serving: a continuously batched request loop over band solves (arh_server_* in include/arh.h).
//
// Each worker thread owns an arh_ctx (its own HIP stream). Requests wait in one FIFO; the head
// is admitted as soon as a worker is free and its estimated HBM footprint fits the budget (a
// request larger than the whole budget runs alone). Admitted requests overlap on the device: a
// band's latency-bound ACA assembly shares the GPU with other bands' BiCGSTAB iterations. Each
// request runs the same deterministic kernels as arh_band_solve, so its result does not depend on
// what else is in flight.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arh_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct SMesh {
    std::vector<double> P;
    std::vector<int> E;  // reordered
    std::vector<int> blocks;
    int nd = 0, na = 0, V = 0, N = 0;
    int64_t dense_entries = 0, aca_scratch = 0;  // entries (complex)
};

struct Request {
    int64_t id = 0;
    int mesh = 0;
    double k = 0, aca_tol = 0, tol = 0;
    int storage = 0, kind = 0, nrhs = 0;
    std::vector<double> vecs, L;
    std::vector<double> d;
    int64_t footprint = 0;
    bool exclusive = false;  // out of memory once: runs again alone, after a trim of the memory cache
    arh_request_info info{};
    std::string err;
};

}  // namespace

struct arh_server_s {
    arh_server_opts opts{};
    int64_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<SMesh>> meshes;
    std::vector<std::thread> threads;
    std::vector<std::string> worker_err;
    int64_t next_id = 1;
    int live = 0;  // workers that started
    int waiters = 0;  // callers blocked in arh_server_wait
    bool stop = false;
    bool cancelling = false;  // set by arh_server_destroy: nothing new is admitted
    bool profiling = false;
    arh_profile prof{};
    Clock::time_point t0 = Clock::now();
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace arh {
void set_thread_error(const std::string &m);
}

template <class F>
static int sguard(F &&f) {
    try {
        f();
        return ARH_OK;
    } catch (const arh::Error &e) {
        arh::set_thread_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        arh::set_thread_error(e.what());
        return ARH_EINVAL;
    }
}

// HBM footprint estimate of one request: dense panels (+ FP32 copy), ACA scratch and packed factors
// (+ FP32 copy), solver vectors, the receiver matrix and per-element records.
static int64_t footprint(const SMesh &m, int storage, int nrhs, int LL) {
    const bool mixed = storage == ARH_MIXED;
    // FP64 operator + its FP64 MFMA layout (true residuals) + the FP32 MFMA layout when mixed; the
    // low-rank part bounded by the ACA scratch (rank <= lim), plus the scratch itself
    const double dense = 16.0 * m.dense_entries * (mixed ? 2.5 : 2.0);
    const double lowrank = 16.0 * m.aca_scratch * (mixed ? 3.5 : 3.0);
    const double vec = 16.0 * m.N * (double)nrhs * 14.0;
    const double q = 16.0 * (double)LL * m.N;
    return (int64_t)(dense + lowrank + vec + q + 256.0 * m.N + (64 << 20));
}

static void worker(arh_server *s, int wid) {
    arh_ctx *ctx = nullptr;
    if (arh_ctx_create(s->opts.device, &ctx) != ARH_OK) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->worker_err[wid] = arh_last_error(nullptr);
            s->live--;
        }
        s->cv_done.notify_all();
        return;
    }
    int cached = -1;
    arh_geom *geom = nullptr;
    bool prof_on = false;
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                if (s->queue.empty()) return false;
                const Request &h = *s->reqs[s->queue.front()];
                return s->used == 0 || (!h.exclusive && s->used + h.footprint <= s->budget);
            });
            if (s->stop) break;
            r = s->reqs[s->queue.front()].get();
            s->queue.pop_front();
            s->used += r->footprint;
            r->info.status = 1;
            r->info.t_start = s->now();
            prof_on = s->profiling;
        }
        arh_ctx_profile(ctx, prof_on ? 1 : 0);
        int rc = ARH_OK;
        const SMesh &m = *s->meshes[r->mesh];
        double geom_ms = 0.0;
        if (cached != r->mesh) {
            const auto g0 = std::chrono::steady_clock::now();
            arh_geom_destroy(geom);
            geom = nullptr;
            cached = -1;
            rc = arh_geom_create(ctx, m.P.data(), m.V, m.E.data(), m.N, m.blocks.data(), m.nd, m.na, &geom);
            if (rc == ARH_OK) cached = r->mesh;
            geom_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g0).count();
        }
        arh_solve_stats st{};
        arh_hmat_info hi{};
        std::vector<double> d(r->nrhs, 0.0);
        if (rc == ARH_OK) {
            arh_solve_opts o{};
            o.tol = r->tol;
            o.maxit = s->opts.maxit;
            o.true_residual = 1;
            rc = arh_band_solve(geom, r->k, r->aca_tol, 100, r->storage, r->kind, r->nrhs, r->vecs.data(),
                                r->L.data(), (int)r->L.size() / 3, &o, d.data(), nullptr, nullptr, &st, &hi);
        }
        if (rc == ARH_ENOMEM && !r->exclusive) {
            // the footprint estimate let in more than the device holds: free what the memory cache
            // keeps and run the request again when nothing else is in flight (unless the server is
            // being destroyed: then it ends cancelled)
            arh_device_trim(s->opts.device, nullptr);
            bool requeued = false;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->used -= r->footprint;
                if (!s->cancelling) {
                    r->exclusive = true;
                    r->info.status = 0;
                    s->queue.push_front(r->id);
                    requeued = true;
                } else {
                    r->err = "cancelled: the server was destroyed";
                    r->info.t_done = s->now();
                    r->info.status = ARH_ECANCEL;
                }
            }
            s->cv_done.notify_all();
            s->cv_work.notify_all();
            (void)requeued;
            continue;
        }
        arh_profile pr{};
        if (prof_on) arh_ctx_get_profile(ctx, &pr);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (prof_on && s->profiling) {
                auto add = [](arh_kernel_prof &a, const arh_kernel_prof &b) {
                    a.launches += b.launches, a.ms += b.ms, a.bytes += b.bytes, a.flops += b.flops;
                };
                add(s->prof.rows_fp64, pr.rows_fp64);
                add(s->prof.lowrank_fp64, pr.lowrank_fp64);
                add(s->prof.rows_fp32, pr.rows_fp32);
                add(s->prof.lowrank_fp32, pr.lowrank_fp32);
            }
            s->used -= r->footprint;
            r->d = std::move(d);
            r->info.iters = st.iters;
            r->info.max_true_res = st.max_true_res;
            r->info.assemble_ms = hi.assemble_ms;
            r->info.solve_ms = st.solve_ms;
            r->info.post_ms = st.post_ms;
            r->info.geom_ms = geom_ms;
            r->info.t_done = s->now();
            if (rc != ARH_OK) r->err = arh_last_error(ctx);
            r->info.status = rc == ARH_OK ? 2 : rc;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    arh_geom_destroy(geom);
    arh_ctx_destroy(ctx);
}

extern "C" {

int arh_server_create(const arh_server_opts *opts, arh_server **out) {
    return sguard([&] {
        ARH_CHECK(opts && out, ARH_EINVAL, "NULL argument");
        ARH_CHECK(opts->workers >= 1 && opts->workers <= 16, ARH_EINVAL, "workers must be in [1, 16]");
        ARH_CHECK(opts->maxit >= 0, ARH_EINVAL, "maxit must be >= 0");
        auto s = std::make_unique<arh_server>();
        s->opts = *opts;
        ARH_HIP(hipSetDevice(opts->device));
        size_t fr = 0, tot = 0;
        ARH_HIP(hipMemGetInfo(&fr, &tot));
        s->budget = opts->hbm_budget > 0 ? opts->hbm_budget : (int64_t)(0.9 * (double)fr);
        s->worker_err.assign(opts->workers, "");
        s->live = opts->workers;
        for (int w = 0; w < opts->workers; w++) s->threads.emplace_back(worker, s.get(), w);
        *out = s.release();
    });
}

int arh_server_add_mesh(arh_server *s, const double *Ps, int V, const int *Es, int N, int leaf_size,
                        double radius_factor, int *mesh_id) {
    return sguard([&] {
        ARH_CHECK(s && Ps && Es && mesh_id, ARH_EINVAL, "NULL argument");
        ARH_CHECK(V > 0 && N > 0, ARH_EINVAL, "empty mesh");
        auto m = std::make_unique<SMesh>();
        m->V = V, m->N = N;
        m->P.assign(Ps, Ps + 3 * (size_t)V);
        std::vector<double> C(3 * (size_t)N);
        for (int e = 0; e < N; e++)
            for (int d = 0; d < 3; d++) {
                for (int c = 0; c < 3; c++) ARH_CHECK(Es[3 * e + c] >= 0 && Es[3 * e + c] < V, ARH_EINVAL, "vertex index");
                C[3 * e + d] = (Ps[3 * Es[3 * e] + d] + Ps[3 * Es[3 * e + 1] + d] + Ps[3 * Es[3 * e + 2] + d]) / 3;
            }
        std::vector<int> perm(N);
        int nd = 0, na = 0;
        if (arh_partition(C.data(), N, leaf_size, radius_factor, perm.data(), nullptr, 0, &nd, &na) != ARH_OK)
            throw arh::Error(ARH_EINVAL, arh_last_error(nullptr));
        m->blocks.resize(4 * (size_t)(nd + na));
        if (arh_partition(C.data(), N, leaf_size, radius_factor, perm.data(), m->blocks.data(), nd + na, &nd, &na) !=
            ARH_OK)
            throw arh::Error(ARH_EINVAL, arh_last_error(nullptr));
        m->nd = nd, m->na = na;
        m->E.resize(3 * (size_t)N);
        for (int i = 0; i < N; i++)
            for (int d = 0; d < 3; d++) m->E[3 * i + d] = Es[3 * (size_t)perm[i] + d];
        for (int b = 0; b < nd + na; b++) {
            const int *q = &m->blocks[4 * b];
            const int64_t rm = q[1] - q[0], cn = q[3] - q[2];
            if (b < nd)
                m->dense_entries += rm * cn;
            else
                m->aca_scratch += std::max<int64_t>(0, std::min<int64_t>(100, std::min(rm - 1, cn - 1))) * (rm + cn);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->meshes.push_back(std::move(m));
        *mesh_id = (int)s->meshes.size() - 1;
    });
}

int arh_server_submit(arh_server *s, int mesh_id, double k, double aca_tol, int storage, int kind, int nrhs,
                      const double *vecs, const double *L, int LL, double tol, int64_t *request_id) {
    return sguard([&] {
        ARH_CHECK(s && vecs && L && request_id, ARH_EINVAL, "NULL argument");
        ARH_CHECK(nrhs >= 1 && LL >= 2, ARH_EINVAL, "nrhs >= 1 and LL >= 2 required");
        ARH_CHECK(kind == ARH_RHS_POINT || kind == ARH_RHS_PLANE, ARH_EINVAL, "unknown rhs kind");
        ARH_CHECK(storage == ARH_FP64 || storage == ARH_MIXED, ARH_EINVAL, "unknown storage");
        ARH_CHECK_K(k);
        std::unique_lock<std::mutex> lk(s->mu);
        ARH_CHECK(mesh_id >= 0 && mesh_id < (int)s->meshes.size(), ARH_EINVAL, "unknown mesh id");
        for (auto &e : s->worker_err) ARH_CHECK(e.empty(), ARH_EHIP, "server worker failed: " + e);
        auto r = std::make_unique<Request>();
        r->id = s->next_id++;
        r->mesh = mesh_id, r->k = k, r->aca_tol = aca_tol, r->tol = tol;
        r->storage = storage, r->kind = kind, r->nrhs = nrhs;
        r->vecs.assign(vecs, vecs + 3 * (size_t)nrhs);
        r->L.assign(L, L + 3 * (size_t)LL);
        r->footprint = footprint(*s->meshes[mesh_id], storage, nrhs, LL);
        r->info.nrhs = nrhs;
        r->info.footprint = r->footprint;
        r->info.t_submit = s->now();
        *request_id = r->id;
        s->queue.push_back(r->id);
        s->reqs[r->id] = std::move(r);
        lk.unlock();
        s->cv_work.notify_all();
    });
}

int arh_server_wait(arh_server *s, int64_t id, double *d, arh_request_info *info) {
    return sguard([&] {
        ARH_CHECK(s, ARH_EINVAL, "NULL argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        ARH_CHECK(it != s->reqs.end(), ARH_EINVAL, "unknown request id");
        Request *r = it->second.get();
        s->waiters++;
        s->cv_done.wait(lk, [&] { return r->info.status >= 2 || r->info.status < 0 || s->live == 0; });
        s->waiters--;
        s->cv_done.notify_all();  // arh_server_destroy may be waiting for the waiters to leave
        ARH_CHECK(r->info.status != 0 && r->info.status != 1, ARH_EHIP, "no server worker could start");
        if (info) *info = r->info;
        ARH_CHECK(r->info.status == 2, r->info.status, "request failed: " + r->err);
        if (d) std::copy(r->d.begin(), r->d.end(), d);
    });
}

int arh_server_poll(arh_server *s, int64_t id, arh_request_info *info) {
    return sguard([&] {
        ARH_CHECK(s && info, ARH_EINVAL, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        ARH_CHECK(it != s->reqs.end(), ARH_EINVAL, "unknown request id");
        *info = it->second->info;
    });
}

int arh_server_profile(arh_server *s, int enable) {
    return sguard([&] {
        ARH_CHECK(s, ARH_EINVAL, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        s->profiling = enable != 0;
        if (enable) s->prof = arh_profile{};
    });
}

int arh_server_get_profile(arh_server *s, arh_profile *out) {
    return sguard([&] {
        ARH_CHECK(s && out, ARH_EINVAL, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->prof;
    });
}

// Drops every queued request (status ARH_ECANCEL); requests already running finish. Caller holds s->mu.
static int cancel_queued(arh_server *s) {
    int n = 0;
    for (int64_t id : s->queue) {
        Request &r = *s->reqs[id];
        r.err = "cancelled before it was admitted";
        r.info.t_done = s->now();
        r.info.status = ARH_ECANCEL;
        n++;
    }
    s->queue.clear();
    return n;
}

int arh_server_cancel(arh_server *s, int *n_cancelled) {
    return sguard([&] {
        ARH_CHECK(s, ARH_EINVAL, "NULL argument");
        int n = 0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            n = cancel_queued(s);
        }
        s->cv_done.notify_all();
        if (n_cancelled) *n_cancelled = n;
    });
}

void arh_server_destroy(arh_server *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // queued requests are cancelled; running ones finish (a kernel in flight cannot be stopped),
        // then the workers leave and every caller still blocked in arh_server_wait has returned
        s->cancelling = true;
        cancel_queued(s);
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [&] {
            if (s->live > 0)
                for (auto &kv : s->reqs)
                    if (kv.second->info.status == 1) return false;
            return s->queue.empty();
        });
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->live = 0;
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [&] { return s->waiters == 0; });
    }
    delete s;
}

}  // extern "C"
