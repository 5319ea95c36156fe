// acoustics3d_impl: pybind11 drop-in for the reference's nanobind module (acoustics3d.cpp:353-383).
// Same class, method, argument names, order and defaults; additive keyword arguments only
// (speed_of_sound, device). All GPU work goes through the C-ABI of libarh (include/arh.h).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "arh.h"

namespace py = pybind11;
using cd = std::complex<double>;
using f64arr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using i32arr = py::array_t<int, py::array::c_style | py::array::forcecast>;

namespace {

constexpr double kPi = 3.14159265358979323846;  // constants.h:12
constexpr int kApproxMaxRank = 100;              // APPROX_ACA_MAX_K, constants.h:16
constexpr int kQAcaMaxRank = 5000;               // Q_ACA_MAX_K, constants.h:17

void check(int rc, arh_ctx *ctx) {
    if (rc != ARH_OK) throw std::runtime_error(std::string("acoustics3d (gfx950): ") + arh_last_error(ctx));
}

// Device buffer owned through the C-ABI.
struct DevBuf {
    arh_ctx *ctx = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf(arh_ctx *c, size_t n) : ctx(c), bytes(n) { check(arh_malloc(c, n, &p), c); }
    ~DevBuf() {
        if (p) arh_free(ctx, p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    double *d() const { return (double *)p; }
    void up(const void *h, size_t n) { check(arh_upload(ctx, p, h, n), ctx); }
    void down(void *h, size_t n) const { check(arh_download(ctx, h, p, n), ctx); }
};

// helpers.cpp:67-83
std::vector<double> get_frequencies(double band, int n) {
    if (n == 1) return {band};
    double bw = std::pow(2.0, 1.0 / 6.0);
    double lo = std::log(band / bw), hi = std::log(band * bw);
    double dl = (hi - lo) / (n - 1.0);
    std::vector<double> f;
    for (int i = 0; i < n; i++) f.push_back(std::exp(lo + i * dl));
    return f;
}

// helpers.cpp:34-65: elevation from +y, (r sin t cos p, r cos t, -r sin t sin p); ds-stepped loops.
std::vector<double> listener_points(double r, double ds) {
    std::vector<double> el{0}, az{0};
    for (double e = ds; e < 91; e += ds)
        for (double a = 0; a < 360; a += ds) el.push_back(e), az.push_back(a);
    std::vector<double> L(3 * el.size());
    for (size_t i = 0; i < el.size(); i++) {
        double t = el[i] * kPi / 180.0, p = az[i] * kPi / 180.0;
        L[3 * i] = r * std::sin(t) * std::cos(p);
        L[3 * i + 1] = r * std::cos(t);
        L[3 * i + 2] = -r * std::sin(t) * std::sin(p);
    }
    return L;
}

// d and u = dd/dyt from yt (adjoint.cpp:388-397)
double diffusion(const std::vector<double> &yt, std::vector<double> &u) {
    const int LL = (int)yt.size();
    double S = 0, Q = 0;
    for (double y : yt) S += y * y, Q += y * y * y * y;
    double c = (S * S - Q) / ((LL - 1) * Q);
    u.resize(LL);
    for (int l = 0; l < LL; l++) {
        double L = yt[l] * yt[l];
        u[l] = 4 * (S * Q - L * S * S) / ((LL - 1) * Q * Q) * yt[l];
    }
    return c;
}

// Minimal OBJ reader (read_obj.cpp:8-45): 'v x y z', 'f i j k' (1-based).
void read_obj(const std::string &fname, std::vector<double> &P, std::vector<int> &E) {
    std::ifstream fs(fname);
    if (!fs) throw std::runtime_error("cannot open " + fname);
    std::string line;
    while (std::getline(fs, line)) {
        if (line.size() < 2) continue;
        if (line[0] == 'v' && (line[1] == ' ' || line[1] == '\t')) {
            std::istringstream is(line.substr(1));
            double x, y, z;
            if (!(is >> x >> y >> z)) throw std::runtime_error("Invalid vertex specification");
            P.insert(P.end(), {x, y, z});
        } else if (line[0] == 'f' && (line[1] == ' ' || line[1] == '\t')) {
            std::istringstream is(line.substr(1));
            std::string a, b, c;
            if (!(is >> a >> b >> c)) throw std::runtime_error("Invalid face specification (only triangles supported)");
            E.insert(E.end(), {std::stoi(a) - 1, std::stoi(b) - 1, std::stoi(c) - 1});
        }
    }
}

struct Ctx {
    arh_ctx *c = nullptr;
    explicit Ctx(int device) { check(arh_ctx_create(device, &c), nullptr); }
    ~Ctx() { arh_ctx_destroy(c); }
};

// Mesh + partition + device geometry.
struct Mesh {
    std::shared_ptr<Ctx> ctx;
    std::vector<double> P;
    std::vector<int> E;  // reordered
    std::vector<int> blocks;
    int nd = 0, na = 0;
    arh_geom *g = nullptr;
    ~Mesh() { arh_geom_destroy(g); }
    int V() const { return (int)P.size() / 3; }
    int N() const { return (int)E.size() / 3; }
};

void centers(const std::vector<double> &P, const std::vector<int> &E, std::vector<double> &C) {
    const size_t N = E.size() / 3;
    C.resize(3 * N);
    for (size_t e = 0; e < N; e++)
        for (int d = 0; d < 3; d++)
            C[3 * e + d] = (P[3 * E[3 * e] + d] + P[3 * E[3 * e + 1] + d] + P[3 * E[3 * e + 2] + d]) / 3;
}

// set_mesh (acoustics3d.cpp:36-67): partition on the input order, reorder Es.
std::unique_ptr<Mesh> make_mesh(std::shared_ptr<Ctx> ctx, std::vector<double> P, const std::vector<int> &E0, int leaf,
                                double factor, bool print) {
    auto m = std::make_unique<Mesh>();
    m->ctx = ctx;
    m->P = std::move(P);
    const int N = (int)E0.size() / 3, V = (int)m->P.size() / 3;
    for (int v : E0)
        if (v < 0 || v >= V) throw std::runtime_error("element vertex index out of range");
    if (print) std::cout << V << " vertices, " << N << " elements" << std::endl;
    std::vector<double> C;
    centers(m->P, E0, C);
    std::vector<int> perm(N);
    int nd = 0, na = 0;
    check(arh_partition(C.data(), N, leaf, factor, perm.data(), nullptr, 0, &nd, &na), nullptr);
    m->blocks.resize(4 * (size_t)(nd + na));
    check(arh_partition(C.data(), N, leaf, factor, perm.data(), m->blocks.data(), nd + na, &nd, &na), nullptr);
    m->nd = nd, m->na = na;
    m->E.resize(3 * (size_t)N);
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) m->E[3 * i + d] = E0[3 * (size_t)perm[i] + d];
    if (print) {
        std::cout << "tree ok" << std::endl;  // the partition checks ball containment by construction
        std::cout << "Finding direct/approx pairs..." << std::endl;
        std::cout << nd << " direct, " << na << " approx\n" << std::endl;
    }
    check(arh_geom_create(ctx->c, m->P.data(), V, m->E.data(), N, m->blocks.data(), nd, na, &m->g), ctx->c);
    return m;
}

struct Forward {
    double k;
    std::vector<cd> y;     // LL
    std::shared_ptr<DevBuf> x;
    arh_hmat *h = nullptr;
    arh_qaca *q = nullptr;  // listener ACA kept for the backward (stored matrices)
};

}  // namespace

class DiffBEM {
public:
    DiffBEM(int cluster_size, double radius_factor, const std::vector<double> &freq_bands, int n_freqs,
            double approx_ACA_tol, double Q_ACA_tol, double solver_tol, const std::vector<double> &src_pt,
            double listener_radius, double listener_ds, bool recompute_matrices, double speed_of_sound, int device)
        : cluster_size_(cluster_size), radius_factor_(radius_factor), freq_bands_(freq_bands), n_freqs_(n_freqs),
          approx_tol_(approx_ACA_tol), q_tol_(Q_ACA_tol), solver_tol_(solver_tol), listener_radius_(listener_radius),
          listener_ds_(listener_ds), recompute_(recompute_matrices), c0_(speed_of_sound), device_(device) {
        if (src_pt.size() != 3) throw std::invalid_argument("src_pt must have 3 entries");
        src_ = src_pt;
    }

    py::tuple set_mesh(f64arr Ps, i32arr Es) {
        if (Ps.ndim() != 2 || Ps.shape(1) != 3) throw std::invalid_argument("Ps must be (n, 3)");
        if (Es.ndim() != 2 || Es.shape(1) != 3) throw std::invalid_argument("Es must be (n, 3)");
        if (!ctx_) ctx_ = std::make_shared<Ctx>(device_);
        std::vector<double> P(Ps.data(), Ps.data() + Ps.size());
        std::vector<int> E(Es.data(), Es.data() + Es.size());
        {
            py::gil_scoped_release nogil;
            base_dev_.reset();
            mesh_ = make_mesh(ctx_, std::move(P), E, cluster_size_, radius_factor_, true);
            dense_mesh_.reset();
        }
        L_ = listener_points(listener_radius_, listener_ds_);
        mesh_set_ = true;
        return get_mesh();
    }

    void set_diff_points(i32arr diff) {
        // acoustics3d.cpp:69-101: slot per listed vertex (a repeated vertex keeps its last slot)
        Hs_.clear();
        diff_.clear();
        HH_ = 0;
        for (ssize_t t = 0; t < diff.size(); t++) {
            Hs_[diff.data()[t]] = HH_;
            HH_++;
        }
        diff_.assign(HH_, -1);
        for (auto &kv : Hs_) diff_[kv.second] = kv.first;
        diff_set_ = true;
    }

    py::tuple precompute(f64arr Ps, i32arr Es, i32arr diff) {
        set_mesh(Ps, Es);
        set_diff_points(diff);
        return get_mesh();
    }

    double value0() {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return -1;
        }
        double c = 0;
        for (double F : freq_bands_) c += band(mesh_->P, F, true, nullptr);
        return c / freq_bands_.size();
    }

    double value1(f64arr x) {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return 0;
        }
        auto P = deformed(x);
        double c = 0;
        for (double F : freq_bands_) c += band(P, F, true, nullptr);
        return c / freq_bands_.size();
    }

    double band_value(double F) {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return -1;
        }
        return band(mesh_->P, F, true, nullptr);
    }

    std::vector<double> values() {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return {};
        }
        std::vector<double> out;
        for (double F : freq_bands_) out.push_back(band(mesh_->P, F, true, nullptr));
        return out;
    }

    // acoustics3d.cpp:199-234: sum over bands of values and gradients. Each band uses its own
    // forward state (the reference reuses band 0's x and y for later bands, a bug not kept here).
    py::tuple gradient(f64arr x) {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return py::make_tuple(0.0, py::array_t<double>(0));
        }
        if (!diff_set_) {
            std::cerr << "Error! No differentiable points set. Set them using set_diff_pts or precompute." << std::endl;
            return py::make_tuple(-1.0, py::array_t<double>(0));
        }
        auto P = deformed(x);
        std::vector<double> g(HH_, 0.0), gb;
        double c = 0;
        for (double F : freq_bands_) {
            c += band(P, F, false, &gb);
            for (int h = 0; h < HH_; h++) g[h] += gb[h];
        }
        py::array_t<double> out(HH_);
        std::copy(g.begin(), g.end(), out.mutable_data());
        return py::make_tuple(c, out);
    }

    // gradient(x) for heights and a gradient that live on the device (a torch optimiser's tensors): x_ptr and
    // grad_ptr are device addresses of HH float64 values, read and written on the library's stream; the call
    // returns once grad is written. Same arithmetic, in the same order, as gradient(x): the heights offset the
    // base mesh on the device, each band's sum over its frequencies is formed on the device and added to grad.
    double gradient_device(uintptr_t x_ptr, uintptr_t grad_ptr) {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return 0.0;
        }
        if (!diff_set_) {
            std::cerr << "Error! No differentiable points set. Set them using set_diff_pts or precompute." << std::endl;
            return -1.0;
        }
        if (!x_ptr || !grad_ptr) throw std::invalid_argument("x_ptr and grad_ptr must be device addresses");
        double c = 0;
        const double *xd = reinterpret_cast<const double *>(x_ptr);
        double *gd = reinterpret_cast<double *>(grad_ptr);
        std::vector<double> zeros(std::max(1, HH_), 0.0);
        DevBuf band_grad(ctx_->c, sizeof(double) * std::max(1, HH_));
        if (HH_ > 0) check(arh_upload(ctx_->c, gd, zeros.data(), sizeof(double) * HH_), ctx_->c);
        for (double F : freq_bands_) {  // band() releases the GIL around its GPU work
            if (HH_ > 0) band_grad.up(zeros.data(), sizeof(double) * HH_);
            c += band(mesh_->P, F, false, nullptr, xd, band_grad.d());
            if (HH_ > 0) check(arh_add(ctx_->c, HH_, band_grad.d(), gd), ctx_->c);
        }
        check(arh_ctx_synchronize(ctx_->c), ctx_->c);
        return c;
    }

    py::array_t<cd> pvals1(double f) { return pvals2(f, f64arr(py::array_t<double>({(ssize_t)L_.size() / 3, (ssize_t)3}, L_.data()))); }

    py::array_t<cd> pvals2(double f, f64arr listeners) {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return py::array_t<cd>(0);
        }
        if (listeners.ndim() != 2 || listeners.shape(1) != 3) throw std::invalid_argument("listeners must be (n, 3)");
        std::vector<double> L(listeners.data(), listeners.data() + listeners.size());
        const int LL = (int)L.size() / 3;
        std::vector<cd> y(LL);
        {
            py::gil_scoped_release nogil;
            Mesh &m = hmat_mesh(mesh_->P);
            const double k = 2 * kPi * f / c0_;
            auto x = solve_forward(m, k, nullptr, false);
            DevBuf yd(ctx_->c, sizeof(cd) * LL);
            // acoustics3d.cpp:275-276: y = U (V x) from the listener ACA at Q_ACA_tol
            arh_qaca *q = listener_aca(m, k, L);
            const int rc = arh_qaca_apply(q, 0, 1, x->d(), yd.d());
            arh_qaca_destroy(q);
            check(rc, ctx_->c);
            yd.down(y.data(), sizeof(cd) * LL);
        }
        py::array_t<cd> out(LL);
        std::copy(y.begin(), y.end(), out.mutable_data());
        return out;
    }

    py::array_t<cd> surface_vals(double f) {
        if (!mesh_set_) {
            std::cerr << "Error! No mesh set. Set one using set_mesh or precompute." << std::endl;
            return py::array_t<cd>(0);
        }
        const int N = mesh_->N();
        std::vector<cd> xs(N);
        {
            py::gil_scoped_release nogil;
            Mesh &m = hmat_mesh(mesh_->P);
            const double k = 2 * kPi * f / c0_;
            auto x = solve_forward(m, k, nullptr, false);
            x->down(xs.data(), sizeof(cd) * N);
        }
        py::array_t<cd> out(N);
        std::copy(xs.begin(), xs.end(), out.mutable_data());
        return out;
    }

    py::array_t<double> get_listeners() {
        py::array_t<double> out({(ssize_t)L_.size() / 3, (ssize_t)3});
        std::copy(L_.begin(), L_.end(), out.mutable_data());
        return out;
    }

    py::tuple get_mesh() {
        if (!mesh_set_) std::cout << "Warning! No mesh set. Behavior might not be as expected." << std::endl;
        if (!mesh_) return py::make_tuple(py::array_t<double>(std::vector<ssize_t>{0, 3}), py::array_t<int>(std::vector<ssize_t>{0, 3}));
        return mesh_tuple(mesh_->P);
    }

    py::tuple get_mesh_x(f64arr x) {
        if (!mesh_set_) std::cout << "Warning! No mesh set. Behavior might not be as expected." << std::endl;
        if (!diff_set_) std::cout << "Warning! No diff points set. Behavior might not be as expected." << std::endl;
        if (!mesh_) return py::make_tuple(py::array_t<double>(std::vector<ssize_t>{0, 3}), py::array_t<int>(std::vector<ssize_t>{0, 3}));
        return mesh_tuple(deformed(x));
    }

    std::unordered_map<int, int> get_Hs() { return Hs_; }

    void set_band(double F) { freq_bands_.assign(1, F); }

    bool silent = false;
    bool use_actual = false;

private:
    py::tuple mesh_tuple(const std::vector<double> &P) {
        py::array_t<double> Po({(ssize_t)P.size() / 3, (ssize_t)3});
        std::copy(P.begin(), P.end(), Po.mutable_data());
        py::array_t<int> Eo({(ssize_t)mesh_->E.size() / 3, (ssize_t)3});
        std::copy(mesh_->E.begin(), mesh_->E.end(), Eo.mutable_data());
        return py::make_tuple(Po, Eo);
    }

    // acoustics3d.cpp:139-142: y coordinate of each differentiable vertex += x[slot]
    std::vector<double> deformed(f64arr x) {
        std::vector<double> P = mesh_->P;
        if (x.size() < HH_) throw std::invalid_argument("x must have one entry per differentiable point");
        for (auto &kv : Hs_) P[3 * (size_t)kv.first + 1] = x.data()[kv.second] + mesh_->P[3 * (size_t)kv.first + 1];
        return P;
    }

    // The mesh the solves run on: the H-matrix partition, or (use_actual) every leaf pair dense,
    // the stand-in for the reference's dense matrix + colPivHouseholderQr (adjoint.cpp:68-83).
    Mesh &active_mesh(const std::vector<double> &P) {
        Mesh &m = select_mesh();
        check(arh_geom_set_vertices(m.g, P.data()), ctx_->c);
        return m;
    }

    // The H-matrix partition's mesh, or (use_actual) the all-dense one, built on first use.
    Mesh &select_mesh() {
        Mesh *m = mesh_.get();
        if (use_actual) {
            if (!dense_mesh_) {
                dense_mesh_ = std::make_unique<Mesh>();
                dense_mesh_->ctx = ctx_;
                dense_mesh_->P = mesh_->P;
                dense_mesh_->E = mesh_->E;
                std::vector<std::pair<int, int>> leaves;
                for (int b = 0; b < mesh_->nd; b++)
                    if (mesh_->blocks[4 * b] == mesh_->blocks[4 * b + 2])
                        leaves.push_back({mesh_->blocks[4 * b], mesh_->blocks[4 * b + 1]});
                for (auto &r : leaves)
                    for (auto &c : leaves) dense_mesh_->blocks.insert(dense_mesh_->blocks.end(), {r.first, r.second, c.first, c.second});
                dense_mesh_->nd = (int)leaves.size() * (int)leaves.size();
                dense_mesh_->na = 0;
                check(arh_geom_create(ctx_->c, dense_mesh_->P.data(), dense_mesh_->V(), dense_mesh_->E.data(),
                                      dense_mesh_->N(), dense_mesh_->blocks.data(), dense_mesh_->nd, 0, &dense_mesh_->g),
                      ctx_->c);
            }
            m = dense_mesh_.get();
        }
        return *m;
    }

    // active_mesh with the heights on the device: the base mesh (a device copy of the set_mesh vertices, made
    // once) with y of each differentiable point offset by x[slot] (acoustics3d.cpp:210-213).
    Mesh &active_mesh_device(const double *x_dev) {
        Mesh &m = select_mesh();
        if (!base_dev_) {
            base_dev_ = std::make_unique<DevBuf>(ctx_->c, sizeof(double) * mesh_->P.size());
            base_dev_->up(mesh_->P.data(), sizeof(double) * mesh_->P.size());
        }
        check(arh_geom_set_heights_device(m.g, base_dev_->d(), diff_.data(), HH_, x_dev), ctx_->c);
        return m;
    }

    // The H-matrix mesh whatever use_actual says (pvals and surface_vals always take it,
    // acoustics3d.cpp:240-316).
    Mesh &hmat_mesh(const std::vector<double> &P) {
        check(arh_geom_set_vertices(mesh_->g, P.data()), ctx_->c);
        return *mesh_;
    }

    // compute_ACA_listener(..., Q_ACA_tol, Q_ACA_MAX_K) (aca.cpp:84-134) for listeners L at k
    arh_qaca *listener_aca(Mesh &m, double k, const std::vector<double> &L) {
        arh_qaca *q = nullptr;
        check(arh_qaca_create(m.g, k, L.data(), (int)L.size() / 3, q_tol_, kQAcaMaxRank, &q), ctx_->c);
        return q;
    }

    arh_solve_opts opts(bool actual) const {
        arh_solve_opts o{};
        o.tol = actual ? 1e-13 : solver_tol_;
        o.maxit = 0;
        o.true_residual = actual ? 1 : 0;
        o.fail_on_noconv = 0;
        o.hist_cap = 0;
        return o;
    }

    // assemble at k and solve (M - I/2) x = -b for the point source (adjoint.cpp:63-97)
    std::shared_ptr<DevBuf> solve_forward(Mesh &m, double k, arh_hmat **keep, bool actual) {
        arh_hmat *h = nullptr;
        check(arh_hmat_assemble(m.g, k, approx_tol_, kApproxMaxRank, ARH_FP64, &h), ctx_->c);
        const int N = m.N();
        DevBuf b(ctx_->c, sizeof(cd) * N);
        auto x = std::make_shared<DevBuf>(ctx_->c, sizeof(cd) * N);
        check(arh_rhs(m.g, k, ARH_RHS_POINT, 1, src_.data(), b.d()), ctx_->c);
        std::vector<cd> hb(N);
        b.down(hb.data(), sizeof(cd) * N);
        for (auto &v : hb) v = -v;
        b.up(hb.data(), sizeof(cd) * N);
        arh_solve_opts o = opts(actual);
        arh_solve_stats st{};
        int rc = arh_solve(h, 0, 1, b.d(), x->d(), &o, &st);
        if (rc != ARH_OK) {
            arh_hmat_destroy(h);
            check(rc, ctx_->c);
        }
        if (!silent)
            std::cout << st.matvecs << " matrix mults, " << st.iters << " iters, " << st.max_err << " error" << std::endl;
        if (keep)
            *keep = h;
        else
            arh_hmat_destroy(h);
        return x;
    }

    // bem3d::gradient for one band (adjoint.cpp:24-242) on vertex positions P.
    // x_dev (device heights) replaces P when given; grad_dev (device, HH) receives the band's gradient added
    // frequency by frequency instead of grad.
    double band(const std::vector<double> &P, double F, bool forward_only, std::vector<double> *grad,
                const double *x_dev = nullptr, double *grad_dev = nullptr) {
        py::gil_scoped_release nogil;
        Mesh &m = x_dev ? active_mesh_device(x_dev) : active_mesh(P);
        const int N = m.N(), LL = (int)L_.size() / 3;
        auto freqs = get_frequencies(F, n_freqs_);
        std::vector<Forward> fw;
        fw.reserve(freqs.size());
        // kept operators are released on every exit, error paths included
        struct Release {
            std::vector<Forward> &fw;
            ~Release() {
                for (auto &s : fw) {
                    if (s.h) arh_hmat_destroy(s.h);
                    if (s.q) arh_qaca_destroy(s.q);
                }
            }
        } release{fw};
        std::vector<double> yt(LL, 0.0);
        for (double f : freqs) {
            fw.emplace_back();
            Forward &st = fw.back();
            st.k = 2 * kPi * f / c0_;
            arh_hmat *h = nullptr;
            st.x = solve_forward(m, st.k, (forward_only || recompute_) ? nullptr : &h, use_actual);
            st.h = h;
            DevBuf yd(ctx_->c, sizeof(cd) * LL);
            if (use_actual) {
                // the exact listener matrix (adjoint.cpp:75-77)
                check(arh_receivers(m.g, st.k, L_.data(), LL, 0, 1, st.x->d(), yd.d()), ctx_->c);
            } else {
                // y = U (V x) (adjoint.cpp:101-102); kept for the backward unless recomputing
                arh_qaca *q = listener_aca(m, st.k, L_);
                const int rc = arh_qaca_apply(q, 0, 1, st.x->d(), yd.d());
                if (rc != ARH_OK || forward_only || recompute_)
                    arh_qaca_destroy(q);
                else
                    st.q = q;
                check(rc, ctx_->c);
            }
            st.y.resize(LL);
            yd.down(st.y.data(), sizeof(cd) * LL);
            for (int l = 0; l < LL; l++) yt[l] += std::abs(st.y[l]);
        }
        std::vector<double> u;
        const double c = diffusion(yt, u);
        if (forward_only) return c;
        if (!grad_dev) grad->assign(HH_, 0.0);
        DevBuf vfd(ctx_->c, sizeof(cd) * LL), adj(ctx_->c, sizeof(cd) * N), lam(ctx_->c, sizeof(cd) * N);
        DevBuf A(ctx_->c, sizeof(double) * std::max(1, HH_)), B(ctx_->c, sizeof(double) * std::max(1, HH_)),
            C(ctx_->c, sizeof(double) * std::max(1, HH_));
        std::vector<double> hA(HH_), hB(HH_), hC(HH_);
        for (auto &st : fw) {
            // vf = u conj(y) / |y| (adjoint.cpp:145); adj_b = Q^T vf (:198)
            std::vector<cd> vf(LL);
            for (int l = 0; l < LL; l++) vf[l] = u[l] * std::conj(st.y[l]) / std::abs(st.y[l]);
            vfd.up(vf.data(), sizeof(cd) * LL);
            if (use_actual) {
                check(arh_receivers(m.g, st.k, L_.data(), LL, 1, 1, vfd.d(), adj.d()), ctx_->c);
            } else {
                // adj_b = V^T (U^T vf) (adjoint.cpp:177-178 recomputed, :195 stored)
                arh_qaca *q = st.q ? st.q : listener_aca(m, st.k, L_);
                st.q = nullptr;
                const int rc = arh_qaca_apply(q, 1, 1, vfd.d(), adj.d());
                arh_qaca_destroy(q);
                check(rc, ctx_->c);
            }
            arh_hmat *h = st.h;
            if (!h) check(arh_hmat_assemble(m.g, st.k, approx_tol_, kApproxMaxRank, ARH_FP64, &h), ctx_->c);
            arh_solve_opts o = opts(use_actual);
            arh_solve_stats ss{};
            int rc = arh_solve(h, 1, 1, adj.d(), lam.d(), &o, &ss);
            arh_hmat_destroy(h);
            st.h = nullptr;
            check(rc, ctx_->c);
            if (!silent)
                std::cout << ss.matvecs << " matrix mults, " << ss.iters << " iters, " << ss.max_err << " error"
                          << std::endl;
            if (HH_ > 0 && grad_dev) {
                check(arh_shape_grad_accumulate(m.g, st.k, L_.data(), LL, diff_.data(), HH_, 1, st.x->d(), lam.d(),
                                                vfd.d(), ARH_RHS_POINT, src_.data(), grad_dev),
                      ctx_->c);
            } else if (HH_ > 0) {
                check(arh_shape_grad(m.g, st.k, L_.data(), LL, diff_.data(), HH_, 1, st.x->d(), lam.d(), vfd.d(),
                                     ARH_RHS_POINT, src_.data(), A.d(), B.d(), C.d()),
                      ctx_->c);
                A.down(hA.data(), sizeof(double) * HH_);
                B.down(hB.data(), sizeof(double) * HH_);
                C.down(hC.data(), sizeof(double) * HH_);
                for (int h2 = 0; h2 < HH_; h2++) (*grad)[h2] += hA[h2] - hB[h2] + hC[h2];
            }
        }
        return c;
    }

    int cluster_size_;
    double radius_factor_;
    std::vector<double> freq_bands_;
    int n_freqs_;
    double approx_tol_, q_tol_, solver_tol_;
    std::vector<double> src_;
    double listener_radius_, listener_ds_;
    bool recompute_;
    double c0_;
    int device_;
    std::shared_ptr<Ctx> ctx_;
    std::unique_ptr<Mesh> mesh_, dense_mesh_;
    std::unique_ptr<DevBuf> base_dev_;  // device copy of mesh_->P (gradient_device)
    std::vector<double> L_;
    std::unordered_map<int, int> Hs_;
    std::vector<int> diff_;
    int HH_ = 0;
    bool mesh_set_ = false, diff_set_ = false;
};

// sphere_test::sphere (sphere.h:28-85): plane wave e^{-ikx} on an OBJ sphere, leaf 128, factor 1.5,
// ACA and BiCGSTAB tolerance 1e-10; returns the scattered field on (r cos t, 0, -r sin t).
py::array_t<cd> sphere(const std::string &path, double freq, int LL, double lrad, bool actual, double speed_of_sound,
                       int device) {
    std::vector<cd> y(LL);
    {
        py::gil_scoped_release nogil;
        std::vector<double> P;
        std::vector<int> E;
        read_obj(path, P, E);
        auto ctx = std::make_shared<Ctx>(device);
        auto m = make_mesh(ctx, P, E, 128, 1.5, true);
        if (actual) {  // all leaf pairs dense + tight BiCGSTAB in place of the dense QR solve
            std::vector<int> dense;
            std::vector<std::pair<int, int>> leaves;
            for (int b = 0; b < m->nd; b++)
                if (m->blocks[4 * b] == m->blocks[4 * b + 2]) leaves.push_back({m->blocks[4 * b], m->blocks[4 * b + 1]});
            for (auto &r : leaves)
                for (auto &c : leaves) dense.insert(dense.end(), {r.first, r.second, c.first, c.second});
            arh_geom_destroy(m->g);
            m->g = nullptr;
            m->blocks = dense;
            m->nd = (int)dense.size() / 4, m->na = 0;
            check(arh_geom_create(ctx->c, m->P.data(), m->V(), m->E.data(), m->N(), m->blocks.data(), m->nd, 0, &m->g),
                  ctx->c);
        }
        const double k = 2 * kPi * freq / speed_of_sound;
        const double tols = 1e-10;
        arh_hmat *h = nullptr;
        check(arh_hmat_assemble(m->g, k, tols, kApproxMaxRank, ARH_FP64, &h), ctx->c);
        const int N = m->N();
        DevBuf b(ctx->c, sizeof(cd) * N), x(ctx->c, sizeof(cd) * N), yd(ctx->c, sizeof(cd) * LL);
        const double dir[3] = {1.0, 0.0, 0.0};
        check(arh_rhs(m->g, k, ARH_RHS_PLANE, 1, dir, b.d()), ctx->c);
        std::vector<cd> hb(N);
        b.down(hb.data(), sizeof(cd) * N);
        for (auto &v : hb) v = -v;
        b.up(hb.data(), sizeof(cd) * N);
        arh_solve_opts o{};
        o.tol = actual ? 1e-13 : tols;
        o.true_residual = actual ? 1 : 0;
        arh_solve_stats st{};
        int rc = arh_solve(h, 0, 1, b.d(), x.d(), &o, &st);
        arh_hmat_destroy(h);
        check(rc, ctx->c);
        std::cout << st.matvecs << " matrix mults, " << st.iters << " iters, " << st.max_err << " error" << std::endl;
        std::vector<double> L(3 * (size_t)LL, 0.0);
        for (int l = 0; l < LL; l++) {
            double t = (double)l / (double)LL * 2 * kPi;
            L[3 * l] = lrad * std::cos(t);
            L[3 * l + 2] = -lrad * std::sin(t);
        }
        if (actual) {
            // the exact listener matrix (sphere.h:65-67)
            check(arh_receivers(m->g, k, L.data(), LL, 0, 1, x.d(), yd.d()), ctx->c);
        } else {
            // listener ACA at tols (sphere.h:80-81)
            arh_qaca *q = nullptr;
            check(arh_qaca_create(m->g, k, L.data(), LL, tols, kQAcaMaxRank, &q), ctx->c);
            const int rc2 = arh_qaca_apply(q, 0, 1, x.d(), yd.d());
            arh_qaca_destroy(q);
            check(rc2, ctx->c);
        }
        yd.down(y.data(), sizeof(cd) * LL);
    }
    py::array_t<cd> out(LL);
    std::copy(y.begin(), y.end(), out.mutable_data());
    return out;
}

PYBIND11_MODULE(acoustics3d_impl, m) {
    m.doc() = "MI355X (gfx950) drop-in for acoustics3d_cpp (DiffBEM, sphere)";
    py::class_<DiffBEM>(m, "DiffBEM")
        .def(py::init<int, double, const std::vector<double> &, int, double, double, double, const std::vector<double> &,
                      double, double, bool, double, int>(),
             py::arg("cluster_size") = 64, py::arg("radius_factor") = 1.5,
             py::arg("freq_bands") = std::vector<double>{1000}, py::arg("n_freqs") = 1, py::arg("approx_ACA_tol") = 1e-3,
             py::arg("Q_ACA_tol") = 1e-3, py::arg("solver_tol") = 1e-3,
             py::arg("src_pt") = std::vector<double>{0, 100, 0}, py::arg("listener_radius") = 50,
             py::arg("listener_ds") = 5, py::arg("recompute_matrices") = false, py::kw_only(),
             py::arg("speed_of_sound") = 344.0, py::arg("device") = 0)
        .def("set_mesh", &DiffBEM::set_mesh, py::arg("Ps"), py::arg("Es"))
        .def("set_diff_points", &DiffBEM::set_diff_points, py::arg("diff_pts"))
        .def("precompute", &DiffBEM::precompute, py::arg("Ps"), py::arg("Es"), py::arg("diff_pts"))
        .def("value", &DiffBEM::value0)
        .def("value", &DiffBEM::value1, py::arg("x"))
        .def("band_value", &DiffBEM::band_value, py::arg("freq_band"))
        .def("values", &DiffBEM::values)
        .def("gradient", &DiffBEM::gradient, py::arg("x"))
        .def("gradient_device", &DiffBEM::gradient_device, py::arg("x_ptr"), py::arg("grad_ptr"))
        .def("pvals", &DiffBEM::pvals1, py::arg("frequency"))
        .def("pvals", &DiffBEM::pvals2, py::arg("frequency"), py::arg("listeners"))
        .def("get_listeners", &DiffBEM::get_listeners)
        .def("surface_vals", &DiffBEM::surface_vals, py::arg("frequency"))
        .def("get_mesh", &DiffBEM::get_mesh)
        .def("get_mesh", &DiffBEM::get_mesh_x, py::arg("x"))
        .def("get_Hs", &DiffBEM::get_Hs)
        .def("set_band", &DiffBEM::set_band, py::arg("freq_band"))
        .def_readwrite("silent", &DiffBEM::silent)
        .def_readwrite("use_actual", &DiffBEM::use_actual);
    m.def("sphere", &sphere, py::arg("sphere_mesh"), py::arg("freq"), py::arg("LL") = 100, py::arg("lrad") = 10,
          py::arg("actual") = false, py::kw_only(), py::arg("speed_of_sound") = 344.0, py::arg("device") = 0);
}
