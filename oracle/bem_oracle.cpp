// CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// An independent FP64 C++17 + OpenMP restatement of the acoustics3d_cpp Helmholtz BEM
// (mickey1356/acoustic_reliefs). Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker or the timed CPU
// baseline. The product (acoustic_reliefs_amd) never links, loads or calls it.
//
// Every function cites the reference file:line it restates. Where the reference relies on
// Eigen (absent from the reference snapshot: acoustics3d_cpp/extern/eigen is an empty, unpinned
// git submodule, Eigen >= 3.4 is implied by Eigen::seq/placeholders::all), the published
// Eigen 3.4 algorithm is restated (BiCGSTAB) or replaced by a mathematically equivalent
// routine (dense LU with partial pivoting instead of colPivHouseholderQr).
//
// Pinning (what checks each part; the tests hold them):
//   entry / block / row / col   : k=0 Gauss-law sums, degree-4 exactness of the rule,
//                                 adaptive brute-force quadrature, M_ii == 0
//   tree / partition            : ball check, exact cover of N^2, permutation
//   ACA                         : SVD error bound, exact recovery of synthetic rank-r blocks
//                                 (rank exactly r), a hand-traced 3x3 complex cross sequence
//                                 (non-conjugated estimate, discarded stopping cross)
//   H-matvec                    : dense brute force (M - I/2) x
//   BiCGSTAB                    : dense LU solution, residual <= tol, finite termination on
//                                 matrices with 1, 2 and 3 distinct eigenvalues
//   right-hand sides            : point source -> plane wave in the far field, by hand near by
//   listener points, bands      : entries by hand, closed form of the log spacing
//   diffusion coefficient / u   : closed forms (uniform -> 1, single receiver -> 0), FD of u
//   derivatives / gradient      : central FD of compute_M, FD of the whole forward
//   full solve                  : analytic rigid-sphere series
// Parity with the reference's own outputs is unpinned: the reference ships no golden
// vectors, no tests and cannot be built here.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;  // constants.h:12

// Degree-4, 6-point triangle rule (weight, bary_b, bary_c), constants.h:39-46. The table must be
// numerically identical to the reference's (these are the published Strang-Fix / Dunavant values
// truncated exactly as the reference truncates them).
static const double kQuad[6][3] = {
    {0.223381589678011, 0.445948490915965, 0.108103018168070},
    {0.223381589678011, 0.445948490915965, 0.445948490915965},
    {0.223381589678011, 0.108103018168070, 0.445948490915965},
    {0.109951743655322, 0.091576213509771, 0.816847572980459},
    {0.109951743655322, 0.091576213509771, 0.091576213509771},
    {0.109951743655322, 0.816847572980459, 0.091576213509771},
};

struct V3 {
    double x, y, z;
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct Mesh {
    const double *P;  // V x 3 row-major
    const int *E;     // N x 3 row-major (0-based)
    int V, N;
    V3 vert(int v) const { return {P[3 * v], P[3 * v + 1], P[3 * v + 2]}; }
    V3 a(int e) const { return vert(E[3 * e]); }
    V3 b(int e) const { return vert(E[3 * e + 1]); }
    V3 c(int e) const { return vert(E[3 * e + 2]); }
};

// helpers.cpp:16-32: centroid (a+b+c)/3 and the unnormalised normal ab x ac.
static void intermediates(const Mesh &m, std::vector<V3> &C, std::vector<V3> &Nrm) {
    C.resize(m.N);
    Nrm.resize(m.N);
    for (int e = 0; e < m.N; e++) {
        V3 a = m.a(e), b = m.b(e), c = m.c(e);
        C[e] = {(a.x + b.x + c.x) / 3, (a.y + b.y + c.y) / 3, (a.z + b.z + c.z) / 3};
        V3 ab = b - a, ac = c - a;
        Nrm[e] = {ac.z * ab.y - ac.y * ab.z, ac.x * ab.z - ac.z * ab.x, ac.y * ab.x - ac.x * ab.y};
    }
}

// integrate.h:11-19: K(p, q, n) = dG/dr * (-(p-q).n / r), dG/dr = e^{-ikr}(-ikr-1)/(4 pi r^2).
static inline cd kernel(double k, V3 p, V3 q, V3 n) {
    V3 d = p - q;
    double r = norm(d);
    cd e = std::exp(cd(0.0, -k * r));
    cd dGdr = e * cd(-1.0, -k * r) / (4 * kPi * r * r);
    return dGdr * (-dot(d, n) / r);
}

// integrate.cpp:5-14: M = 1/2 sum_g w_g K(p, a + xi_g ab + eta_g ac, n).
static inline cd entry(double k, V3 p, V3 qa, V3 qb, V3 qc, V3 n) {
    V3 ab = qb - qa, ac = qc - qa;
    cd res = 0;
    for (int g = 0; g < 6; g++) {
        V3 pt = qa + kQuad[g][1] * ab + kQuad[g][2] * ac;
        res += kQuad[g][0] * kernel(k, p, pt, n);
    }
    return res * 0.5;
}

struct Geo {
    Mesh m;
    std::vector<V3> C, Nrm;
    cd M(double k, V3 p, int j) const { return entry(k, p, m.a(j), m.b(j), m.c(j), Nrm[j]); }
};

// ---------------------------------------------------------------- cluster tree (clustertree.cpp)
struct Node {
    int s, e;  // half-open element range in the reordered numbering
    V3 center;
    double radius;
    bool leaf;
    int left = -1, right = -1;
};

struct Tree {
    std::vector<Node> nodes;
    std::vector<int> perm;  // reordering: new position -> original element id
    int max_pts;

    // clustertree.cpp:60-101. Mean of the centroids (sequential sum / n), radius = max distance,
    // split the longest bbox axis (first on ties) at the mean, '<' goes left, order preserved.
    // The reference recurses forever when one side is empty; here that is an error.
    int build(int start, const std::vector<int> &idx, const std::vector<V3> &C) {
        const int n = (int)idx.size();
        V3 sum = {0, 0, 0};
        for (int i : idx) sum = sum + C[i];
        V3 ctr = {sum.x / n, sum.y / n, sum.z / n};
        double rad = 0;
        for (int i : idx) rad = std::max(rad, norm(C[i] - ctr));
        int id = (int)nodes.size();
        nodes.push_back(Node{start, start + n, ctr, rad, true});
        if (n <= max_pts) {
            for (int t = 0; t < n; t++) perm[start + t] = idx[t];
            return id;
        }
        nodes[id].leaf = false;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i : idx) {
            double v[3] = {C[i].x, C[i].y, C[i].z};
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], v[d]);
                hi[d] = std::max(hi[d], v[d]);
            }
        }
        int dim = 0;
        double best = hi[0] - lo[0];
        for (int d = 1; d < 3; d++)
            if (hi[d] - lo[d] > best) {
                best = hi[d] - lo[d];
                dim = d;
            }
        double gamma = dim == 0 ? ctr.x : (dim == 1 ? ctr.y : ctr.z);
        std::vector<int> L, R;
        for (int i : idx) {
            double v = dim == 0 ? C[i].x : (dim == 1 ? C[i].y : C[i].z);
            (v < gamma ? L : R).push_back(i);
        }
        if (L.empty() || R.empty()) throw std::runtime_error("cluster tree: degenerate split (coincident centroids)");
        int l = build(start, L, C);
        int r = build(start + (int)L.size(), R, C);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }

    // clustertree.cpp:160-187. Admissible iff dist(centres) > factor (r_s + r_t).
    void traverse(int s, int t, double factor, std::vector<int> &dir, std::vector<int> &app) const {
        const Node &S = nodes[s], &T = nodes[t];
        double dist = norm(S.center - T.center);
        if (dist > factor * (S.radius + T.radius)) {
            app.insert(app.end(), {S.s, S.e, T.s, T.e});
        } else if (S.leaf && T.leaf) {
            dir.insert(dir.end(), {S.s, S.e, T.s, T.e});
        } else {
            bool split_src = T.leaf || ((S.e - S.s) > (T.e - T.s));
            if (split_src) {
                traverse(S.left, t, factor, dir, app);
                traverse(S.right, t, factor, dir, app);
            } else {
                traverse(s, T.left, factor, dir, app);
                traverse(s, T.right, factor, dir, app);
            }
        }
    }
};

// ---------------------------------------------------------------- ACA (aca.cpp)
// aca.cpp:9-36: first maximum of |v|^2 (strict '>', start at -1 so NaN never wins).
static int argmax_abs(const cd *v, int n, const std::vector<char> *used) {
    int best = 0;
    double bv = -1;
    for (int i = 0; i < n; i++) {
        if (used && (*used)[i]) continue;
        double a = std::norm(v[i]);
        if (a > bv) {
            bv = a;
            best = i;
        }
    }
    return best;
}

// Partially pivoted ACA of an m x n block whose rows/cols come from callbacks (aca.cpp:38-82,
// 84-134). U is m x r column-major, V is r x n row-major. guard_zero_pivot selects the block
// variant (aca.cpp:56-60); the listener variant divides without a guard (aca.cpp:104).
template <class RowF, class ColF>
static int aca(int m, int n, double eps, int max_iter, RowF row, ColF col, bool guard_zero_pivot, std::vector<cd> &U,
               std::vector<cd> &Vr) {
    const int lim = std::min(max_iter, std::min(m - 1, n - 1));
    if (lim <= 0) {
        U.clear();
        Vr.clear();
        return 0;
    }
    U.assign((size_t)m * lim, 0.0);
    Vr.assign((size_t)lim * n, 0.0);
    std::vector<char> used(m, 0);
    used[0] = 1;
    std::vector<cd> v(n), u(m);
    int i = 0, it = 0;
    double err = 0;
    while (it < lim) {
        row(i, v.data());  // aca.cpp:53-54: residual row
        for (int p = 0; p < it; p++) {
            cd uip = U[(size_t)p * m + i];
            const cd *vp = &Vr[(size_t)p * n];
            for (int c = 0; c < n; c++) v[c] -= uip * vp[c];
        }
        int j = argmax_abs(v.data(), n, nullptr);
        cd piv = v[j];
        cd *vrow = &Vr[(size_t)it * n];
        if (guard_zero_pivot && piv == 0.0) {
            for (int c = 0; c < n; c++) vrow[c] = v[c];
        } else {
            for (int c = 0; c < n; c++) vrow[c] = v[c] / piv;
        }
        col(j, u.data());  // aca.cpp:61-63: residual column
        for (int p = 0; p < it; p++) {
            cd vpj = Vr[(size_t)p * n + j];
            const cd *up = &U[(size_t)p * m];
            for (int q = 0; q < m; q++) u[q] -= up[q] * vpj;
        }
        cd *ucol = &U[(size_t)it * m];
        for (int q = 0; q < m; q++) ucol[q] = u[q];
        double vn = 0, un = 0;
        for (int c = 0; c < n; c++) vn += std::norm(vrow[c]);
        for (int q = 0; q < m; q++) un += std::norm(ucol[q]);
        // aca.cpp:64-69: non-conjugated products, absolute values.
        double dots = 0;
        for (int p = 0; p < it; p++) {
            cd su = 0, sv = 0;
            const cd *up = &U[(size_t)p * m];
            const cd *vp = &Vr[(size_t)p * n];
            for (int q = 0; q < m; q++) su += ucol[q] * up[q];
            for (int c = 0; c < n; c++) sv += vrow[c] * vp[c];
            dots += std::abs(su) * std::abs(sv);
        }
        err += un * vn + 2 * dots;
        if (std::sqrt(un * vn / err) <= eps) break;  // aca.cpp:71, current cross discarded
        i = argmax_abs(ucol, m, &used);
        used[i] = 1;
        it++;
    }
    U.resize((size_t)m * it);
    Vr.resize((size_t)it * n);
    return it;
}

// ---------------------------------------------------------------- H-matrix (hmatrix.h/.cpp)
struct HMat {
    int N;
    std::vector<int> dir, app;                // (rs,re,cs,ce) quadruples
    std::vector<std::vector<cd>> D;           // dense leaves, column-major (integrate.cpp:87-100)
    std::vector<std::vector<cd>> U, Vr;       // low-rank factors
    std::vector<int> rank;
};

static void hmat_assemble(HMat &H, const Geo &g, double k, double aca_tol, int max_rank) {
    const int nd = (int)H.dir.size() / 4, na = (int)H.app.size() / 4;
    H.D.assign(nd, {});
    H.U.assign(na, {});
    H.Vr.assign(na, {});
    H.rank.assign(na, 0);
    int nan_found = 0;
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < nd; b++) {
        int rs = H.dir[4 * b], re = H.dir[4 * b + 1], cs = H.dir[4 * b + 2], ce = H.dir[4 * b + 3];
        int m = re - rs, n = ce - cs;
        auto &blk = H.D[b];
        blk.resize((size_t)m * n);
        for (int c = 0; c < n; c++)
            for (int r = 0; r < m; r++) {
                cd v = g.M(k, g.C[rs + r], cs + c);
                if (std::isnan(v.real()) || std::isnan(v.imag())) nan_found = 1;
                blk[(size_t)c * m + r] = v;
            }
    }
    if (nan_found) throw std::runtime_error("NaN in a direct block (coincident collocation point and panel)");
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < na; b++) {
        int rs = H.app[4 * b], re = H.app[4 * b + 1], cs = H.app[4 * b + 2], ce = H.app[4 * b + 3];
        int m = re - rs, n = ce - cs;
        auto row = [&](int i, cd *out) {  // integrate.cpp:59-72
            for (int c = 0; c < n; c++) out[c] = g.M(k, g.C[rs + i], cs + c);
        };
        auto col = [&](int j, cd *out) {  // integrate.cpp:74-85
            for (int r = 0; r < m; r++) out[r] = g.M(k, g.C[rs + r], cs + j);
        };
        H.rank[b] = aca(m, n, aca_tol, max_rank, row, col, true, H.U[b], H.Vr[b]);
    }
}

// hmatrix.h:84-140: y = -x/2 + sum_b D_b x + sum_b U_b (V_b x); transposed: D^T, V^T (U^T x).
// Blocks are accumulated in list order per thread and thread partials summed in thread order.
static void hmat_matvec(const HMat &H, bool transpose, const cd *x, cd *y) {
    const int N = H.N;
    const int nd = (int)H.dir.size() / 4, na = (int)H.app.size() / 4;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<cd>> part(nt, std::vector<cd>(N, 0.0));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        cd *acc = part[tid].data();
        std::vector<cd> t;
#pragma omp for schedule(static)
        for (int b = 0; b < nd + na; b++) {
            if (b < nd) {
                int rs = H.dir[4 * b], re = H.dir[4 * b + 1], cs = H.dir[4 * b + 2], ce = H.dir[4 * b + 3];
                int m = re - rs, n = ce - cs;
                const cd *D = H.D[b].data();
                if (!transpose) {
                    for (int c = 0; c < n; c++)
                        for (int r = 0; r < m; r++) acc[rs + r] += D[(size_t)c * m + r] * x[cs + c];
                } else {
                    for (int c = 0; c < n; c++) {
                        cd s = 0;
                        for (int r = 0; r < m; r++) s += D[(size_t)c * m + r] * x[rs + r];
                        acc[cs + c] += s;
                    }
                }
            } else {
                int a = b - nd;
                int rs = H.app[4 * a], re = H.app[4 * a + 1], cs = H.app[4 * a + 2], ce = H.app[4 * a + 3];
                int m = re - rs, n = ce - cs, r = H.rank[a];
                const cd *U = H.U[a].data();
                const cd *Vr = H.Vr[a].data();
                t.assign(r, 0.0);
                if (!transpose) {
                    for (int p = 0; p < r; p++)
                        for (int c = 0; c < n; c++) t[p] += Vr[(size_t)p * n + c] * x[cs + c];
                    for (int p = 0; p < r; p++)
                        for (int q = 0; q < m; q++) acc[rs + q] += U[(size_t)p * m + q] * t[p];
                } else {
                    for (int p = 0; p < r; p++)
                        for (int q = 0; q < m; q++) t[p] += U[(size_t)p * m + q] * x[rs + q];
                    for (int p = 0; p < r; p++)
                        for (int c = 0; c < n; c++) acc[cs + c] += Vr[(size_t)p * n + c] * t[p];
                }
            }
        }
    }
    for (int i = 0; i < N; i++) {
        cd s = -0.5 * x[i];
        for (int t = 0; t < nt; t++) s += part[t][i];
        y[i] = s;
    }
}

// ---------------------------------------------------------------- BiCGSTAB (Eigen 3.4 BiCGSTAB.h)
// Published algorithm restated: x0 = 0, identity preconditioner, restart when rho is too small,
// stop on the recursive residual; dot(a,b) = sum conj(a) b. hist receives ||r_i|| / ||b|| for
// i = 0..iters (the entry criterion value before each iteration and the final one).
template <class Op>
static int bicgstab(int N, Op A, const cd *b, cd *x, double tol, int maxit, std::vector<double> &hist, double &err) {
    auto dotc = [N](const std::vector<cd> &a, const std::vector<cd> &c) {
        cd s = 0;
        for (int i = 0; i < N; i++) s += std::conj(a[i]) * c[i];
        return s;
    };
    auto sq = [N](const std::vector<cd> &a) {
        double s = 0;
        for (int i = 0; i < N; i++) s += std::norm(a[i]);
        return s;
    };
    std::vector<cd> xv(N, 0.0), r(N), r0, v(N, 0.0), p(N, 0.0), s(N), t(N), Ax(N);
    A(xv.data(), Ax.data());  // r = b - A x with x = 0 (Eigen evaluates the product)
    for (int i = 0; i < N; i++) r[i] = b[i] - Ax[i];
    r0 = r;
    double r0_sq = sq(r0);
    double rhs_sq = 0;
    for (int i = 0; i < N; i++) rhs_sq += std::norm(b[i]);
    hist.clear();
    if (rhs_sq == 0) {
        for (int i = 0; i < N; i++) x[i] = 0;
        err = 0;
        return 0;
    }
    cd rho = 1, alpha = 1, w = 1;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = std::numeric_limits<double>::epsilon();
    const double eps2 = eps * eps;
    int i = 0, restarts = 0;
    double rn = sq(r);
    hist.push_back(std::sqrt(rn / rhs_sq));
    while (rn > tol2 && i < maxit) {
        cd rho_old = rho;
        rho = dotc(r0, r);
        if (std::abs(rho) < eps2 * r0_sq) {
            A(xv.data(), Ax.data());
            for (int q = 0; q < N; q++) r[q] = b[q] - Ax[q];
            r0 = r;
            rho = r0_sq = sq(r);
            if (restarts++ == 0) i = 0;
        }
        cd beta = (rho / rho_old) * (alpha / w);
        for (int q = 0; q < N; q++) p[q] = r[q] + beta * (p[q] - w * v[q]);
        A(p.data(), v.data());
        alpha = rho / dotc(r0, v);
        for (int q = 0; q < N; q++) s[q] = r[q] - alpha * v[q];
        A(s.data(), t.data());
        double tmp = sq(t);
        if (tmp > 0)
            w = dotc(t, s) / tmp;
        else
            w = 0;
        for (int q = 0; q < N; q++) {
            xv[q] += alpha * p[q] + w * s[q];
            r[q] = s[q] - w * t[q];
        }
        ++i;
        rn = sq(r);
        hist.push_back(std::sqrt(rn / rhs_sq));
    }
    err = std::sqrt(rn / rhs_sq);
    for (int q = 0; q < N; q++) x[q] = xv[q];
    return i;
}

// Dense LU with partial pivoting (stands in for Eigen's colPivHouseholderQr, adjoint.cpp:74).
static void dense_solve(int N, std::vector<cd> A /* col-major */, const cd *b, cd *x) {
    std::vector<int> piv(N);
    for (int c = 0; c < N; c++) {
        int pr = c;
        double best = std::abs(A[(size_t)c * N + c]);
        for (int r = c + 1; r < N; r++)
            if (std::abs(A[(size_t)c * N + r]) > best) {
                best = std::abs(A[(size_t)c * N + r]);
                pr = r;
            }
        piv[c] = pr;
        if (pr != c)
            for (int cc = 0; cc < N; cc++) std::swap(A[(size_t)cc * N + c], A[(size_t)cc * N + pr]);
        cd d = A[(size_t)c * N + c];
        for (int r = c + 1; r < N; r++) A[(size_t)c * N + r] /= d;
#pragma omp parallel for schedule(static)
        for (int cc = c + 1; cc < N; cc++) {
            cd f = A[(size_t)cc * N + c];
            if (f == 0.0) continue;
            for (int r = c + 1; r < N; r++) A[(size_t)cc * N + r] -= A[(size_t)c * N + r] * f;
        }
    }
    std::vector<cd> y(b, b + N);
    for (int c = 0; c < N; c++) std::swap(y[c], y[piv[c]]);
    for (int c = 0; c < N; c++)
        for (int r = c + 1; r < N; r++) y[r] -= A[(size_t)c * N + r] * y[c];
    for (int c = N - 1; c >= 0; c--) {
        y[c] /= A[(size_t)c * N + c];
        for (int r = 0; r < c; r++) y[r] -= A[(size_t)c * N + r] * y[c];
    }
    for (int i = 0; i < N; i++) x[i] = y[i];
}

// ---------------------------------------------------------------- shape derivatives
// Derived here from the kernel (not transcribed): with d = p - q, r = |d|, s = -(d . n),
// K = F(r) s, F(r) = e^{-ikr}(-ikr-1) / (4 pi r^3), F'(r) = e^{-ikr}(-k^2/r^2 + 3(1+ikr)/r^4) / (4 pi).
// (a) collocation element moves (p = centroid, dp/dh = e_y/3):
//        dK/dh = (F'(r) d_y s / r - F(r) n_y) / 3                     [restates dMij_dpy]
// (b) vertex c of the integration element moves (dq/dh = beta_c e_y, beta = (1-xi-eta, xi, eta),
//     n = ab x ac, dn/dh = e_y x ac | ab x e_y | e_y x (b - c) for c = b | c | a):
//        dK/dh = -F'(r) beta_c d_y s / r + F(r)(beta_c n_y - d . dn/dh)   [restates dMij_dq{1,2,3}y]
// derivatives.cpp:5-131 holds the same derivatives as sympy CSE output; the quadrature wrappers
// (derivatives.cpp:147-177) are 1/2 sum_g w_g (.), as here.
static inline void FdF(double k, double r, cd &F, cd &dF) {
    cd e = std::exp(cd(0.0, -k * r));
    double r2 = r * r, r3 = r2 * r, r4 = r2 * r2;
    F = e * cd(-1.0, -k * r) / (4 * kPi * r3);
    dF = e * (cd(-k * k / r2, 0.0) + cd(3.0, 3.0 * k * r) / r4) / (4 * kPi);
}

// dM_ij / dh for h = y of any vertex of collocation element i (p = centroid of i).
static cd dM_dp(double k, V3 p, V3 qa, V3 qb, V3 qc) {
    V3 ab = qb - qa, ac = qc - qa;
    V3 n = {ac.z * ab.y - ac.y * ab.z, ac.x * ab.z - ac.z * ab.x, ac.y * ab.x - ac.x * ab.y};
    cd res = 0;
    for (int g = 0; g < 6; g++) {
        V3 q = qa + kQuad[g][1] * ab + kQuad[g][2] * ac;
        V3 d = p - q;
        double r = norm(d), s = -dot(d, n);
        cd F, dF;
        FdF(k, r, F, dF);
        res += kQuad[g][0] * ((dF * (d.y * s / r) - F * n.y) / 3.0);
    }
    return 0.5 * res;
}

// dM_ij / dh for h = y of vertex `corner` (0,1,2 = a,b,c) of integration element j.
static cd dM_dq(double k, V3 p, V3 qa, V3 qb, V3 qc, int corner) {
    V3 ab = qb - qa, ac = qc - qa;
    V3 n = {ac.z * ab.y - ac.y * ab.z, ac.x * ab.z - ac.z * ab.x, ac.y * ab.x - ac.x * ab.y};
    V3 dn;
    if (corner == 0) {
        V3 bc = qb - qc;
        dn = {bc.z, 0.0, -bc.x};
    } else if (corner == 1) {
        dn = {ac.z, 0.0, -ac.x};
    } else {
        dn = {-ab.z, 0.0, ab.x};
    }
    cd res = 0;
    for (int g = 0; g < 6; g++) {
        double xi = kQuad[g][1], eta = kQuad[g][2];
        double beta = corner == 0 ? 1.0 - xi - eta : (corner == 1 ? xi : eta);
        V3 q = qa + xi * ab + eta * ac;
        V3 d = p - q;
        double r = norm(d), s = -dot(d, n);
        cd F, dF;
        FdF(k, r, F, dF);
        res += kQuad[g][0] * (-dF * (beta * d.y * s / r) + F * (beta * n.y - dot(d, dn)));
    }
    return 0.5 * res;
}

// helpers.cpp:116-120: b = e^{-ikr}/(4 pi r), r = |c - src|.
static cd point_source(double k, V3 c, V3 src) {
    double r = norm(c - src);
    return std::exp(cd(0.0, -k * r)) / (4 * kPi * r);
}
// derivatives.cpp:133-145 (dbi_dpy): d b_i / dh for any vertex of element i (centroid moves 1/3).
static cd dpoint_source(double k, V3 c, V3 src) {
    V3 d = c - src;
    double r = norm(d);
    cd e = std::exp(cd(0.0, -k * r));
    return e * cd(-1.0, -k * r) / (4 * kPi * r * r) * (d.y / (3.0 * r));
}
// sphere.h:22-26 generalised to direction dir: b = e^{-ik dir . c}; db/dh = -ik dir_y / 3 * b.
static cd plane_wave(double k, V3 c, V3 dir) { return std::exp(cd(0.0, -k * dot(dir, c))); }

// adjoint.cpp:388-397: d = ((sum L)^2 - sum L^2) / ((LL-1) sum L^2), u = dd/dyt.
static double diffusion(int LL, const double *yt, double *u) {
    double S = 0, Q = 0;
    for (int l = 0; l < LL; l++) {
        double L = yt[l] * yt[l];
        S += L;
        Q += L * L;
    }
    double c = (S * S - Q) / ((LL - 1) * Q);
    if (u)
        for (int l = 0; l < LL; l++) {
            double L = yt[l] * yt[l];
            u[l] = 4 * (S * Q - L * S * S) / ((LL - 1) * Q * Q) * yt[l];
        }
    return c;
}

// The state the C API hands out: geometry + partition (frozen at set-up) + optional H-matrix.
struct Ctx {
    std::vector<double> P;
    std::vector<int> E;
    Geo g;
    HMat H;
    double k = 0;
    bool assembled = false;
    void refresh() {
        g.m = Mesh{P.data(), E.data(), (int)P.size() / 3, (int)E.size() / 3};
        intermediates(g.m, g.C, g.Nrm);
    }
};

static thread_local std::string g_err;

}  // namespace orc

using namespace orc;

template <class F>
static int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#endif
}

// helpers.cpp:16-32
int orc_intermediates(const double *P, int V, const int *E, int N, double *C, double *Nrm) {
    return guarded([&] {
        Mesh m{P, E, V, N};
        std::vector<V3> c, n;
        intermediates(m, c, n);
        for (int e = 0; e < N; e++) {
            C[3 * e] = c[e].x, C[3 * e + 1] = c[e].y, C[3 * e + 2] = c[e].z;
            Nrm[3 * e] = n[e].x, Nrm[3 * e + 1] = n[e].y, Nrm[3 * e + 2] = n[e].z;
        }
    });
}

// One entry M(p, triangle) (integrate.cpp:5-14). tri = 9 doubles (a, b, c); n = ab x ac.
int orc_entry(double k, const double *p, const double *tri, double *out) {
    return guarded([&] {
        V3 a{tri[0], tri[1], tri[2]}, b{tri[3], tri[4], tri[5]}, c{tri[6], tri[7], tri[8]};
        V3 ab = b - a, ac = c - a;
        V3 n = {ac.z * ab.y - ac.y * ab.z, ac.x * ab.z - ac.z * ab.x, ac.y * ab.x - ac.x * ab.y};
        cd v = entry(k, {p[0], p[1], p[2]}, a, b, c, n);
        out[0] = v.real(), out[1] = v.imag();
    });
}

// Derivatives of one entry: which = 0 -> collocation element moves, 1..3 -> corner a,b,c of
// the integration element moves (all in y).
int orc_dentry(double k, const double *p3 /* 9: collocation triangle */, const double *tri, int which, double *out) {
    return guarded([&] {
        V3 p1{p3[0], p3[1], p3[2]}, p2{p3[3], p3[4], p3[5]}, p3v{p3[6], p3[7], p3[8]};
        V3 p{(p1.x + p2.x + p3v.x) / 3, (p1.y + p2.y + p3v.y) / 3, (p1.z + p2.z + p3v.z) / 3};
        V3 a{tri[0], tri[1], tri[2]}, b{tri[3], tri[4], tri[5]}, c{tri[6], tri[7], tri[8]};
        cd v = which == 0 ? dM_dp(k, p, a, b, c) : dM_dq(k, p, a, b, c, which - 1);
        out[0] = v.real(), out[1] = v.imag();
    });
}

// Cluster tree + block partition (clustertree.cpp:49-137). perm[N] out; blocks out as
// (rs,re,cs,ce) quadruples into dir_out / app_out (capacity cap quadruples each);
// leaves out as (s,e) pairs in order (capacity N).
int orc_partition(const double *C, int N, int leaf, double factor, int *perm, int *dir_out, int *nd, int *app_out,
                  int *na, int cap, int *leaves_out, int *nleaves) {
    return guarded([&] {
        std::vector<V3> Cv(N);
        for (int e = 0; e < N; e++) Cv[e] = {C[3 * e], C[3 * e + 1], C[3 * e + 2]};
        Tree t;
        t.max_pts = leaf;
        t.perm.assign(N, 0);
        std::vector<int> idx(N);
        std::iota(idx.begin(), idx.end(), 0);
        t.build(0, idx, Cv);
        // The reference traverses with the centroids in reordered numbering; node geometry is
        // identical because nodes hold sets, not positions.
        std::vector<int> dir, app;
        t.traverse(0, 0, factor, dir, app);
        if ((int)dir.size() / 4 > cap || (int)app.size() / 4 > cap) throw std::runtime_error("partition: capacity");
        std::copy(t.perm.begin(), t.perm.end(), perm);
        std::copy(dir.begin(), dir.end(), dir_out);
        std::copy(app.begin(), app.end(), app_out);
        *nd = (int)dir.size() / 4;
        *na = (int)app.size() / 4;
        int nl = 0;
        for (auto &n : t.nodes)
            if (n.leaf) {
                leaves_out[2 * nl] = n.s;
                leaves_out[2 * nl + 1] = n.e;
                nl++;
            }
        *nleaves = nl;
    });
}

// Tree check (clustertree.cpp:33-47) on the reordered centroids: every node's members lie in
// its ball (+1e-6). Returns 1 if ok.
int orc_tree_check(const double *C, int N, int leaf) {
    std::vector<V3> Cv(N);
    for (int e = 0; e < N; e++) Cv[e] = {C[3 * e], C[3 * e + 1], C[3 * e + 2]};
    Tree t;
    t.max_pts = leaf;
    t.perm.assign(N, 0);
    std::vector<int> idx(N);
    std::iota(idx.begin(), idx.end(), 0);
    try {
        t.build(0, idx, Cv);
    } catch (...) {
        return 0;
    }
    std::vector<V3> Cr(N);
    for (int i = 0; i < N; i++) Cr[i] = Cv[t.perm[i]];
    for (auto &n : t.nodes)
        for (int i = n.s; i < n.e; i++)
            if (norm(Cr[i] - n.center) > n.radius + 1e-6) return 0;
    return 1;
}

// A context: mesh (already reordered) + partition.
void *orc_ctx_create(const double *P, int V, const int *E, int N, const int *dir, int nd, const int *app, int na) {
    auto *c = new Ctx();
    c->P.assign(P, P + 3 * V);
    c->E.assign(E, E + 3 * N);
    c->refresh();
    c->H.N = N;
    c->H.dir.assign(dir, dir + 4 * nd);
    c->H.app.assign(app, app + 4 * na);
    return c;
}
void orc_ctx_destroy(void *h) { delete (Ctx *)h; }

int orc_set_vertices(void *h, const double *P) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        std::copy(P, P + c->P.size(), c->P.begin());
        c->refresh();
        c->assembled = false;
    });
}

// hmatrix.cpp:10-85
int orc_assemble(void *h, double k, double aca_tol, int max_rank) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        hmat_assemble(c->H, c->g, k, aca_tol, max_rank);
        c->k = k;
        c->assembled = true;
    });
}

int orc_ranks(void *h, int *out) {
    auto *c = (Ctx *)h;
    return guarded([&] { std::copy(c->H.rank.begin(), c->H.rank.end(), out); });
}

// Low-rank factors of approx block a: U (m x r col-major) and V (r x n row-major), interleaved complex.
int orc_factors(void *h, int a, double *U, double *V) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        std::memcpy(U, c->H.U[a].data(), c->H.U[a].size() * sizeof(cd));
        std::memcpy(V, c->H.Vr[a].data(), c->H.Vr[a].size() * sizeof(cd));
    });
}

int orc_dense_block(void *h, int b, double *out) {
    auto *c = (Ctx *)h;
    return guarded([&] { std::memcpy(out, c->H.D[b].data(), c->H.D[b].size() * sizeof(cd)); });
}

int orc_matvec(void *h, int transpose, const double *x, double *y) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        if (!c->assembled) throw std::runtime_error("not assembled");
        hmat_matvec(c->H, transpose != 0, (const cd *)x, (cd *)y);
    });
}

// BiCGSTAB on the assembled H-matrix (adjoint.cpp:92-95,206-209). hist capacity maxit+1.
int orc_bicgstab(void *h, int transpose, const double *b, double *x, double tol, int maxit, double *hist, int *iters,
                 double *err) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        if (!c->assembled) throw std::runtime_error("not assembled");
        int N = c->H.N;
        if (maxit < 0) maxit = 2 * N;  // Eigen default: 2 * cols
        std::vector<double> hs;
        auto op = [&](const cd *xx, cd *yy) { hmat_matvec(c->H, transpose != 0, xx, yy); };
        *iters = bicgstab(N, op, (const cd *)b, (cd *)x, tol, maxit, hs, *err);
        if (hist) std::copy(hs.begin(), hs.end(), hist);
    });
}

// The block ACA (aca.cpp:38-82; guard = 1) or the listener ACA's unguarded pivot division
// (aca.cpp:104; guard = 0) on an explicit m x n matrix A (column-major, interleaved complex), so
// the cross sequence can be checked on matrices whose crosses are known by hand. U (m x lim
// column-major) and V (lim x n row-major) with lim = min(max_iter, m - 1, n - 1) are written
// when non-null; *rank receives the number of crosses kept.
int orc_aca_dense(int m, int n, const double *A, double eps, int max_iter, int guard, double *U, double *V,
                  int *rank) {
    return guarded([&] {
        const cd *Am = (const cd *)A;
        auto row = [&](int i, cd *out) {
            for (int c = 0; c < n; c++) out[c] = Am[(size_t)c * m + i];
        };
        auto col = [&](int j, cd *out) {
            for (int r = 0; r < m; r++) out[r] = Am[(size_t)j * m + r];
        };
        std::vector<cd> Uv, Vv;
        *rank = aca(m, n, eps, max_iter, row, col, guard != 0, Uv, Vv);
        if (U) std::memcpy(U, Uv.data(), Uv.size() * sizeof(cd));
        if (V) std::memcpy(V, Vv.data(), Vv.size() * sizeof(cd));
    });
}

// BiCGSTAB (the restatement used on the H-matrix) on an explicit N x N matrix A (column-major,
// interleaved complex), so finite termination and hand-traced recurrences can pin it.
// maxit < 0: 2 N. hist capacity maxit + 2.
int orc_bicgstab_dense(int N, const double *A, const double *b, double *x, double tol, int maxit, double *hist,
                       int *iters, double *err) {
    return guarded([&] {
        const cd *Am = (const cd *)A;
        if (maxit < 0) maxit = 2 * N;
        auto op = [&](const cd *xx, cd *yy) {
            for (int i = 0; i < N; i++) yy[i] = 0;
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) yy[i] += Am[(size_t)j * N + i] * xx[j];
        };
        std::vector<double> hs;
        *iters = bicgstab(N, op, (const cd *)b, (cd *)x, tol, maxit, hs, *err);
        if (hist) std::copy(hs.begin(), hs.end(), hist);
    });
}

// Dense boundary matrix M - I/2 (integrate.cpp:32-44), column-major, interleaved complex.
int orc_dense_matrix(void *h, double k, double *out) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        int N = c->g.m.N;
        cd *A = (cd *)out;
#pragma omp parallel for schedule(dynamic)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) A[(size_t)j * N + i] = c->g.M(k, c->g.C[i], j) - (i == j ? 0.5 : 0.0);
    });
}

int orc_dense_solve(int N, const double *A, const double *b, double *x) {
    return guarded([&] {
        std::vector<cd> Am((const cd *)A, (const cd *)A + (size_t)N * N);
        dense_solve(N, Am, (const cd *)b, (cd *)x);
    });
}

// Listener matrix Q (LL x N, column-major) (integrate.cpp:46-57).
int orc_listener_matrix(void *h, double k, const double *L, int LL, double *out) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        int N = c->g.m.N;
        cd *Q = (cd *)out;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; j++)
            for (int l = 0; l < LL; l++) Q[(size_t)j * LL + l] = c->g.M(k, {L[3 * l], L[3 * l + 1], L[3 * l + 2]}, j);
    });
}

// Listener ACA (aca.cpp:84-134): returns the rank; U (LL x r col-major), V (r x N row-major)
// are written when non-null (capacity LL*maxr, maxr*N).
int orc_aca_listener(void *h, double k, const double *L, int LL, double eps, int max_rank, double *U, double *V,
                     int *rank) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        int N = c->g.m.N;
        auto row = [&](int i, cd *out) {
            V3 p{L[3 * i], L[3 * i + 1], L[3 * i + 2]};
#pragma omp parallel for schedule(static)
            for (int j = 0; j < N; j++) out[j] = c->g.M(k, p, j);
        };
        auto col = [&](int j, cd *out) {
            for (int l = 0; l < LL; l++) out[l] = c->g.M(k, {L[3 * l], L[3 * l + 1], L[3 * l + 2]}, j);
        };
        std::vector<cd> Uv, Vv;
        int r = aca(LL, N, eps, max_rank, row, col, false, Uv, Vv);
        *rank = r;
        if (U) std::memcpy(U, Uv.data(), Uv.size() * sizeof(cd));
        if (V) std::memcpy(V, Vv.data(), Vv.size() * sizeof(cd));
    });
}

// Right-hand sides. kind 0: point source at src (helpers.cpp:46-50); kind 1: plane wave along
// dir (sphere.h:22-26, there dir = +x).
int orc_rhs(void *h, double k, int kind, const double *v, double *out) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        int N = c->g.m.N;
        V3 s{v[0], v[1], v[2]};
        for (int i = 0; i < N; i++) {
            cd b = kind == 0 ? point_source(k, c->g.C[i], s) : plane_wave(k, c->g.C[i], s);
            out[2 * i] = b.real(), out[2 * i + 1] = b.imag();
        }
    });
}

double orc_diffusion(int LL, const double *yt, double *u) { return diffusion(LL, yt, u); }

// Shape gradient of one (frequency, incidence) (adjoint.cpp:219-231, 244-386) restructured
// per element. diff[HH] are vertex ids; x, lam (N) and vf (LL) interleaved complex.
// src_kind/src as orc_rhs. Writes A, B, C (HH each); dcdh = A - B + C.
int orc_shape_grad(void *h, double k, const double *L, int LL, const int *diff, int HH, const double *x,
                   const double *lam, const double *vf, int src_kind, const double *src, double *A, double *B,
                   double *Cc) {
    auto *c = (Ctx *)h;
    return guarded([&] {
        const Geo &g = c->g;
        int N = g.m.N;
        const cd *X = (const cd *)x, *Lm = (const cd *)lam, *VF = (const cd *)vf;
        std::vector<int> slot(g.m.V, -1);
        for (int t = 0; t < HH; t++) slot[diff[t]] = t;
        // elements touching a diff vertex (set_diff_points, acoustics3d.cpp:69-101)
        std::vector<int> touch;
        for (int e = 0; e < N; e++)
            for (int q = 0; q < 3; q++)
                if (slot[g.m.E[3 * e + q]] >= 0) {
                    touch.push_back(e);
                    break;
                }
        const int T = (int)touch.size();
        std::vector<cd> S1(T), S2(3 * T), SA(3 * T);
        V3 s{src[0], src[1], src[2]};
#pragma omp parallel for schedule(dynamic)
        for (int t = 0; t < T; t++) {
            int i = touch[t];
            V3 pa = g.m.a(i), pb = g.m.b(i), pc = g.m.c(i);
            cd s1 = 0, s2[3] = {0, 0, 0}, sa[3] = {0, 0, 0};
            for (int j = 0; j < N; j++) {
                s1 += X[j] * dM_dp(k, g.C[i], g.m.a(j), g.m.b(j), g.m.c(j));           // adjoint.cpp:321
                for (int q = 0; q < 3; q++) s2[q] += Lm[j] * dM_dq(k, g.C[j], pa, pb, pc, q);  // :329-333
            }
            for (int l = 0; l < LL; l++) {  // adjoint.cpp:268-276
                V3 p{L[3 * l], L[3 * l + 1], L[3 * l + 2]};
                for (int q = 0; q < 3; q++) sa[q] += VF[l] * dM_dq(k, p, pa, pb, pc, q);
            }
            S1[t] = s1;
            for (int q = 0; q < 3; q++) S2[3 * t + q] = s2[q], SA[3 * t + q] = sa[q];
        }
        for (int t = 0; t < HH; t++) A[t] = B[t] = Cc[t] = 0;
        for (int t = 0; t < T; t++) {
            int i = touch[t];
            cd db = src_kind == 0 ? dpoint_source(k, g.C[i], s) : cd(0.0, -k * s.y / 3.0) * plane_wave(k, g.C[i], s);
            for (int q = 0; q < 3; q++) {
                int h2 = slot[g.m.E[3 * i + q]];
                if (h2 < 0) continue;
                A[h2] += (SA[3 * t + q] * X[i]).real();                              // adjoint.cpp:277
                B[h2] += (S1[t] * Lm[i]).real() + (S2[3 * t + q] * X[i]).real();    // adjoint.cpp:339
                Cc[h2] -= (db * Lm[i]).real();                                       // adjoint.cpp:373
            }
        }
    });
}

}  // extern "C"
