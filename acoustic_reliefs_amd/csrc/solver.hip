// Batched BiCGSTAB on the H-matrix, one independent iteration per right-hand side.
//
// Per rhs it follows Eigen 3.4's BiCGSTAB (the solver the reference instantiates as
// Eigen::BiCGSTAB<HMatrix, Eigen::IdentityPreconditioner>, adjoint.cpp:92-95):
//   x = 0, r = b, r0 = r, rho = alpha = w = 1, v = p = 0
//   while ||r||^2 > tol^2 ||b||^2 and i < maxit:
//     rho_old = rho; rho = r0^H r
//     if |rho| < eps^2 ||r0||^2: r = b - A x; r0 = r; rho = ||r0||^2 (first restart resets i)
//     beta = (rho / rho_old)(alpha / w); p = r + beta (p - w v); v = A p
//     alpha = rho / (r0^H v); s = r - alpha v; t = A s
//     w = t^H s / ||t||^2 (0 if ||t|| = 0); x += alpha p + w s; r = s - w t; i++
// The reductions are two-stage with a fixed split, so every rhs gets the same bits whatever
// the batch. The loop control (convergence, maxit, breakdown restarts) runs on the device in
// k_finish, so the host launches K iterations between synchronisations, K chosen from the
// observed convergence rate; a right-hand side that meets a restart pauses until the host
// restarts it at the next synchronisation, which leaves its arithmetic unchanged.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstring>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

// ARH_TRACE=1 in the environment: one line per solver sweep on stderr (precision, right-hand sides
// in the sweep, iterations, true residuals) for diagnosing refinement on hard problems.
static bool trace_on() {
    static const bool on = [] {
        const char *e = std::getenv("ARH_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}

static void trace_sweep(const char *what, int sweep, const std::vector<int> &in, const std::vector<int> &iters,
                        const std::vector<double> &tres) {
    if (!trace_on()) return;
    int n = 0, imax = 0;
    double rmax = 0.0, rmin = 1e300;
    for (size_t j = 0; j < in.size(); j++) {
        if (!in[j]) continue;
        n++;
        imax = std::max(imax, j < iters.size() ? iters[j] : 0);
    }
    for (double r : tres) rmax = std::max(rmax, r), rmin = std::min(rmin, r);
    std::fprintf(stderr, "[arh solve] %s sweep %d: %d rhs, max iters %d, true residual %.3e .. %.3e\n", what, sweep, n,
                 imax, tres.empty() ? 0.0 : rmin, rmax);
}

struct RhsState {
    double2 rho;      // rho of the previous iteration (Eigen's rho_old at the next step)
    double2 rho_new;  // r0^H r for the coming iteration
    double2 alpha, w;
    double r0sq, rhs_sq, rn, tol2;  // tol2 = tol^2 ||b||^2
    int active;    // iterating
    int restart;   // restart pending (set at init, or by k_finish on breakdown: then paused)
    int it;        // iterations counted toward maxit (reset at the first restart)
    int iters;     // iterations done
    int nrestart;  // restarts done
    int maxit;
    int pad[2];
};

constexpr int kVT = 256;      // threads per block of vector kernels
constexpr double kInnerFloor = 1e-7;  // mixed refinement: tightest inner (FP32-operator) tolerance
#ifndef ARH_INNER_FRAC
#define ARH_INNER_FRAC 0.5
#endif
constexpr double kInnerFrac = ARH_INNER_FRAC;  // mixed refinement: inner target as a fraction of the remaining one
constexpr int kInnerBudget = 256;      // mixed refinement: iterations an inner solve may take
#ifndef ARH_SOLVER_MAXK
#define ARH_SOLVER_MAXK 8
#endif
constexpr int kMaxK = ARH_SOLVER_MAXK;  // iterations per host synchronisation, at most
constexpr int kMaxNB = 128;   // blocks per rhs of a reduction

__device__ __forceinline__ void block_reduce4(double v[4], double *sh) {
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; k++) sh[w * 4 + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) {
            double s = sh[k];
            for (int q = 1; q < kVT / ARH_WAVE; q++) s += sh[q * 4 + k];
            v[k] = s;
        }
}

// fixed-order sum of NB partial quadruples
__device__ __forceinline__ void sum_parts(const double *part, int NB, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    for (int b = 0; b < NB; b++)
#pragma unroll
        for (int k = 0; k < 4; k++) out[k] += part[b * 4 + k];
}

struct Chunk {
    int lo, hi;
};
__device__ __forceinline__ Chunk chunk_of(int N, int NB) {
    int len = (N + NB - 1) / NB;
    int lo = blockIdx.x * len;
    return Chunk{lo, min(N, lo + len)};
}

// partial sums of conj(a) b (complex) and |c|^2 over the block's chunk
__global__ __launch_bounds__(kVT) void k_dot(const RhsState *__restrict__ S, int N, int NB,
                                             const double2 *__restrict__ a, const double2 *__restrict__ b,
                                             const double2 *__restrict__ c, double *__restrict__ part) {
    __shared__ double sh[4 * kVT / ARH_WAVE];
    const int j = blockIdx.y;
    if (!S[j].active) return;
    const Chunk ch = chunk_of(N, NB);
    const int64_t o = (int64_t)j * N;
    double v[4] = {0, 0, 0, 0};
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += kVT) {
        double2 ai = a[o + i], bi = b[o + i];
        v[0] += ai.x * bi.x + ai.y * bi.y;
        v[1] += ai.x * bi.y - ai.y * bi.x;
        if (c) {
            double2 ci = c[o + i];
            v[2] += ci.x * ci.x + ci.y * ci.y;
        }
    }
    block_reduce4(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[((int64_t)j * NB + blockIdx.x) * 4 + k] = v[k];
}

// p = r + beta (p - w v), beta = (rho_new / rho)(alpha / w)
__global__ __launch_bounds__(kVT) void k_p_update(const RhsState *__restrict__ S, int N, int NB,
                                                  const double2 *__restrict__ r, double2 *__restrict__ p,
                                                  const double2 *__restrict__ v) {
    const int j = blockIdx.y;
    const RhsState st = S[j];
    if (!st.active) return;
    const double2 beta = cmul(cdiv(st.rho_new, st.rho), cdiv(st.alpha, st.w));
    const Chunk ch = chunk_of(N, NB);
    const int64_t o = (int64_t)j * N;
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += kVT) {
        double2 pi = p[o + i], vi = v[o + i], ri = r[o + i];
        double2 q = csub(pi, cmul(st.w, vi));
        p[o + i] = cadd(ri, cmul(beta, q));
    }
}

// s = r - alpha v, alpha = rho_new / (r0^H v) from the partials
__global__ __launch_bounds__(kVT) void k_s_update(const RhsState *__restrict__ S, int N, int NB,
                                                  const double *__restrict__ part_r0v, const double2 *__restrict__ r,
                                                  const double2 *__restrict__ v, double2 *__restrict__ s) {
    const int j = blockIdx.y;
    const RhsState st = S[j];
    if (!st.active) return;
    double d[4];
    sum_parts(part_r0v + (int64_t)j * NB * 4, NB, d);
    const double2 alpha = cdiv(st.rho_new, make_double2(d[0], d[1]));
    const Chunk ch = chunk_of(N, NB);
    const int64_t o = (int64_t)j * N;
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += kVT) s[o + i] = csub(r[o + i], cmul(alpha, v[o + i]));
}

// x += alpha p + w s; r = s - w t; partials of r0^H r and |r|^2
__global__ __launch_bounds__(kVT) void k_xr_update(const RhsState *__restrict__ S, int N, int NB,
                                                   const double *__restrict__ part_r0v,
                                                   const double *__restrict__ part_ts, const double2 *__restrict__ p,
                                                   const double2 *__restrict__ s, const double2 *__restrict__ t,
                                                   const double2 *__restrict__ r0, double2 *__restrict__ x,
                                                   double2 *__restrict__ r, double *__restrict__ part_rr) {
    __shared__ double sh[4 * kVT / ARH_WAVE];
    const int j = blockIdx.y;
    const RhsState st = S[j];
    if (!st.active) return;
    double d[4], e[4];
    sum_parts(part_r0v + (int64_t)j * NB * 4, NB, d);
    sum_parts(part_ts + (int64_t)j * NB * 4, NB, e);
    const double2 alpha = cdiv(st.rho_new, make_double2(d[0], d[1]));
    const double tt = e[2];
    const double2 w = tt > 0.0 ? make_double2(e[0] / tt, e[1] / tt) : make_double2(0.0, 0.0);
    const Chunk ch = chunk_of(N, NB);
    const int64_t o = (int64_t)j * N;
    double v[4] = {0, 0, 0, 0};
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += kVT) {
        double2 si = s[o + i], ti = t[o + i];
        double2 xi = x[o + i];
        xi = cadd(xi, cadd(cmul(alpha, p[o + i]), cmul(w, si)));
        x[o + i] = xi;
        double2 ri = csub(si, cmul(w, ti));
        r[o + i] = ri;
        double2 qi = r0[o + i];
        v[0] += qi.x * ri.x + qi.y * ri.y;
        v[1] += qi.x * ri.y - qi.y * ri.x;
        v[2] += ri.x * ri.x + ri.y * ri.y;
    }
    block_reduce4(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part_rr[((int64_t)j * NB + blockIdx.x) * 4 + k] = v[k];
}

// end of an iteration: commit alpha, w, rho; rho_new = r0^H r, rn = |r|^2; then Eigen's loop
// control for the next iteration: stop when ||r||^2 <= tol2 or it == maxit, pause for a restart
// when |r0^H r| < eps^2 ||r0||^2. Optional history of ||r|| / ||b||. One workgroup; gate[0]
// receives the number of right-hand sides still active.
__global__ __launch_bounds__(256) void k_finish(RhsState *__restrict__ S, int R, int NB,
                                                const double *__restrict__ part_r0v,
                                                const double *__restrict__ part_ts, const double *__restrict__ part_rr,
                                                double *__restrict__ hist, int hist_cap, int *__restrict__ gate,
                                                int *__restrict__ gate_hist) {
    __shared__ int cnt[256 / ARH_WAVE];
    const double eps2 = 4.930380657631324e-32;  // (2^-52)^2, Eigen's NumTraits<double>::epsilon() squared
    int nact = 0;
    for (int j = threadIdx.x; j < R; j += blockDim.x) {
        RhsState st = S[j];
        if (!st.active) continue;
        double d[4], e[4], f[4];
        sum_parts(part_r0v + (int64_t)j * NB * 4, NB, d);
        sum_parts(part_ts + (int64_t)j * NB * 4, NB, e);
        sum_parts(part_rr + (int64_t)j * NB * 4, NB, f);
        st.alpha = cdiv(st.rho_new, make_double2(d[0], d[1]));
        st.w = e[2] > 0.0 ? make_double2(e[0] / e[2], e[1] / e[2]) : make_double2(0.0, 0.0);
        st.rho = st.rho_new;
        st.rho_new = make_double2(f[0], f[1]);
        st.rn = f[2];
        st.it++;
        st.iters++;
        if (hist && st.iters < hist_cap) hist[(int64_t)j * hist_cap + st.iters] = sqrt(st.rn / st.rhs_sq);
        if (!(st.rn > st.tol2 && st.it < st.maxit)) {
            st.active = 0;
        } else if (hypot(st.rho_new.x, st.rho_new.y) < eps2 * st.r0sq) {
            st.active = 0;
            st.restart = 1;
        } else {
            nact++;
        }
        S[j] = st;
    }
    // integer count: the same whatever the order of the additions
    for (int off = ARH_WAVE / 2; off > 0; off >>= 1) nact += __shfl_xor(nact, off, ARH_WAVE);
    if (threadIdx.x % ARH_WAVE == 0) cnt[threadIdx.x / ARH_WAVE] = nact;
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int k = 0; k < (int)blockDim.x / ARH_WAVE; k++) c += cnt[k];
        gate[0] = c;
        gate_hist[0] = c;
    }
}

// (re)start of rhs j: r = b - Ax (Ax given, or zero), r0 = r, partials of |r|^2 (and b^2 on init)
__global__ __launch_bounds__(kVT) void k_restart(const RhsState *__restrict__ S, int N, int NB,
                                                 const double2 *__restrict__ b, const double2 *__restrict__ Ax,
                                                 double2 *__restrict__ r, double2 *__restrict__ r0,
                                                 double *__restrict__ part) {
    __shared__ double sh[4 * kVT / ARH_WAVE];
    const int j = blockIdx.y;
    if (!S[j].restart) return;
    const Chunk ch = chunk_of(N, NB);
    const int64_t o = (int64_t)j * N;
    double v[4] = {0, 0, 0, 0};
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += kVT) {
        double2 bi = b[o + i];
        double2 ri = Ax ? csub(bi, Ax[o + i]) : bi;
        r[o + i] = ri;
        r0[o + i] = ri;
        v[0] += ri.x * ri.x + ri.y * ri.y;
        v[1] += bi.x * bi.x + bi.y * bi.y;
    }
    block_reduce4(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[((int64_t)j * NB + blockIdx.x) * 4 + k] = v[k];
}

// full restart: rho = r0sq = |r|^2; fresh: also alpha = w = rho_old = 1, rhs_sq, and the caller
// zeroes p, v (the host decides activity after a fresh start); a breakdown restart resumes the
// iteration (active again) and resets the iteration count on the first restart (Eigen)
__global__ void k_restart_finish(RhsState *__restrict__ S, int R, int NB, const double *__restrict__ part,
                                 int fresh) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= R || !S[j].restart) return;
    RhsState st = S[j];
    double f[4];
    sum_parts(part + (int64_t)j * NB * 4, NB, f);
    st.r0sq = f[0];
    st.rho_new = make_double2(f[0], 0.0);
    st.rn = f[0];
    if (fresh) {
        st.rhs_sq = f[1];
        st.rho = make_double2(1.0, 0.0);
        st.alpha = make_double2(1.0, 0.0);
        st.w = make_double2(1.0, 0.0);
    } else {
        if (st.nrestart++ == 0) st.it = 0;
        st.active = 1;
    }
    st.restart = 0;
    S[j] = st;
}

// masked zeroing of columns (p, v on fresh starts; x on init)
__global__ void k_zero_cols(const RhsState *__restrict__ S, int N, double2 *__restrict__ a, int use_restart_mask) {
    const int j = blockIdx.y;
    if (use_restart_mask && !S[j].restart) return;
    const int64_t o = (int64_t)j * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        a[o + i] = make_double2(0.0, 0.0);
}

// x += f d (refinement; f = -1 reverts a correction), masked per column
__global__ void k_axpy_cols(const int *__restrict__ mask, int N, double2 *__restrict__ x,
                            const double2 *__restrict__ d, double f) {
    const int j = blockIdx.y;
    if (!mask[j]) return;
    const int64_t o = (int64_t)j * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double2 di = d[o + i];
        x[o + i] = cadd(x[o + i], make_double2(f * di.x, f * di.y));
    }
}

// r = b - Ax, per rhs partials of |r|^2 and |b|^2 (true residual)
__global__ __launch_bounds__(kVT) void k_resid(int N, int NB, const double2 *__restrict__ b,
                                               const double2 *__restrict__ Ax, double2 *__restrict__ r,
                                               double *__restrict__ part) {
    __shared__ double sh[4 * kVT / ARH_WAVE];
    const int j = blockIdx.y;
    const Chunk ch = chunk_of(N, NB);
    const int64_t o = (int64_t)j * N;
    double v[4] = {0, 0, 0, 0};
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += kVT) {
        double2 bi = b[o + i];
        double2 ri = csub(bi, Ax[o + i]);
        if (r) r[o + i] = ri;
        v[0] += ri.x * ri.x + ri.y * ri.y;
        v[1] += bi.x * bi.x + bi.y * bi.y;
    }
    block_reduce4(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[((int64_t)j * NB + blockIdx.x) * 4 + k] = v[k];
}

// ------------------------------------------------------------------ host driver
namespace {

struct Work {
    DBuf buf, hbuf;
    double2 *r, *r0, *p, *v, *s, *t, *Ax;
    double *part;
    RhsState *S;
    int *mask;
    int *ctl;  // [0]: gate (active count), [1..]: column list
    int *ghist;  // active count after each iteration of a batch
    void init(int64_t N, int R, int NB) {
        const int64_t vec = N * R;
        const size_t vb = sizeof(double2) * vec;
        const size_t pb = sizeof(double) * 4 * (size_t)R * NB;
        // every piece starts on a 256-byte boundary: the total is the sum of the rounded sizes
        auto rnd = [](size_t n) { return (n + 255) / 256 * 256; };
        const size_t total = 7 * rnd(vb) + rnd(4 * pb) + rnd(sizeof(RhsState) * R) + rnd(sizeof(int) * R) +
                             rnd(sizeof(int) * (R + 1)) + rnd(sizeof(int) * kMaxK);
        buf.alloc(total);
        char *c = buf.as<char>();
        char *const end = c + total;
        auto take = [&](size_t n) {
            char *q = c;
            c += rnd(n);
            ARH_CHECK(c <= end, ARH_EINVAL, "internal: solver workspace overflow");
            return q;
        };
        r = (double2 *)take(vb), r0 = (double2 *)take(vb), p = (double2 *)take(vb), v = (double2 *)take(vb);
        s = (double2 *)take(vb), t = (double2 *)take(vb), Ax = (double2 *)take(vb);
        part = (double *)take(4 * pb);
        S = (RhsState *)take(sizeof(RhsState) * R);
        mask = (int *)take(sizeof(int) * R);
        ctl = (int *)take(sizeof(int) * (R + 1));
        ghist = (int *)take(sizeof(int) * kMaxK);
    }
};

#ifndef ARH_SOLVER_ROWS
#define ARH_SOLVER_ROWS 2048
#endif
// blocks per right-hand side of the vector kernels: about ARH_SOLVER_ROWS rows each
int pick_nb(int N) { return std::max(1, std::min(kMaxNB, (N + ARH_SOLVER_ROWS - 1) / ARH_SOLVER_ROWS)); }

}  // namespace

// BiCGSTAB on operator `prec` (ARH_FP64 / ARH_MIXED) for the rhs with want[j] != 0, from x = 0.
// Returns per-rhs iterations; fills recursive residual history.
// stall_window > 0 (inner solves of the refinement only): a right-hand side whose recursive
// residual has not halved within stall_window iterations is stopped, and so is one that has run
// budget iterations (budget > 0) without converging (both checked at synchronisations); 0 keeps
// Eigen's loop control exactly.
static void bicgstab(arh_hmat *h, int transpose, int prec, int R, const double2 *b, double2 *x,
                     const std::vector<int> &want, const std::vector<double> &tolj, int maxit, const arh_solve_opts &o,
                     arh_solve_stats *st, std::vector<int> &iters, std::vector<double> &err, Work &W, int NB,
                     bool record_hist, int stall_window = 0, int budget = 0) {
    arh_ctx *ctx = h->g->ctx;
    hipStream_t s = ctx->stream;
    const int N = h->g->N;
    const dim3 gv(NB, R), bv(kVT);
    const int pblk = (R + 63) / 64;
    const int zblk = std::max(1, std::min(1024, (N + 255) / 256));
    double *P1 = W.part, *P2 = W.part + 4 * (size_t)R * NB, *P3 = W.part + 8 * (size_t)R * NB;
    const bool hist_on = record_hist && st && st->hist && st->hist_len && o.hist_cap > 0;
    const int cap = hist_on ? o.hist_cap : 0;
    double *hist_dev = nullptr;
    if (hist_on) {
        W.hbuf.alloc(sizeof(double) * (size_t)R * cap);
        hist_dev = W.hbuf.as<double>();
    }
    std::vector<RhsState> hs(R);
    for (int j = 0; j < R; j++) {
        std::memset(&hs[j], 0, sizeof(RhsState));
        hs[j].restart = want[j] ? 1 : 0;
    }
    ARH_HIP(hipMemcpyAsync(W.S, hs.data(), sizeof(RhsState) * R, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_zero_cols, dim3(zblk, R), dim3(256), 0, s, W.S, N, x, 1);
    hipLaunchKernelGGL(k_zero_cols, dim3(zblk, R), dim3(256), 0, s, W.S, N, W.p, 1);
    hipLaunchKernelGGL(k_zero_cols, dim3(zblk, R), dim3(256), 0, s, W.S, N, W.v, 1);
    hipLaunchKernelGGL(k_restart, gv, bv, 0, s, W.S, N, NB, b, (const double2 *)nullptr, W.r, W.r0, P1);
    hipLaunchKernelGGL(k_restart_finish, dim3(pblk), dim3(64), 0, s, W.S, R, NB, P1, 1);
    ARH_HIP(hipGetLastError());
    ARH_HIP(hipMemcpyAsync(hs.data(), W.S, sizeof(RhsState) * R, hipMemcpyDeviceToHost, s));
    ARH_HIP(hipStreamSynchronize(s));
    iters.assign(R, 0);
    err.assign(R, 0.0);
    std::vector<double> hist0(R, 0.0);
    int nact = 0;
    for (int j = 0; j < R; j++) {
        hs[j].active = 0;
        if (!want[j]) continue;
        if (hs[j].rhs_sq == 0.0) {  // Eigen: x = 0, done
            err[j] = 0.0;
            continue;
        }
        hs[j].tol2 = tolj[j] * tolj[j] * hs[j].rhs_sq;
        hs[j].maxit = maxit;
        hist0[j] = std::sqrt(hs[j].rn / hs[j].rhs_sq);
        if (hs[j].rn > hs[j].tol2 && maxit > 0) hs[j].active = 1, nact++;
        err[j] = std::sqrt(hs[j].rn / hs[j].rhs_sq);
    }
    ARH_HIP(hipMemcpyAsync(W.S, hs.data(), sizeof(RhsState) * R, hipMemcpyHostToDevice, s));
    // K from the observed convergence: iterations until the fastest active rhs is predicted to
    // reach its tolerance (so that few iterations run past convergence), 1..kMaxK
    std::vector<double> e_prev(R, 0.0);
    std::vector<int> i_prev(R, 0);
    for (int j = 0; j < R; j++) e_prev[j] = err[j];
    int K = 2;
    // separate host buffers for the two lists: a list is rewritten only after the stream has been
    // synchronised since its last copy
    std::vector<int> ctl(R + 1), ctl_rs(R + 1);
    std::vector<double> best(err);
    std::vector<int> best_it(R, 0);
    while (nact > 0) {
        int nc = 0;
        for (int j = 0; j < R; j++)
            if (hs[j].active) ctl[1 + nc++] = j;
        ctl[0] = nc;
        ARH_HIP(hipMemcpyAsync(W.ctl, ctl.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice, s));
        const int *cols = W.ctl + 1, *gate = W.ctl;
        for (int k = 0; k < K; k++) {
            ctx->prof.cur_slot = k;
            hipLaunchKernelGGL(k_p_update, gv, bv, 0, s, W.S, N, NB, W.r, W.p, W.v);
            hmat_apply(h, transpose, prec, R, W.p, W.v, s, cols, nc, gate);
            hipLaunchKernelGGL(k_dot, gv, bv, 0, s, W.S, N, NB, W.r0, W.v, (const double2 *)nullptr, P1);
            hipLaunchKernelGGL(k_s_update, gv, bv, 0, s, W.S, N, NB, P1, W.r, W.v, W.s);
            hmat_apply(h, transpose, prec, R, W.s, W.t, s, cols, nc, gate);
            hipLaunchKernelGGL(k_dot, gv, bv, 0, s, W.S, N, NB, W.t, W.s, W.t, P2);
            hipLaunchKernelGGL(k_xr_update, gv, bv, 0, s, W.S, N, NB, P1, P2, W.p, W.s, W.t, W.r0, x, W.r, P3);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, W.S, R, NB, P1, P2, P3, hist_dev, cap, W.ctl,
                               W.ghist + k);
            ARH_HIP(hipGetLastError());
        }
        ctx->prof.cur_slot = -1;
        int gh[kMaxK];
        ARH_HIP(hipMemcpyAsync(gh, W.ghist, sizeof(int) * K, hipMemcpyDeviceToHost, s));
        ARH_HIP(hipMemcpyAsync(hs.data(), W.S, sizeof(RhsState) * R, hipMemcpyDeviceToHost, s));
        ARH_HIP(hipStreamSynchronize(s));
        if (ctx->prof.on) {
            // iteration k's products ran iff some rhs was active after iteration k - 1
            ctx->prof.slot_ok.assign(K, 1);
            for (int k = 1; k < K; k++) ctx->prof.slot_ok[k] = gh[k - 1] > 0;
            ctx->prof.flush(false);
            ctx->prof.slot_ok.clear();
        }
        // stalled inner solves stop here (nothing has run on the device since the state was read)
        if (stall_window > 0 || budget > 0) {
            bool changed = false;
            for (int j = 0; j < R; j++) {
                if (!(hs[j].active || hs[j].restart) || hs[j].rhs_sq == 0.0) continue;
                const double e = std::sqrt(hs[j].rn / hs[j].rhs_sq);
                if (budget > 0 && hs[j].iters >= budget) {
                    hs[j].active = 0;
                    hs[j].restart = 0;
                    changed = true;
                } else if (e < 0.5 * best[j]) {
                    best[j] = e;
                    best_it[j] = hs[j].iters;
                } else if (stall_window > 0 && hs[j].iters - best_it[j] >= stall_window) {
                    hs[j].active = 0;
                    hs[j].restart = 0;
                    changed = true;
                }
            }
            if (changed) ARH_HIP(hipMemcpyAsync(W.S, hs.data(), sizeof(RhsState) * R, hipMemcpyHostToDevice, s));
        }
        // breakdown restarts: r = b - A x, r0 = r for the paused right-hand sides
        int nrs = 0;
        for (int j = 0; j < R; j++)
            if (hs[j].restart) ctl_rs[1 + nrs++] = j;
        if (nrs) {
            ctl_rs[0] = nrs;
            ARH_HIP(hipMemcpyAsync(W.ctl, ctl_rs.data(), sizeof(int) * (nrs + 1), hipMemcpyHostToDevice, s));
            hmat_apply(h, transpose, prec, R, x, W.Ax, s, W.ctl + 1, nrs, W.ctl);
            if (st) st->matvecs++;
            hipLaunchKernelGGL(k_restart, gv, bv, 0, s, W.S, N, NB, b, W.Ax, W.r, W.r0, P1);
            hipLaunchKernelGGL(k_restart_finish, dim3(pblk), dim3(64), 0, s, W.S, R, NB, P1, 0);
            ARH_HIP(hipGetLastError());
            for (int j = 0; j < R; j++)
                if (hs[j].restart) {
                    hs[j].restart = 0;
                    hs[j].active = 1;
                    if (st) st->restarts++;
                }
        }
        if (st) st->matvecs += 2 * K;
        nact = 0;
        double kpred = kMaxK;
        for (int j = 0; j < R; j++) {
            if (!want[j] || hs[j].rhs_sq == 0.0) continue;
            err[j] = std::sqrt(hs[j].rn / hs[j].rhs_sq);
            if (!hs[j].active) continue;
            nact++;
            // geometric extrapolation of the residual since the previous synchronisation
            const int di = hs[j].iters - i_prev[j];
            if (di > 0 && err[j] > 0.0 && e_prev[j] > err[j]) {
                const double rate = std::log(e_prev[j] / err[j]) / di;
                kpred = std::min(kpred, std::log(err[j] / tolj[j]) / rate);
            }
            e_prev[j] = err[j];
            i_prev[j] = hs[j].iters;
        }
        K = std::max(1, std::min(kMaxK, (int)std::floor(kpred)));
    }
    for (int j = 0; j < R; j++) iters[j] = hs[j].iters;
    if (hist_on) {
        std::vector<double> hh((size_t)R * cap);
        ARH_HIP(hipMemcpyAsync(hh.data(), hist_dev, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, s));
        ARH_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < R; j++) {
            if (!want[j] || hs[j].rhs_sq == 0.0) continue;
            const int n = std::min(cap, 1 + hs[j].iters);
            double *dst = st->hist + (size_t)j * cap;
            dst[0] = hist0[j];
            for (int k = 1; k < n; k++) dst[k] = hh[(size_t)j * cap + k];
            st->hist_len[j] = n;
        }
    }
}

// true residual ||b - A x|| / ||b|| per rhs (FP64 operator); r receives b - A x if non-null
static void true_residual(arh_hmat *h, int transpose, int R, const double2 *b, const double2 *x, double2 *r, Work &W,
                          int NB, std::vector<double> &res, arh_solve_stats *st) {
    hipStream_t s = h->g->ctx->stream;
    const int N = h->g->N;
    hmat_apply(h, transpose, ARH_FP64, R, x, W.Ax, s);
    if (st) st->matvecs++;
    hipLaunchKernelGGL(k_resid, dim3(NB, R), dim3(kVT), 0, s, N, NB, b, W.Ax, r, W.part);
    ARH_HIP(hipGetLastError());
    std::vector<double> part(4 * (size_t)R * NB);
    ARH_HIP(hipMemcpyAsync(part.data(), W.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    ARH_HIP(hipStreamSynchronize(s));
    res.assign(R, 0.0);
    for (int j = 0; j < R; j++) {
        double rr = 0, bb = 0;
        for (int q = 0; q < NB; q++) rr += part[((size_t)j * NB + q) * 4], bb += part[((size_t)j * NB + q) * 4 + 1];
        res[j] = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    }
}

void solve(arh_hmat *h, int transpose, int R, const double2 *b, double2 *x, const arh_solve_opts &o,
           arh_solve_stats *st) {
    arh_geom *g = h->g;
    hipStream_t s = g->ctx->stream;
    const int N = g->N;
    const int NB = pick_nb(N);
    const int maxit = o.maxit > 0 ? o.maxit : 2 * N;
    const double tol = o.tol;
    ARH_CHECK(R >= 1, ARH_EINVAL, "nrhs must be >= 1");
    ARH_CHECK(tol >= 0.0, ARH_EINVAL, "tol must be >= 0");
    const auto t0 = std::chrono::steady_clock::now();
    Work W;
    W.init(N, R, NB);
    if (st) {
        st->iters = st->matvecs = st->refinements = st->restarts = 0;
        st->max_err = st->max_true_res = 0.0;
        if (st->hist_len)
            for (int j = 0; j < R; j++) st->hist_len[j] = 0;
    }
    std::vector<int> want(R, 1), iters, iters_tot(R, 0);
    std::vector<double> err, tres;
    if (h->storage != ARH_MIXED) {
        bicgstab(h, transpose, ARH_FP64, R, b, x, want, std::vector<double>(R, tol), maxit, o, st, iters, err, W, NB, true);
        for (int j = 0; j < R; j++) iters_tot[j] = iters[j];
        // optional: continue from x until the FP64 true residual meets tol; the correction solves
        // share what is left of maxit and stop when one no longer halves the worst residual
        int sweeps = 0;
        int left = maxit - *std::max_element(iters.begin(), iters.end());
        std::vector<double> last_err = err;
        if (o.true_residual || st) true_residual(h, transpose, R, b, x, nullptr, W, NB, tres, st);
        trace_sweep("fp64", 0, want, iters, tres);
        while (o.true_residual && sweeps < 8 && left > 0) {
            std::vector<int> redo(R, 0);
            int nre = 0;
            for (int j = 0; j < R; j++)
                if (tres[j] > tol) redo[j] = 1, nre++;
            if (!nre) break;
            const double worst = *std::max_element(tres.begin(), tres.end());
            // correction solve A d = b - A x for the unconverged rhs, then x += d
            DBuf rb, dx;
            rb.alloc(sizeof(double2) * (size_t)N * R);
            dx.alloc(sizeof(double2) * (size_t)N * R);
            true_residual(h, transpose, R, b, x, rb.as<double2>(), W, NB, tres, nullptr);
            // per-rhs tolerance relative to the current residual (no coupling between right-hand sides)
            std::vector<double> ctol(R);
            for (int j = 0; j < R; j++) ctol[j] = std::min(0.5, 0.5 * tol / std::max(tres[j], 1e-300));
            bicgstab(h, transpose, ARH_FP64, R, rb.as<double2>(), dx.as<double2>(), redo, ctol, left, o, st, iters,
                     err, W, NB, false);
            left -= *std::max_element(iters.begin(), iters.end());
            ARH_HIP(hipMemcpyAsync(W.mask, redo.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_axpy_cols, dim3(std::max(1, std::min(1024, (N + 255) / 256)), R), dim3(256), 0, s,
                               W.mask, N, x, dx.as<double2>(), 1.0);
            ARH_HIP(hipGetLastError());
            for (int j = 0; j < R; j++) iters_tot[j] += redo[j] ? iters[j] : 0;
            true_residual(h, transpose, R, b, x, nullptr, W, NB, tres, st);
            sweeps++;
            if (st) st->refinements++;
            if (*std::max_element(tres.begin(), tres.end()) > 0.5 * worst) break;
        }
        err = last_err;
    } else {
        // mixed: FP32-operator BiCGSTAB inside FP64 true-residual refinement. A sweep's correction is
        // kept for a right-hand side only if it cuts the FP64 true residual; an inner solve that
        // stalls, or a correction that does not pay, sends the right-hand side to FP64 correction
        // solves (the FP64 operator, Eigen's loop control), which converge whenever the FP64 path does.
        DBuf rb, dx;
        rb.alloc(sizeof(double2) * (size_t)N * R);
        dx.alloc(sizeof(double2) * (size_t)N * R);
        ARH_HIP(hipMemcpyAsync(rb.p, b, sizeof(double2) * (size_t)N * R, hipMemcpyDeviceToDevice, s));
        ARH_HIP(hipMemsetAsync(x, 0, sizeof(double2) * (size_t)N * R, s));
        std::vector<double> bnorm_rel(R, 1.0);  // ||r|| / ||b|| of the current residual
        std::vector<int> redo(R, 1), fp64(R, 0);
        const dim3 ga(std::max(1, std::min(1024, (N + 255) / 256)), R);
        const int stall = std::max(200, std::min(maxit, 400));
        for (int sweep = 0; sweep < 10; sweep++) {
            // per-rhs inner tolerance relative to its current residual, aiming at half the target (not
            // below kInnerFloor); no coupling between right-hand sides. An inner solve gets at most
            // kInnerBudget iterations: on a hard (ill-conditioned) problem the FP32 operator's
            // rounding slows BiCGSTAB to two or three times the FP64 iteration count, so such a
            // right-hand side goes to FP64 after a bounded FP32 attempt
            std::vector<double> itol(R);
            for (int j = 0; j < R; j++)
                itol[j] = std::min(0.5, std::max(kInnerFloor, kInnerFrac * tol / std::max(bnorm_rel[j], 1e-300)));
            bicgstab(h, transpose, ARH_MIXED, R, rb.as<double2>(), dx.as<double2>(), redo, itol, maxit, o, st, iters,
                     err, W, NB, sweep == 0, stall, kInnerBudget);
            for (int j = 0; j < R; j++) iters_tot[j] += redo[j] ? iters[j] : 0;
            ARH_HIP(hipMemcpyAsync(W.mask, redo.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_axpy_cols, ga, dim3(256), 0, s, W.mask, N, x, dx.as<double2>(), 1.0);
            ARH_HIP(hipGetLastError());
            true_residual(h, transpose, R, b, x, rb.as<double2>(), W, NB, tres, st);
            if (st) st->refinements++;
            trace_sweep("fp32", sweep, redo, iters, tres);
            std::vector<int> undo(R, 0);
            int nundo = 0, nre = 0;
            for (int j = 0; j < R; j++) {
                // an inner solve that stopped short of its tolerance (stalled or hit maxit) marks a
                // problem the FP32 operator cannot resolve: finish in FP64 from here
                if (redo[j] && tres[j] > tol && err[j] > itol[j]) fp64[j] = 1;
                if (redo[j] && tres[j] > tol && tres[j] > 0.5 * bnorm_rel[j]) {
                    fp64[j] = 1;  // the FP32 correction did not pay: finish in FP64
                    if (tres[j] > bnorm_rel[j]) undo[j] = 1, nundo++;
                }
                if (!undo[j]) bnorm_rel[j] = tres[j];
                redo[j] = (bnorm_rel[j] > tol && !fp64[j]) ? 1 : 0;
                nre += redo[j];
            }
            if (nundo) {  // revert corrections that made the residual worse
                ARH_HIP(hipMemcpyAsync(W.mask, undo.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_axpy_cols, ga, dim3(256), 0, s, W.mask, N, x, dx.as<double2>(), -1.0);
                ARH_HIP(hipGetLastError());
                true_residual(h, transpose, R, b, x, rb.as<double2>(), W, NB, tres, st);
                for (int j = 0; j < R; j++) bnorm_rel[j] = tres[j];
            }
            if (!nre) break;
        }
        // FP64 correction solves for the right-hand sides still above tol; together they get maxit
        // iterations and stop when one no longer halves the worst residual
        int left = maxit;
        for (int sweep = 0; sweep < 8 && left > 0; sweep++) {
            std::vector<int> f64(R, 0);
            int nf = 0;
            double worst = 0.0;
            for (int j = 0; j < R; j++)
                if (tres[j] > tol) f64[j] = 1, nf++, worst = std::max(worst, tres[j]);
            if (!nf) break;
            std::vector<double> ctol(R);
            for (int j = 0; j < R; j++) ctol[j] = std::min(0.5, 0.5 * tol / std::max(tres[j], 1e-300));
            bicgstab(h, transpose, ARH_FP64, R, rb.as<double2>(), dx.as<double2>(), f64, ctol, left, o, st, iters,
                     err, W, NB, false);
            for (int j = 0; j < R; j++) iters_tot[j] += f64[j] ? iters[j] : 0;
            left -= *std::max_element(iters.begin(), iters.end());
            ARH_HIP(hipMemcpyAsync(W.mask, f64.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_axpy_cols, ga, dim3(256), 0, s, W.mask, N, x, dx.as<double2>(), 1.0);
            ARH_HIP(hipGetLastError());
            true_residual(h, transpose, R, b, x, rb.as<double2>(), W, NB, tres, st);
            if (st) st->refinements++;
            trace_sweep("fp64 correction", sweep, f64, iters, tres);
            double now = 0.0;
            for (int j = 0; j < R; j++)
                if (f64[j]) now = std::max(now, tres[j]);
            if (now > 0.5 * worst) break;
        }
        err = tres;
    }
    if (tres.empty()) true_residual(h, transpose, R, b, x, nullptr, W, NB, tres, st);
    if (st) {
        for (int j = 0; j < R; j++) {
            st->iters = std::max(st->iters, iters_tot[j]);
            st->max_err = std::max(st->max_err, err[j]);
            st->max_true_res = std::max(st->max_true_res, tres[j]);
            if (st->iters_per_rhs) st->iters_per_rhs[j] = iters_tot[j];
            if (st->true_res) st->true_res[j] = tres[j];
        }
        st->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (o.fail_on_noconv) {
        // true-residual modes are judged on the FP64 true residual, Eigen semantics on the recursive one
        const bool use_true = o.true_residual || h->storage == ARH_MIXED;
        for (int j = 0; j < R; j++) {
            const double e = use_true ? tres[j] : err[j];
            ARH_CHECK(e <= tol, ARH_ECONV,
                      "BiCGSTAB did not reach the tolerance (rhs " + std::to_string(j) + ", residual " +
                          std::to_string(e) + " > " + std::to_string(tol) + ")");
        }
    }
}

}  // namespace arh
