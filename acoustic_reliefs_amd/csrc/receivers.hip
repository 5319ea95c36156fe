// Right-hand sides, the receiver (listener) matrix product and the diffusion coefficient.
#include <algorithm>
#include <mutex>
#include <string>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

// b_j,i: point source e^{-ikr}/(4 pi r), r = |c_i - s_j| (helpers.cpp:46-50), or plane wave
// e^{-ik d_j . c_i} (sphere.h:22-26).
__global__ void k_rhs(const Tri *__restrict__ tri, int N, double k, int kind, const double *__restrict__ vecs,
                      double2 *__restrict__ b) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const Tri t = tri[i];
    const double vx = vecs[3 * j], vy = vecs[3 * j + 1], vz = vecs[3 * j + 2];
    double2 out;
    if (kind == ARH_RHS_POINT) {
        double dx = t.gx - vx, dy = t.gy - vy, dz = t.gz - vz;
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double sn, cs;
        trig_sincos(k * r, &sn, &cs);
        double f = 1.0 / (4.0 * kPi * r);
        out = make_double2(cs * f, -sn * f);
    } else {
        double ph = k * (vx * t.gx + vy * t.gy + vz * t.gz);
        double sn, cs;
        trig_sincos(ph, &sn, &cs);
        out = make_double2(cs, -sn);
    }
    b[(int64_t)j * N + i] = out;
}

void launch_rhs(arh_geom *g, double k, int kind, int nrhs, const double *vecs_dev, double2 *b) {
    hipLaunchKernelGGL(k_rhs, dim3((g->N + 255) / 256, nrhs), dim3(256), 0, g->ctx->stream, g->tri.as<Tri>(), g->N, k,
                       kind, vecs_dev, b);
    ARH_HIP(hipGetLastError());
}

// Q (LL x N, row-major): Q_l,i = M(L_l, element i) (integrate.cpp:46-57).
__global__ void k_qmat(const Tri *__restrict__ tri, int N, double k, const double *__restrict__ L,
                       double2 *__restrict__ Q) {
    const int l = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    Q[(int64_t)l * N + i] = bem_entry(k, L[3 * l], L[3 * l + 1], L[3 * l + 2], tri[i]);
}

constexpr int kRT = 256;
constexpr int kRC = 8;

// Sources of the receiver-matrix entries Q_l,i: read from a materialised Q (LL x N, row-major) or
// computed on the fly from the receiver coordinates (the same bem_entry k_qmat evaluates).
struct QStored {
    const double2 *__restrict__ Q;
    int N;
    __device__ double2 operator()(int l, int i) const { return Q[(int64_t)l * N + i]; }
};
struct QOnTheFly {
    const Tri *__restrict__ tri;
    const double *__restrict__ L;
    double k;
    __device__ double2 operator()(int l, int i) const {
        return bem_entry(k, L[3 * l], L[3 * l + 1], L[3 * l + 2], tri[i]);
    }
};

// y[j, l] = sum_i Q[l, i] x[j, i]; one workgroup per receiver, rhs in chunks of kRC.
template <class Src>
__global__ __launch_bounds__(kRT) void k_qx(Src q, int N, int LL, int R, const double2 *__restrict__ x,
                                            double2 *__restrict__ y) {
    __shared__ double2 sh[kRT / ARH_WAVE][kRC];
    const int l = blockIdx.x;
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
    for (int j0 = 0; j0 < R; j0 += kRC) {
        double2 acc[kRC];
#pragma unroll
        for (int c = 0; c < kRC; c++) acc[c] = make_double2(0.0, 0.0);
        for (int i = threadIdx.x; i < N; i += kRT) {
            double2 a = q(l, i);
#pragma unroll
            for (int c = 0; c < kRC; c++)
                if (j0 + c < R) cfma(acc[c], a, x[(int64_t)(j0 + c) * N + i]);
        }
#pragma unroll
        for (int c = 0; c < kRC; c++) acc[c] = wave_sum2(acc[c]);
        __syncthreads();
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < kRC; c++) sh[w][c] = acc[c];
        __syncthreads();
        if (threadIdx.x < kRC && j0 + threadIdx.x < R) {
            double2 s = sh[0][threadIdx.x];
            for (int ww = 1; ww < kRT / ARH_WAVE; ww++) s = cadd(s, sh[ww][threadIdx.x]);
            y[(int64_t)(j0 + threadIdx.x) * LL + l] = s;
        }
    }
}

// y[j, i] = sum_l Q[l, i] x[j, l] (the adjoint right-hand side Q^T vf, adjoint.cpp:198).
template <class Src>
__global__ __launch_bounds__(kRT) void k_qtx(Src q, int N, int LL, int R, const double2 *__restrict__ x,
                                             double2 *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= N) return;
    double2 acc = make_double2(0.0, 0.0);
    for (int l = 0; l < LL; l++) cfma(acc, q(l, i), x[(int64_t)j * LL + l]);
    y[(int64_t)j * N + i] = acc;
}

void launch_receivers(arh_geom *g, double k, const double *L_dev, int LL, int transpose, int nrhs, const double2 *x,
                      double2 *y) {
    arh_ctx *ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const int N = g->N;
    DBuf Q;
    Q.alloc(sizeof(double2) * (size_t)LL * N);
    hipLaunchKernelGGL(k_qmat, dim3((N + 255) / 256, LL), dim3(256), 0, s, g->tri.as<Tri>(), N, k, L_dev,
                       Q.as<double2>());
    const QStored src{Q.as<double2>(), N};
    if (!transpose)
        hipLaunchKernelGGL(k_qx<QStored>, dim3(LL), dim3(kRT), 0, s, src, N, LL, nrhs, x, y);
    else
        hipLaunchKernelGGL(k_qtx<QStored>, dim3((N + kRT - 1) / kRT, nrhs), dim3(kRT), 0, s, src, N, LL, nrhs, x, y);
    ARH_HIP(hipGetLastError());
    ARH_HIP(hipStreamSynchronize(s));  // Q is freed on return
}

// ---------------------------------------------------------------- receivers without allocation
// arh_receivers_noalloc keeps the receiver coordinates in this file-scope device array (one copy
// per GPU, held until the process ends) and evaluates the receiver-matrix entries on the fly, so it
// needs no device memory beyond its inputs and outputs. Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU (other streams, other threads) race.
static __device__ __align__(256) double scratch[3 * kNoAllocMaxReceivers];

namespace {
std::mutex scratch_mu;
double *scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

// The device address of this GPU's copy of scratch (the array's name is only a host-side stand-in);
// looked up once per device, with that device current.
static double *scratch_on(int device) {
    ARH_CHECK(device >= 0 && device < kNoAllocMaxDevices, ARH_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        ARH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ARH_CHECK(p != nullptr, ARH_EHIP, "hipGetSymbolAddress returned no address for the receiver scratch");
        scratch_ptr[device] = static_cast<double *>(p);
    }
    return scratch_ptr[device];
}

void launch_receivers_noalloc(arh_geom *g, double k, const double *L_host, int LL, int transpose, int nrhs,
                              const double2 *x, double2 *y) {
    ARH_CHECK(LL <= kNoAllocMaxReceivers, ARH_EINVAL,
              "arh_receivers_noalloc takes at most " + std::to_string(kNoAllocMaxReceivers) + " receivers, got " +
                  std::to_string(LL));
    arh_ctx *ctx = g->ctx;
    hipStream_t s = ctx->stream;
    double *Ld = scratch_on(ctx->device);
    // stream-ordered after earlier work on this stream that may still read the previous coordinates
    ARH_HIP(hipMemcpyAsync(Ld, L_host, sizeof(double) * 3 * (size_t)LL, hipMemcpyHostToDevice, s));
    const QOnTheFly src{g->tri.as<Tri>(), Ld, k};
    const int N = g->N;
    if (!transpose)
        hipLaunchKernelGGL(k_qx<QOnTheFly>, dim3(LL), dim3(kRT), 0, s, src, N, LL, nrhs, x, y);
    else
        hipLaunchKernelGGL(k_qtx<QOnTheFly>, dim3((N + kRT - 1) / kRT, nrhs), dim3(kRT), 0, s, src, N, LL, nrhs, x,
                           y);
    ARH_HIP(hipGetLastError());
}

// Diffusion coefficient per rhs with n_freqs = 1 (adjoint.cpp:112-127, 145, 388-397).
__global__ __launch_bounds__(256) void k_diffusion(int LL, const double2 *__restrict__ y, double *__restrict__ d,
                                                   double *__restrict__ u, double2 *__restrict__ vf) {
    __shared__ double sh[2][4];
    const int j = blockIdx.x;
    const double2 *yj = y + (int64_t)j * LL;
    double s = 0.0, q = 0.0;
    for (int l = threadIdx.x; l < LL; l += 256) {
        double a = cabs_(yj[l]);
        double L = a * a;
        s += L;
        q += L * L;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
    __shared__ double part[2][4];
    if (lane == 0) part[0][w] = s, part[1][w] = q;
    __syncthreads();
    const double S = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    const double Qs = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    (void)sh;
    if (threadIdx.x == 0) d[j] = (S * S - Qs) / ((LL - 1) * Qs);
    for (int l = threadIdx.x; l < LL; l += 256) {
        double2 yl = yj[l];
        double a = cabs_(yl);
        double L = a * a;
        double ul = 4.0 * (S * Qs - L * S * S) / ((LL - 1) * Qs * Qs) * a;
        if (u) u[(int64_t)j * LL + l] = ul;
        if (vf) vf[(int64_t)j * LL + l] = make_double2(ul * yl.x / a, -ul * yl.y / a);
    }
}

__global__ void k_scale(double2 *__restrict__ a, int64_t n, double f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = make_double2(f * a[i].x, f * a[i].y);
}

__global__ void k_add(const double *__restrict__ x, double *__restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = y[i] + x[i];
}

void launch_add(hipStream_t s, const double *x, double *y, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add, dim3((unsigned)std::min<int64_t>(1024, (n + 255) / 256)), dim3(256), 0, s, x, y, n);
    ARH_HIP(hipGetLastError());
}

void launch_scale(hipStream_t s, double2 *a, int64_t n, double f) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)std::min<int64_t>(1024, (n + 255) / 256)), dim3(256), 0, s, a, n, f);
    ARH_HIP(hipGetLastError());
}

void launch_diffusion(hipStream_t s, int LL, int nrhs, const double2 *y, double *d, double *u, double2 *vf) {
    hipLaunchKernelGGL(k_diffusion, dim3(nrhs), dim3(256), 0, s, LL, y, d, u, vf);
    ARH_HIP(hipGetLastError());
}

}  // namespace arh
