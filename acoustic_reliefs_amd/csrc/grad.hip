// Shape gradient of the diffusion coefficient w.r.t. vertex heights (y coordinates).
//
// Restates compute_der_A/B/C (adjoint.cpp:244-386) restructured per "touched" element i (an
// element with at least one differentiable vertex) so each derivative is evaluated once and
// shared by the three corners and by every rhs of a batch:
//   A_(i,c) = sum_rhs Re( x_i  sum_l  vf_l  dQ_li / dq_(i,c) )                 (adjoint.cpp:268-277)
//   B_(i,c) = sum_rhs Re( lam_i sum_e x_e dM_ie / dp_i ) + Re( x_i sum_e lam_e dM_ei / dq_(i,c) )   (:314-339)
//   C_(i,c) = -sum_rhs Re( lam_i db_i / dp_i )                                  (:373)
// and gradient_h = sum_{(i,c) : E[i][c] = h} A - B + C (element order, then corner order, as the
// reference's elements[h] lists, acoustics3d.cpp:83-98).
// Derivatives (derived from the kernel, not transcribed from derivatives.cpp): with d = p - q,
// r = |d|, s = -(d . n), F = e^{-ikr}(-ikr-1)/(4 pi r^3), F' = e^{-ikr}(-k^2/r^2 + 3(1+ikr)/r^4)/(4 pi):
//   collocation centroid moves (dp/dh = e_y/3):   dK = (F' d_y s / r - F n_y) / 3
//   corner c of the integration triangle moves (dq/dh = beta_c e_y, beta = (1-xi-eta, xi, eta)):
//                                                 dK = -F' beta_c d_y s / r + F (beta_c n_y - d . dn_c)
//   dn_a = e_y x (b - c), dn_b = e_y x (c - a), dn_c = (b - a) x e_y.
#include <algorithm>
#include <vector>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

constexpr int kGT = 128;  // threads per workgroup (A and C: one touched element each; B: a quad of lanes each)
// Rhs per lane and pass. A quad of lanes shares one touched element: the lanes split the quadrature points
// of its derivatives and sum them with lane swaps, then each lane accumulates every fourth rhs, so a pass
// covers 4 kRQ right-hand sides and evaluates each derivative once (37 plane waves: 2 passes; 5 per lane keeps
// the kernels at 2 waves per SIMD, which measured faster than one pass at 1 wave).
#ifndef ARH_GRAD_RQ
#define ARH_GRAD_RQ 5
#endif
constexpr int kRQ = ARH_GRAD_RQ;

__device__ __forceinline__ void FdF(double k, double r, double2 &F, double2 &dF) {
    double sn, cs;
    trig_sincos(k * r, &sn, &cs);
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2;
    const double kr = k * r;
    const double inv4pi = 1.0 / (4.0 * kPi);
    F = make_double2((-cs - kr * sn) / r3 * inv4pi, (sn - kr * cs) / r3 * inv4pi);
    const double A = -k * k / r2 + 3.0 / r4, B = 3.0 * k / r3;
    dF = make_double2((cs * A + sn * B) * inv4pi, (cs * B - sn * A) * inv4pi);
}

// The value of another lane of the quad (DPP quad permutation CTRL; every lane of the quad is active).
template <int CTRL>
__device__ __forceinline__ double quad_swap(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum of v over the 4 lanes of a quad: (l0 + l1) + (l2 + l3), the same bits in every lane.
__device__ __forceinline__ double quad_sum(double v) {
    v += quad_swap<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    v += quad_swap<0x4E>(v);  // quad_perm [2, 3, 0, 1]
    return v;
}

// One lane's share of the two derivatives of pair (collocation i, element e): quadrature points
// g = q4, q4 + 4, q4 + 8 of the list {dM_ie/dp_i: rule points 0..5 on triangle e, collocation centroid i;
// dM_ei/dq_(i,c): rule points 0..5 on triangle i, collocation centroid e}. Both kinds share the distance
// and kernel factors; each point's weight goes to its own kind only (the other is multiplied by 0).
__device__ __forceinline__ void dM_share(double k, const Tri &ti, const Tri &te, int q4, double2 &dp, double2 dq[3]) {
    const double bcx = ti.bx - ti.cx, bcz = ti.bz - ti.cz;  // dn of triangle i's corners (x and z only)
    const double dnx[3] = {bcz, ti.cz, -ti.bz};
    const double dnz[3] = {-bcx, -ti.cx, ti.bx};
    dp = make_double2(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < 3; c++) dq[c] = make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const int g = q4 + 4 * m;
        const bool isq = g >= 6;
        double w, xi, eta;
        quad_point(isq ? g - 6 : g, w, xi, eta);
        const Tri &t = isq ? ti : te;                // integration triangle
        const double px = isq ? te.gx : ti.gx, py = isq ? te.gy : ti.gy, pz = isq ? te.gz : ti.gz;
        const double dx = px - (t.ax + xi * t.bx + eta * t.cx);
        const double dy = py - (t.ay + xi * t.by + eta * t.cy);
        const double dz = pz - (t.az + xi * t.bz + eta * t.cz);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double sn_ = -(dx * t.nx + dy * t.ny + dz * t.nz);
        double2 F, dF;
        FdF(k, r, F, dF);
        const double a = dy * sn_ / r;
        const double wp = isq ? 0.0 : w, wq = isq ? w : 0.0;
        dp.x += wp * (dF.x * a - F.x * t.ny);
        dp.y += wp * (dF.y * a - F.y * t.ny);
        const double beta[3] = {1.0 - xi - eta, xi, eta};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double f1 = -beta[c] * a;
            const double f2 = beta[c] * t.ny - (dx * dnx[c] + dz * dnz[c]);
            dq[c].x += wq * (dF.x * f1 + F.x * f2);
            dq[c].y += wq * (dF.y * f1 + F.y * f2);
        }
    }
    dp = make_double2(quad_sum(dp.x) * (0.5 / 3.0), quad_sum(dp.y) * (0.5 / 3.0));
#pragma unroll
    for (int c = 0; c < 3; c++) dq[c] = make_double2(quad_sum(dq[c].x) * 0.5, quad_sum(dq[c].y) * 0.5);
}

// Element-major copy of x and lambda: xl[(e * R + j) * 2] = x_j[e], xl[(e * R + j) * 2 + 1] = lambda_j[e], so the
// rhs values a quad reads for one element share a few cache lines.
__global__ void k_grad_pack(int N, int R, const double2 *__restrict__ x, const double2 *__restrict__ lam,
                            double2 *__restrict__ xl) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= (int64_t)N * R) return;
    const int e = (int)(q / R), j = (int)(q % R);
    xl[2 * q] = x[(int64_t)j * N + e];
    xl[2 * q + 1] = lam[(int64_t)j * N + e];
}

// B partials: grid.x over touched elements (a quad of lanes each), grid.y splits the element loop; adds
// to partB[(split * T + t) * 3 + c] (one pass over rhs j0..j0 + 4 kRQ - 1; passes accumulate).
__global__ __launch_bounds__(kGT) void k_grad_B(const Tri *__restrict__ tri, int N, int T, const int *__restrict__ touched,
                                                double k, int R, int j0, const double2 *__restrict__ x,
                                                const double2 *__restrict__ lam, const double2 *__restrict__ xl,
                                                int esplit, double *__restrict__ partB) {
    const int t = (int)((blockIdx.x * kGT + threadIdx.x) >> 2), q4 = threadIdx.x & 3;
    const int sp = blockIdx.y;
    if (t >= T) return;  // the whole quad leaves together
    const int i = touched[t];
    const Tri ti = tri[i];
    const int len = (N + esplit - 1) / esplit;
    const int e0 = sp * len, e1 = min(N, e0 + len);
    double2 s1[kRQ], s2[kRQ][3];
#pragma unroll
    for (int m = 0; m < kRQ; m++) {
        s1[m] = make_double2(0.0, 0.0);
        for (int c = 0; c < 3; c++) s2[m][c] = make_double2(0.0, 0.0);
    }
    for (int e = e0; e < e1; e++) {
        const Tri te = tri[e];
        double2 dp, dq[3];
        dM_share(k, ti, te, q4, dp, dq);  // dM_ie/dp_i (collocation i moves), dM_ei/dq_(i,c) (corner c moves)
#pragma unroll
        for (int m = 0; m < kRQ; m++) {
            const int j = j0 + q4 + 4 * m;
            if (j < R) {
                const double2 xe = xl[((int64_t)e * R + j) * 2];
                const double2 le = xl[((int64_t)e * R + j) * 2 + 1];
                cfma(s1[m], xe, dp);
#pragma unroll
                for (int c = 0; c < 3; c++) cfma(s2[m][c], le, dq[c]);
            }
        }
    }
    double out[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < kRQ; m++) {
        const int j = j0 + q4 + 4 * m;
        if (j < R) {
            const double2 li = lam[(int64_t)j * N + i];
            const double2 xi = x[(int64_t)j * N + i];
            const double a = cmul(s1[m], li).x;
#pragma unroll
            for (int c = 0; c < 3; c++) out[c] += a + cmul(s2[m][c], xi).x;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = quad_sum(out[c]);
    if (q4 == 0)
#pragma unroll
        for (int c = 0; c < 3; c++) partB[((int64_t)sp * T + t) * 3 + c] += out[c];
}

// dM(p, tri)/dq_c for corner c = 0, 1, 2 of the integration triangle moving in y (sum over the rule, times
// 1/2): lane q4 of the quad takes rule points q4 and q4 + 4 (< 6), and the quad sums the shares.
__device__ __forceinline__ void dM_dq3_quad(double k, double px, double py, double pz, const Tri &t, int q4,
                                            double2 out[3]) {
    const double bcx = t.bx - t.cx, bcz = t.bz - t.cz;
    const double dnx[3] = {bcz, t.cz, -t.bz};
    const double dnz[3] = {-bcx, -t.cx, t.bx};
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const int g = q4 + 4 * m;
        double w, xi, eta;
        quad_point(g < 6 ? g : 0, w, xi, eta);
        if (g >= 6) w = 0.0;  // lanes 2 and 3 have one point
        const double dx = px - (t.ax + xi * t.bx + eta * t.cx);
        const double dy = py - (t.ay + xi * t.by + eta * t.cy);
        const double dz = pz - (t.az + xi * t.bz + eta * t.cz);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double sn_ = -(dx * t.nx + dy * t.ny + dz * t.nz);
        double2 F, dF;
        FdF(k, r, F, dF);
        const double beta[3] = {1.0 - xi - eta, xi, eta};
        const double a = dy * sn_ / r;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double f1 = -beta[c] * a;
            const double f2 = beta[c] * t.ny - (dx * dnx[c] + dz * dnz[c]);
            out[c].x += w * (dF.x * f1 + F.x * f2);
            out[c].y += w * (dF.y * f1 + F.y * f2);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = make_double2(quad_sum(out[c].x) * 0.5, quad_sum(out[c].y) * 0.5);
}

// A and C per touched element, a quad of lanes each (listeners: LL is small). The lanes split each
// listener derivative's rule points, then lane q4 accumulates rhs j = q4, q4 + 4, ... (passes of 4 kRQ),
// so a derivative is evaluated once per pass rather than once per rhs.
__global__ __launch_bounds__(kGT) void k_grad_AC(const Tri *__restrict__ tri, int N, int T,
                                                 const int *__restrict__ touched, double k, int R,
                                                 const double2 *__restrict__ x, const double2 *__restrict__ lam,
                                                 const double2 *__restrict__ vf, const double *__restrict__ L, int LL,
                                                 int src_kind, const double *__restrict__ src,
                                                 double *__restrict__ cA, double *__restrict__ cC) {
    const int t = (int)((blockIdx.x * kGT + threadIdx.x) >> 2), q4 = threadIdx.x & 3;
    if (t >= T) return;  // the whole quad leaves together
    const int i = touched[t];
    const Tri ti = tri[i];
    double a[3] = {0.0, 0.0, 0.0}, cc = 0.0;
    for (int j0 = 0; j0 < R; j0 += 4 * kRQ) {
        double2 sa[kRQ][3];
#pragma unroll
        for (int m = 0; m < kRQ; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) sa[m][c] = make_double2(0.0, 0.0);
        for (int l = 0; l < LL; l++) {
            double2 dq[3];
            dM_dq3_quad(k, L[3 * l], L[3 * l + 1], L[3 * l + 2], ti, q4, dq);
#pragma unroll
            for (int m = 0; m < kRQ; m++) {
                const int j = j0 + q4 + 4 * m;
                if (j < R) {
                    const double2 v = vf[(int64_t)j * LL + l];
#pragma unroll
                    for (int c = 0; c < 3; c++) cfma(sa[m][c], v, dq[c]);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < kRQ; m++) {
            const int j = j0 + q4 + 4 * m;
            if (j >= R) continue;
            const double2 xi = x[(int64_t)j * N + i];
#pragma unroll
            for (int c = 0; c < 3; c++) a[c] += cmul(sa[m][c], xi).x;
            // db_i / dh
            double2 db;
            const double sx = src[3 * j], sy = src[3 * j + 1], sz = src[3 * j + 2];
            if (src_kind == ARH_RHS_POINT) {
                const double dx = ti.gx - sx, dy = ti.gy - sy, dz = ti.gz - sz;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                double sn, cs;
                trig_sincos(k * r, &sn, &cs);
                // e^{-ikr}(-1 - ikr) / (4 pi r^2) * d_y / (3 r)
                const double f = dy / (3.0 * r) / (4.0 * kPi * r * r);
                db = make_double2((-cs - k * r * sn) * f, (sn - k * r * cs) * f);
            } else {
                double sn, cs;
                trig_sincos(k * (sx * ti.gx + sy * ti.gy + sz * ti.gz), &sn, &cs);
                // (i f) e^{-i k d.c} with f = -k d_y / 3
                const double f = -k * sy / 3.0;
                db = make_double2(f * sn, f * cs);
            }
            cc -= cmul(db, lam[(int64_t)j * N + i]).x;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) a[c] = quad_sum(a[c]);
    cc = quad_sum(cc);
    if (q4 == 0)
#pragma unroll
        for (int c = 0; c < 3; c++) cA[t * 3 + c] = a[c], cC[t * 3 + c] = cc;
}

// gradient_h: sum over (touched element, corner) in list order; B summed over the splits first. With G, the
// terms are not stored but added: G_h = G_h + ((A_h - B_h) + C_h), the order of the host sum A - B + C.
__global__ void k_grad_scatter(int HH, int T, int esplit, const int *__restrict__ hptr, const int *__restrict__ hlist,
                               const double *__restrict__ cA, const double *__restrict__ partB,
                               const double *__restrict__ cC, double *__restrict__ A, double *__restrict__ B,
                               double *__restrict__ C, double *__restrict__ G) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= HH) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int e = hptr[h]; e < hptr[h + 1]; e++) {
        const int tc = hlist[e];  // t * 3 + corner
        a += cA[tc];
        double bb = 0.0;
        for (int sp = 0; sp < esplit; sp++) bb += partB[(int64_t)sp * T * 3 + tc];
        b += bb;
        c += cC[tc];
    }
    if (G)
        G[h] = G[h] + ((a - b) + c);
    else
        A[h] = a, B[h] = b, C[h] = c;
}

void launch_shape_grad(arh_geom *g, double k, const double *L_dev, int LL, const int *diff, int HH, int nrhs,
                       const double2 *x, const double2 *lam, const double2 *vf, int src_kind, const double *src_dev,
                       double *A, double *B, double *C, double *G) {
    hipStream_t s = g->ctx->stream;
    const int N = g->N, V = g->V;
    // slots of differentiable vertices; touched elements; per-h (element, corner) lists
    std::vector<int> slot(V, -1);
    for (int h = 0; h < HH; h++) {
        ARH_CHECK(diff[h] >= 0 && diff[h] < V, ARH_EINVAL, "diff vertex id out of range");
        slot[diff[h]] = h;  // a repeated id keeps its last slot, as the reference's map assignment does
    }
    std::vector<int> Eh(3 * (size_t)N);
    ARH_HIP(hipMemcpyAsync(Eh.data(), g->E.p, sizeof(int) * 3 * N, hipMemcpyDeviceToHost, s));
    ARH_HIP(hipStreamSynchronize(s));
    std::vector<int> touched, tpos(N, -1);
    std::vector<std::vector<int>> lists(HH);
    for (int e = 0; e < N; e++)
        for (int c = 0; c < 3; c++) {
            const int h = slot[Eh[3 * e + c]];
            if (h < 0) continue;
            if (tpos[e] < 0) tpos[e] = (int)touched.size(), touched.push_back(e);
            lists[h].push_back(tpos[e] * 3 + c);
        }
    const int T = (int)touched.size();
    std::vector<int> hptr(HH + 1, 0), hlist;
    for (int h = 0; h < HH; h++) {
        hlist.insert(hlist.end(), lists[h].begin(), lists[h].end());
        hptr[h + 1] = (int)hlist.size();
    }
    if (T == 0 && G) return;  // nothing to add
    if (T == 0) {
        ARH_HIP(hipMemsetAsync(A, 0, sizeof(double) * HH, s));
        ARH_HIP(hipMemsetAsync(B, 0, sizeof(double) * HH, s));
        ARH_HIP(hipMemsetAsync(C, 0, sizeof(double) * HH, s));
        ARH_HIP(hipStreamSynchronize(s));
        return;
    }
    const int gxb = (4 * T + kGT - 1) / kGT;  // a quad of lanes per touched element
    const int esplit = std::max(1, std::min(std::max(1, N / 64), (2048 + gxb - 1) / gxb));
    DBuf d_touched, d_hptr, d_hlist, cA, cC, pB;
    upload(d_touched, touched, s);
    upload(d_hptr, hptr, s);
    upload(d_hlist, hlist, s);
    cA.alloc(sizeof(double) * 3 * T);
    cC.alloc(sizeof(double) * 3 * T);
    pB.alloc(sizeof(double) * 3 * (size_t)T * esplit);
    DBuf xl;
    xl.alloc(sizeof(double2) * 2 * (size_t)N * nrhs);
    hipLaunchKernelGGL(k_grad_pack, dim3((unsigned)(((int64_t)N * nrhs + 255) / 256)), dim3(256), 0, s, N, nrhs, x, lam,
                       xl.as<double2>());
    ARH_HIP(hipMemsetAsync(pB.p, 0, sizeof(double) * 3 * (size_t)T * esplit, s));
    for (int j0 = 0; j0 < nrhs; j0 += 4 * kRQ)
        hipLaunchKernelGGL(k_grad_B, dim3(gxb, esplit), dim3(kGT), 0, s, g->tri.as<Tri>(), N, T, d_touched.as<int>(), k,
                           nrhs, j0, x, lam, xl.as<double2>(), esplit, pB.as<double>());
    hipLaunchKernelGGL(k_grad_AC, dim3(gxb), dim3(kGT), 0, s, g->tri.as<Tri>(), N, T, d_touched.as<int>(), k, nrhs, x,
                       lam, vf, L_dev, LL, src_kind, src_dev, cA.as<double>(), cC.as<double>());
    hipLaunchKernelGGL(k_grad_scatter, dim3((HH + 255) / 256), dim3(256), 0, s, HH, T, esplit, d_hptr.as<int>(),
                       d_hlist.as<int>(), cA.as<double>(), pB.as<double>(), cC.as<double>(), A, B, C, G);
    ARH_HIP(hipGetLastError());
    ARH_HIP(hipStreamSynchronize(s));
}

}  // namespace arh
