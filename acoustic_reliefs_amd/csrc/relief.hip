// Relief parameterisation on the device: heights of the differentiable vertices sampled bilinearly from a
// heightfield texture (eval_uv, acoustics_opt.py:24-34) and its adjoint, the texture gradient from the
// per-vertex gradient (acoustic_gradient's scatter, acoustics_opt.py:45-56).
//
// The taps (four texels and weights per vertex) are fixed by the uvs, so they are formed once on the host
// with the reference's arithmetic (np.modf of u * w - 0.5, negative fractions wrapped to 1 + f, indices
// clamped) and the weights (1 - vf)(1 - uf), (1 - vf) uf, vf (1 - uf), vf uf. The adjoint is a gather:
// every texel sums its contributions in (vertex, tap) order from a CSR list -- the order of the
// reference's sequential scatter -- without atomics. Neither side fuses multiply-adds (contraction is
// switched off in the kernels and in the tap set-up), so the results are those of the reference's numpy
// expressions bit for bit.
#include <algorithm>
#include <cmath>

#include "arh_internal.h"

namespace arh {

constexpr int kRT = 256;

__global__ __launch_bounds__(kRT) void k_relief_eval(int n, const int4 *__restrict__ idx,
                                                    const double4 *__restrict__ wts,
                                                    const double *__restrict__ tex, double *__restrict__ out) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * kRT + threadIdx.x;
    if (v >= n) return;
    const int4 t = idx[v];
    const double4 w = wts[v];
    // ((w0 t0 + w1 t1) + w2 t2) + w3 t3, each product rounded (plain operators under the pragma: the
    // operations written here carry no contraction permission)
    double s = w.x * tex[t.x];
    s = s + w.y * tex[t.y];
    s = s + w.z * tex[t.z];
    s = s + w.w * tex[t.w];
    out[v] = s;
}

__global__ __launch_bounds__(kRT) void k_relief_adjoint(int ntex, const int *__restrict__ ptr,
                                                       const int *__restrict__ ent, const double *__restrict__ wflat,
                                                       const double *__restrict__ g, double *__restrict__ out) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * kRT + threadIdx.x;
    if (t >= ntex) return;
    double s = 0.0;
    for (int e = ptr[t]; e < ptr[t + 1]; e++) {
        const int q = ent[e];  // vertex * 4 + tap
        s = s + wflat[q] * g[q >> 2];
    }
    out[t] = s;
}

#pragma clang fp contract(off)
void relief_build(arh_relief *r, const double *uvs) {
    const int n = r->n, h = r->h, w = r->w;
    std::vector<int4> idx(n);
    std::vector<double4> wts(n);
    std::vector<int> ptr((size_t)h * w + 1, 0);
    for (int v = 0; v < n; v++) {
        const double u = uvs[2 * v], vv = uvs[2 * v + 1];
        ARH_CHECK(std::isfinite(u) && std::isfinite(vv) && u >= 0.0 && vv >= 0.0 && u <= 1.0 && vv <= 1.0, ARH_EINVAL,
                  "uvs must lie in [0, 1]");
        const double pu = u * w, pv = vv * h;  // rounded products, then the shift (no fused multiply-add)
        double ui, vi;
        double uf = std::modf(pu - 0.5, &ui);
        double vf = std::modf(pv - 0.5, &vi);
        if (uf < 0) uf = 1 + uf;
        if (vf < 0) vf = 1 + vf;
        const int i0 = std::min(std::max((int)ui, 0), w - 1), j0 = std::min(std::max((int)vi, 0), h - 1);
        const int i1 = std::min(i0 + 1, w - 1), j1 = std::min(j0 + 1, h - 1);
        idx[v] = make_int4(j0 * w + i0, j0 * w + i1, j1 * w + i0, j1 * w + i1);
        const double a = 1 - vf, b = 1 - uf;
        wts[v] = make_double4(a * b, a * uf, vf * b, vf * uf);
        ptr[idx[v].x + 1]++, ptr[idx[v].y + 1]++, ptr[idx[v].z + 1]++, ptr[idx[v].w + 1]++;
    }
    for (size_t t = 1; t < ptr.size(); t++) ptr[t] += ptr[t - 1];
    std::vector<int> ent((size_t)4 * n), fill(ptr.begin(), ptr.end() - 1);
    for (int v = 0; v < n; v++) {
        const int q[4] = {idx[v].x, idx[v].y, idx[v].z, idx[v].w};
        for (int k = 0; k < 4; k++) ent[fill[q[k]]++] = 4 * v + k;  // (vertex, tap) order within a texel
    }
    hipStream_t s = r->ctx->stream;
    upload(r->idx, idx, s);
    upload(r->wts, wts, s);
    upload(r->ptr, ptr, s);
    upload(r->ent, ent, s);
    ARH_HIP(hipStreamSynchronize(s));
}
#pragma clang fp contract(on)

void relief_eval(arh_relief *r, const double *tex, double *heights) {
    hipStream_t s = r->ctx->stream;
    hipLaunchKernelGGL(k_relief_eval, dim3((r->n + kRT - 1) / kRT), dim3(kRT), 0, s, r->n, r->idx.as<int4>(),
                       r->wts.as<double4>(), tex, heights);
    ARH_HIP(hipGetLastError());
}

void relief_adjoint(arh_relief *r, const double *grad, double *tex_grad) {
    hipStream_t s = r->ctx->stream;
    const int nt = r->h * r->w;
    hipLaunchKernelGGL(k_relief_adjoint, dim3((nt + kRT - 1) / kRT), dim3(kRT), 0, s, nt, r->ptr.as<int>(),
                       r->ent.as<int>(), (const double *)r->wts.p, grad, tex_grad);
    ARH_HIP(hipGetLastError());
}

}  // namespace arh
