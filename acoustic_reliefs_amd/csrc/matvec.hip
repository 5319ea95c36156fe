// H-matrix products y = (M - I/2) x and y = (M - I/2)^T x for nrhs vectors (hmatrix.h:84-140).
//
// Forward, two launches, no atomics (every output row has one owner, so the result does not
// depend on the batch it is computed in):
//   1. per low-rank block b:   t_b = V_b x[cs:ce]                      (r_b values per rhs)
//   2. per row leaf l:         y[l] = -x[l]/2 + P_l x[cols(P_l)] + sum_{b covers l} U_b[l,:] t_b
//      P_l is the leaf's dense panel (its dense blocks side by side, column-major), U_b[l,:] the
//      leaf's contiguous segment of U_b. Lanes cover (row, column) pairs of a column-major tile,
//      so every load of a wavefront is one contiguous run.
// Transposed (set_transposed, hmatrix.h:95-115):
//   1. per low-rank block b:   s_b = U_b^T x[rs:re]
//   2. per column leaf l:      y[l] = -x[l]/2 + sum_{dense (l', l)} D^T x[l'] + sum_{b covers l} V_b[:, l]^T s_b
//      Lanes take the columns of D, so D is read from a row-major copy of the panels (a row of a block is
//      then one contiguous run); the copy is made on the first transposed product.
#include <algorithm>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

constexpr int kMvThreads = 256;
constexpr int kMvWaves = kMvThreads / ARH_WAVE;
constexpr int RC = 4;  // right-hand sides per pass

template <class S>
struct CT;
template <>
struct CT<double2> {
    __device__ static double2 ld(const double2 *p) { return *p; }
};
template <>
struct CT<float2> {
    __device__ static double2 ld(const float2 *p) {
        float2 v = *p;
        return make_double2(v.x, v.y);
    }
};

// V entry (p, c) of a block: row-major (FP64 copy) or column-major (FP32 copy)
template <class S>
__device__ __forceinline__ const S *v_at(const S *V, const LRBlock &B, int p, int c, bool vcm) {
    return V + B.v_off + (vcm ? (int64_t)c * B.rank + p : (int64_t)p * (B.ce - B.cs) + c);
}

// ---------------------------------------------------------------- phase 1 (forward): t = V x
template <class S>
__global__ __launch_bounds__(kMvThreads) void k_lr_t(const LRBlock *__restrict__ blocks, const int *__restrict__ active,
                                                     const S *__restrict__ V, int vcm, const double2 *__restrict__ x,
                                                     int N, int R, double2 *__restrict__ t) {
    const LRBlock B = blocks[active[blockIdx.x]];
    const int n = B.ce - B.cs, r = B.rank;
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
    for (int p = w; p < r; p += kMvWaves) {
        for (int j0 = 0; j0 < R; j0 += RC) {
            double2 acc[RC];
#pragma unroll
            for (int q = 0; q < RC; q++) acc[q] = make_double2(0.0, 0.0);
            for (int c = lane; c < n; c += ARH_WAVE) {
                double2 v = CT<S>::ld(v_at(V, B, p, c, vcm));
#pragma unroll
                for (int q = 0; q < RC; q++)
                    if (j0 + q < R) cfma(acc[q], v, x[(int64_t)(j0 + q) * N + B.cs + c]);
            }
#pragma unroll
            for (int q = 0; q < RC; q++) {
                double2 s = wave_sum2(acc[q]);
                if (lane == 0 && j0 + q < R) t[(B.t_off + p) * R + j0 + q] = s;
            }
        }
    }
}

// ---------------------------------------------------------------- phase 1 (transposed): s = U^T x
template <class S>
__global__ __launch_bounds__(kMvThreads) void k_lr_s(const LRBlock *__restrict__ blocks, const int *__restrict__ active,
                                                     const int *__restrict__ leaf_start, const S *__restrict__ U,
                                                     const double2 *__restrict__ x, int N, int R,
                                                     double2 *__restrict__ t) {
    const LRBlock B = blocks[active[blockIdx.x]];
    const int r = B.rank;
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
    for (int p = w; p < r; p += kMvWaves) {
        for (int j0 = 0; j0 < R; j0 += RC) {
            double2 acc[RC];
#pragma unroll
            for (int q = 0; q < RC; q++) acc[q] = make_double2(0.0, 0.0);
            for (int l = B.l0; leaf_start[l] < B.re; l++) {
                const int ls = leaf_start[l], ms = leaf_start[l + 1] - ls;
                const S *col = U + B.u_off + (int64_t)(ls - B.rs) * r + (int64_t)p * ms;
                for (int i = lane; i < ms; i += ARH_WAVE) {
                    double2 u = CT<S>::ld(col + i);
#pragma unroll
                    for (int q = 0; q < RC; q++)
                        if (j0 + q < R) cfma(acc[q], u, x[(int64_t)(j0 + q) * N + ls + i]);
                }
            }
#pragma unroll
            for (int q = 0; q < RC; q++) {
                double2 s = wave_sum2(acc[q]);
                if (lane == 0 && j0 + q < R) t[(B.t_off + p) * R + j0 + q] = s;
            }
        }
    }
}

// ---------------------------------------------------------------- phase 2 (forward): one workgroup per row leaf
template <class S>
__global__ __launch_bounds__(kMvThreads) void k_rows(const int *__restrict__ leaf_start, const int *__restrict__ pcol_ptr,
                                                     const int64_t *__restrict__ panel_off,
                                                     const int *__restrict__ pcol_idx, const S *__restrict__ panel,
                                                     const int *__restrict__ lr_ptr, const int *__restrict__ lr_blk,
                                                     const LRBlock *__restrict__ blocks, const S *__restrict__ U,
                                                     const double2 *__restrict__ t, const double2 *__restrict__ x,
                                                     int N, int R, double2 *__restrict__ y) {
    __shared__ double2 red[kMvWaves * ARH_WAVE * RC];
    const int l = blockIdx.x;
    const int rs = leaf_start[l], m = leaf_start[l + 1] - rs;
    const int ncol = pcol_ptr[l + 1] - pcol_ptr[l];
    const S *P = panel + panel_off[l];
    const int *cidx = pcol_idx + pcol_ptr[l];
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
    const int b0 = lr_ptr[l], b1 = lr_ptr[l + 1];
    for (int rc = 0; rc < m; rc += ARH_WAVE) {
        const int mm = min(ARH_WAVE, m - rc);
        const int cpl = ARH_WAVE / mm;  // columns per wavefront pass
        const int cc = lane / mm, i = lane - cc * mm;
        const bool on = cc < cpl;
        const int step = kMvWaves * cpl;
        for (int j0 = 0; j0 < R; j0 += RC) {
            double2 acc[RC];
#pragma unroll
            for (int q = 0; q < RC; q++) acc[q] = make_double2(0.0, 0.0);
            if (on) {
                for (int c = w * cpl + cc; c < ncol; c += step) {
                    double2 a = CT<S>::ld(P + (int64_t)c * m + rc + i);
                    const int col = cidx[c];
#pragma unroll
                    for (int q = 0; q < RC; q++)
                        if (j0 + q < R) cfma(acc[q], a, x[(int64_t)(j0 + q) * N + col]);
                }
                for (int e = b0; e < b1; e++) {
                    const LRBlock &B = blocks[lr_blk[e]];
                    const int r = B.rank;
                    const S *seg = U + B.u_off + (int64_t)(rs - B.rs) * r;
                    for (int p = w * cpl + cc; p < r; p += step) {
                        double2 a = CT<S>::ld(seg + (int64_t)p * m + rc + i);
#pragma unroll
                        for (int q = 0; q < RC; q++)
                            if (j0 + q < R) cfma(acc[q], a, t[(B.t_off + p) * R + j0 + q]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RC; q++) red[(w * ARH_WAVE + lane) * RC + q] = acc[q];
            __syncthreads();
            if (threadIdx.x < mm * RC) {
                const int ii = threadIdx.x % mm, q = threadIdx.x / mm;
                if (j0 + q < R) {
                    const int64_t yo = (int64_t)(j0 + q) * N + rs + rc + ii;
                    double2 s = make_double2(-0.5 * x[yo].x, -0.5 * x[yo].y);
                    for (int ww = 0; ww < kMvWaves; ww++)
                        for (int c2 = 0; c2 < cpl; c2++) {
                            double2 v = red[(ww * ARH_WAVE + c2 * mm + ii) * RC + q];
                            s.x += v.x, s.y += v.y;
                        }
                    y[yo] = s;
                }
            }
        }
    }
}

// Row-major copy of each leaf's panel: out[off + i ncol + c] = in[off + c m + i] (one workgroup per leaf).
template <class S>
__global__ __launch_bounds__(kMvThreads) void k_panel_rowmajor(const int *__restrict__ leaf_start,
                                                               const int *__restrict__ pcol_ptr,
                                                               const int64_t *__restrict__ panel_off,
                                                               const S *__restrict__ in, S *__restrict__ out) {
    const int l = blockIdx.x;
    const int m = leaf_start[l + 1] - leaf_start[l], ncol = pcol_ptr[l + 1] - pcol_ptr[l];
    const int64_t off = panel_off[l], tot = (int64_t)m * ncol;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const int i = (int)(e / ncol), c = (int)(e - (int64_t)i * ncol);
        out[off + e] = in[off + (int64_t)c * m + i];
    }
}

// ------------------------------------------------------ phase 2 (transposed): workgroup (column leaf, 64 columns)
// W wavefronts take turns over the leaf's dense blocks and over the ranks of its low-rank blocks.
constexpr int kColWaves = 16;
template <class S, int W>
__global__ __launch_bounds__(64 * W) void k_cols(const int *__restrict__ leaf_start, const int *__restrict__ dT_ptr,
                                                     const int *__restrict__ dT_blk, const int *__restrict__ dblk_info,
                                                     const int *__restrict__ pcol_ptr,
                                                     const int64_t *__restrict__ panel_off,
                                                     const S *__restrict__ panelT,
                                                     const int *__restrict__ lr_ptr, const int *__restrict__ lr_blk,
                                                     const LRBlock *__restrict__ blocks, const S *__restrict__ V,
                                                     int vcm, const double2 *__restrict__ t,
                                                     const double2 *__restrict__ x, int N, int R,
                                                     double2 *__restrict__ y) {
    __shared__ double2 red[W * ARH_WAVE * RC];
    const int l = blockIdx.x;
    const int cs = leaf_start[l], n = leaf_start[l + 1] - cs;
    const int lane = threadIdx.x % ARH_WAVE, w = threadIdx.x / ARH_WAVE;
    const int d0 = dT_ptr[l], d1 = dT_ptr[l + 1];
    const int b0 = lr_ptr[l], b1 = lr_ptr[l + 1];
    {
        const int c0 = (int)blockIdx.y * ARH_WAVE;
        if (c0 >= n) return;
        const int c = c0 + lane;
        const bool on = c < n;
        for (int j0 = 0; j0 < R; j0 += RC) {
            double2 acc[RC];
#pragma unroll
            for (int q = 0; q < RC; q++) acc[q] = make_double2(0.0, 0.0);
            if (on) {
                for (int e = d0 + w; e < d1; e += W) {
                    const int b = dT_blk[e];
                    const int lr = dblk_info[2 * b], pc = dblk_info[2 * b + 1];
                    const int rs = leaf_start[lr], m = leaf_start[lr + 1] - rs;
                    const int64_t ncol = pcol_ptr[lr + 1] - pcol_ptr[lr];
                    const S *D = panelT + panel_off[lr] + pc + c;  // row i of the block at D + i ncol
                    for (int i = 0; i < m; i++) {
                        double2 a = CT<S>::ld(D + i * ncol);
#pragma unroll
                        for (int q = 0; q < RC; q++)
                            if (j0 + q < R) cfma(acc[q], a, x[(int64_t)(j0 + q) * N + rs + i]);
                    }
                }
                for (int e = b0; e < b1; e++) {
                    const LRBlock &B = blocks[lr_blk[e]];
                    const int r = B.rank;
                    for (int p = w; p < r; p += W) {
                        double2 a = CT<S>::ld(v_at(V, B, p, (cs - B.cs) + c, vcm));
#pragma unroll
                        for (int q = 0; q < RC; q++)
                            if (j0 + q < R) cfma(acc[q], a, t[(B.t_off + p) * R + j0 + q]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RC; q++) red[(w * ARH_WAVE + lane) * RC + q] = acc[q];
            __syncthreads();
            const int nn = min(ARH_WAVE, n - c0);
            if (threadIdx.x < nn * RC) {
                const int cc = threadIdx.x % nn, q = threadIdx.x / nn;
                if (j0 + q < R) {
                    const int64_t yo = (int64_t)(j0 + q) * N + cs + c0 + cc;
                    double2 s = make_double2(-0.5 * x[yo].x, -0.5 * x[yo].y);
                    for (int ww = 0; ww < W; ww++) {
                        double2 v = red[(ww * ARH_WAVE + cc) * RC + q];
                        s.x += v.x, s.y += v.y;
                    }
                    y[yo] = s;
                }
            }
            __syncthreads();
        }
    }
}

template <class S>
static void apply_t(arh_hmat *h, int transpose, int R, const S *panel, const S *U, const S *V, const double2 *x,
                    double2 *y, hipStream_t s) {
    arh_geom *g = h->g;
    Profiler &pf = g->ctx->prof;
    const int N = g->N;
    h->tbuf.alloc(sizeof(double2) * std::max<int64_t>(1, h->t_total * R));
    double2 *t = h->tbuf.as<double2>();
    const int nact = (int)h->active_lr.size();
    const double es = (double)sizeof(S), vs = 16.0 * R;
    const int64_t lr_entries = transpose ? h->u_total : h->v_total;   // phase-1 factor
    const int64_t rw_entries = g->dense_entries + (transpose ? h->v_total : h->u_total);
    const bool fp32 = sizeof(S) == sizeof(float2);
    auto timed = [&](int kind, double bytes, double flops, auto &&launch) {
        if (!pf.on) {
            launch();
            return;
        }
        Profiler::Entry e{pf.get(), pf.get(), kind, bytes, flops};
        ARH_HIP(hipEventRecord(e.a, s));
        launch();
        ARH_HIP(hipEventRecord(e.b, s));
        pf.pending.push_back(e);
    };
    if (nact)
        timed(fp32 ? PROF_LR32 : PROF_LR64, es * lr_entries + vs * (N + h->t_total), 8.0 * R * lr_entries, [&] {
            if (!transpose)
                hipLaunchKernelGGL(k_lr_t<S>, dim3(nact), dim3(kMvThreads), 0, s, h->blocks.as<LRBlock>(),
                                   h->d_active_lr.as<int>(), V, (int)fp32, x, N, R, t);
            else
                hipLaunchKernelGGL(k_lr_s<S>, dim3(nact), dim3(kMvThreads), 0, s, h->blocks.as<LRBlock>(),
                                   h->d_active_lr.as<int>(), g->d_leaf_start.as<int>(), U, x, N, R, t);
        });
    timed(fp32 ? PROF_ROWS32 : PROF_ROWS64, es * rw_entries + vs * (2.0 * N + h->t_total), 8.0 * R * rw_entries, [&] {
        if (!transpose)
            hipLaunchKernelGGL(k_rows<S>, dim3(g->nleaf), dim3(kMvThreads), 0, s, g->d_leaf_start.as<int>(),
                               g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(), g->d_pcol_idx.as<int>(), panel,
                               g->d_lrR_ptr.as<int>(), g->d_lrR_blk.as<int>(), h->blocks.as<LRBlock>(), U, t, x, N, R,
                               y);
        else
            hipLaunchKernelGGL((k_cols<S, kColWaves>), dim3(g->nleaf, (g->max_leaf + ARH_WAVE - 1) / ARH_WAVE),
                               dim3(64 * kColWaves), 0, s,
                               g->d_leaf_start.as<int>(), g->d_dT_ptr.as<int>(), g->d_dT_blk.as<int>(),
                               g->d_dblk_info.as<int>(), g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(), panel,
                               g->d_lrC_ptr.as<int>(), g->d_lrC_blk.as<int>(), h->blocks.as<LRBlock>(), V, (int)fp32,
                               t, x, N, R, y);
    });
    ARH_HIP(hipGetLastError());
}

// The row-major panel copy for the transposed product (made once per H-matrix and precision).
template <class S>
static const S *rowmajor_panels(arh_hmat *h, DBuf &src, DBuf &dst, hipStream_t s) {
    arh_geom *g = h->g;
    if (!dst.p) {
        dst.alloc(sizeof(S) * std::max<int64_t>(1, g->dense_entries));
        hipLaunchKernelGGL(k_panel_rowmajor<S>, dim3(g->nleaf), dim3(kMvThreads), 0, s, g->d_leaf_start.as<int>(),
                           g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(), src.as<S>(), dst.as<S>());
        ARH_HIP(hipGetLastError());
    }
    return dst.as<S>();
}

void hmat_apply(arh_hmat *h, int transpose, int precision, int nrhs, const double2 *x, double2 *y, hipStream_t s,
                const int *cols, int ncols, const int *gate) {
    ARH_CHECK(nrhs >= 1, ARH_EINVAL, "nrhs must be >= 1");
    if (precision == ARH_MIXED) {
        ARH_CHECK(h->storage == ARH_MIXED, ARH_ESTATE, "FP32 product requested but the H-matrix has no FP32 copy");
        // every product takes the MFMA engine, whatever nrhs: a right-hand side then gets the same
        // bits alone and in any batch
        if (h->g->max_leaf <= mfma_max_leaf()) {
            hmat_apply_mfma(h, transpose, 0, nrhs, x, y, s, cols, ncols, gate);
            return;
        }
        make_generic_fp32(h);
        const float2 *panel = h->dense32.as<float2>();
        if (transpose) panel = rowmajor_panels<float2>(h, h->dense32, h->denseT32, s);
        apply_t<float2>(h, transpose, nrhs, panel, h->U32.as<float2>(), h->V32.as<float2>(), x, y, s);
    } else {
        if (h->g->max_leaf <= mfma_max_leaf()) {
            hmat_apply_mfma(h, transpose, 1, nrhs, x, y, s, cols, ncols, gate);
            return;
        }
        if (!transpose) {
            hmat_apply64(h, nrhs, x, y, s);
            return;
        }
        apply_t<double2>(h, transpose, nrhs, rowmajor_panels<double2>(h, h->dense, h->denseT, s), h->U.as<double2>(),
                         h->V.as<double2>(), x, y, s);
    }
}

}  // namespace arh
