// Forward H-matrix product in complex FP64 for few right-hand sides (HBM-bound), y = (M - I/2) x.
//
// Phase 1, one wavefront per (low-rank block, 4 rank rows): t_b[p] = V_b[p, :] x[cs:ce]; lanes stride
// the columns, the 4 rows share each x load and keep 4 loads in flight.
// Phase 2, one wavefront per work unit: a unit is up to 64 consecutive columns of one row leaf's
// column list (its dense panel, or the leaf's segment of one low-rank block's U). Lanes cover
// (row, column) pairs of the column-major tile so every load is one contiguous run; the unit's
// partial sums go to its own slot.
// Phase 3, one workgroup per leaf: y = -x/2 + the leaf's unit partials in unit order.
// No atomics anywhere: every output has one owner and a fixed summation order.
#include <algorithm>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

constexpr int kUnitCols = 64;  // columns per phase-2 unit
constexpr int kRC64 = 4;       // rhs per pass

// ---------------------------------------------------------------- phase 1
__global__ __launch_bounds__(256) void k_lr_t64(const LRBlock *__restrict__ blocks, const int2 *__restrict__ units,
                                                int nunits, const double2 *__restrict__ V, const double2 *__restrict__ x,
                                                int N, int R, double2 *__restrict__ t) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int u = blockIdx.x * 4 + w;
    if (u >= nunits) return;
    const int2 un = units[u];
    const LRBlock B = blocks[un.x];
    const int n = B.ce - B.cs, r = B.rank;
    const int p0 = un.y * 4;
    const int np = min(4, r - p0);
    const double2 *v0 = V + B.v_off + (int64_t)p0 * n;
    for (int j0 = 0; j0 < R; j0 += kRC64) {
        double2 acc[4][kRC64];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int jj = 0; jj < kRC64; jj++) acc[q][jj] = make_double2(0.0, 0.0);
        for (int c = lane; c < n; c += 64) {
            double2 xv[kRC64];
#pragma unroll
            for (int jj = 0; jj < kRC64; jj++)
                xv[jj] = (j0 + jj < R) ? x[(int64_t)(j0 + jj) * N + B.cs + c] : make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (q < np) {
                    const double2 a = v0[(int64_t)q * n + c];
#pragma unroll
                    for (int jj = 0; jj < kRC64; jj++) cfma(acc[q][jj], a, xv[jj]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int jj = 0; jj < kRC64; jj++) {
                const double2 s = wave_sum2(acc[q][jj]);
                if (lane == 0 && q < np && j0 + jj < R) t[(B.t_off + p0 + q) * R + j0 + jj] = s;
            }
    }
}

// ---------------------------------------------------------------- phase 2
__global__ __launch_bounds__(256) void k_rows64(const MvUnit *__restrict__ units, int nunits,
                                                const int *__restrict__ leaf_start, const int *__restrict__ pcol_ptr,
                                                const int64_t *__restrict__ panel_off, const int *__restrict__ pcol_idx,
                                                const double2 *__restrict__ panel, const LRBlock *__restrict__ blocks,
                                                const double2 *__restrict__ U, const double2 *__restrict__ t,
                                                const double2 *__restrict__ x, int N, int R, int slot,
                                                double2 *__restrict__ part) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int u = blockIdx.x * 4 + w;
    if (u >= nunits) return;
    const MvUnit un = units[u];
    const int rs = leaf_start[un.leaf], m = leaf_start[un.leaf + 1] - rs;
    const double2 *A;
    const int *cidx = nullptr;
    const double2 *tb = nullptr;
    if (un.kind == 0) {
        A = panel + panel_off[un.leaf];
        cidx = pcol_idx + pcol_ptr[un.leaf];
    } else {
        const LRBlock &B = blocks[un.blk];
        A = U + B.u_off + (int64_t)(rs - B.rs) * B.rank;
        tb = t + B.t_off * R;
    }
    double2 *pu = part + (int64_t)u * slot * R;  // slot: m (<= slot) x R partial sums
    for (int rc = 0; rc < m; rc += 64) {
        const int mm = min(64, m - rc);
        const int cpl = 64 / mm;
        const int cc = lane / mm, i = lane - cc * mm;
        const bool on = cc < cpl;
        for (int j0 = 0; j0 < R; j0 += kRC64) {
            double2 acc[kRC64];
#pragma unroll
            for (int jj = 0; jj < kRC64; jj++) acc[jj] = make_double2(0.0, 0.0);
            if (on) {
                const int c1 = un.c0 + un.nc;
                int c = un.c0 + cc;
                // 4 columns in flight per lane
                for (; c + 3 * cpl < c1; c += 4 * cpl) {
                    double2 a[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) a[q] = A[(int64_t)(c + q * cpl) * m + rc + i];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int cq = c + q * cpl;
#pragma unroll
                        for (int jj = 0; jj < kRC64; jj++)
                            if (j0 + jj < R) {
                                const double2 b = cidx ? x[(int64_t)(j0 + jj) * N + cidx[cq]] : tb[(int64_t)cq * R + j0 + jj];
                                cfma(acc[jj], a[q], b);
                            }
                    }
                }
                for (; c < c1; c += cpl) {
                    const double2 a = A[(int64_t)c * m + rc + i];
#pragma unroll
                    for (int jj = 0; jj < kRC64; jj++)
                        if (j0 + jj < R) {
                            const double2 b = cidx ? x[(int64_t)(j0 + jj) * N + cidx[c]] : tb[(int64_t)c * R + j0 + jj];
                            cfma(acc[jj], a, b);
                        }
                }
            }
            // fold the cpl column groups of the wave into rows (fixed order, cross-lane shuffles)
            const int src0 = lane < mm ? lane : 0;
#pragma unroll
            for (int jj = 0; jj < kRC64; jj++) {
                if (j0 + jj >= R) break;
                double2 s = make_double2(0.0, 0.0);
                for (int q = 0; q < cpl; q++) {
                    const int src = q * mm + src0;
                    const double2 v = make_double2(__shfl(acc[jj].x, src, 64), __shfl(acc[jj].y, src, 64));
                    s = q == 0 ? v : cadd(s, v);
                }
                if (lane < mm) pu[(int64_t)(rc + lane) * R + j0 + jj] = s;
            }
        }
    }
}

// ---------------------------------------------------------------- phase 3
__global__ __launch_bounds__(256) void k_rows64_reduce(const int *__restrict__ uptr, const int *__restrict__ leaf_start,
                                                       const double2 *__restrict__ part, const double2 *__restrict__ x,
                                                       int N, int R, int slot, double2 *__restrict__ y) {
    const int l = blockIdx.x;
    const int rs = leaf_start[l], m = leaf_start[l + 1] - rs;
    const int u0 = uptr[l], u1 = uptr[l + 1];
    for (int e = threadIdx.x; e < m * R; e += blockDim.x) {
        const int i = e % m, j = e / m;
        const int64_t yo = (int64_t)j * N + rs + i;
        double2 s = make_double2(-0.5 * x[yo].x, -0.5 * x[yo].y);
        for (int u = u0; u < u1; u++) s = cadd(s, part[((int64_t)u * slot + i) * R + j]);
        y[yo] = s;
    }
}

// ---------------------------------------------------------------- host
static void build_units64(arh_hmat *h, hipStream_t s) {
    arh_geom *g = h->g;
    h->mv_units.clear();
    h->mv_uptr.assign(g->nleaf + 1, 0);
    for (int l = 0; l < g->nleaf; l++) {
        const int ncol = g->pcol_ptr[l + 1] - g->pcol_ptr[l];
        for (int c0 = 0; c0 < ncol; c0 += kUnitCols) h->mv_units.push_back(MvUnit{l, 0, -1, c0, std::min(kUnitCols, ncol - c0)});
        for (int e = g->lrR_ptr[l]; e < g->lrR_ptr[l + 1]; e++) {
            const int b = g->lrR_blk[e], r = h->hblocks[b].rank;
            for (int c0 = 0; c0 < r; c0 += kUnitCols) h->mv_units.push_back(MvUnit{l, 1, b, c0, std::min(kUnitCols, r - c0)});
        }
        h->mv_uptr[l + 1] = (int)h->mv_units.size();
    }
    h->p1units64.clear();
    for (int a : h->active_lr)
        for (int q = 0; q * 4 < h->hblocks[a].rank; q++) h->p1units64.push_back(int2{a, q});
    upload(h->d_mv_units, h->mv_units, s);
    upload(h->d_mv_uptr, h->mv_uptr, s);
    upload(h->d_p1units64, h->p1units64, s);
}

void hmat_apply64(arh_hmat *h, int R, const double2 *x, double2 *y, hipStream_t s) {
    arh_geom *g = h->g;
    Profiler &pf = g->ctx->prof;
    const int N = g->N;
    if (h->mv_uptr.empty()) build_units64(h, s);
    h->tbuf.alloc(sizeof(double2) * std::max<int64_t>(1, h->t_total * R));
    const int nu = (int)h->mv_units.size();
    const int slot = g->max_leaf;
    h->mv_part.alloc(sizeof(double2) * (size_t)std::max(1, nu) * slot * R);
    double2 *t = h->tbuf.as<double2>();
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
    const int n1 = (int)h->p1units64.size();
    if (n1)
        timed(PROF_LR64, 16.0 * h->v_total + 16.0 * R * (N + h->t_total), 8.0 * R * h->v_total, [&] {
            hipLaunchKernelGGL(k_lr_t64, dim3((n1 + 3) / 4), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                               h->d_p1units64.as<int2>(), n1, h->V.as<double2>(), x, N, R, t);
        });
    const int64_t ent = g->dense_entries + h->u_total;
    timed(PROF_ROWS64, 16.0 * ent + 16.0 * R * (2.0 * N + h->t_total), 8.0 * R * ent, [&] {
        if (nu)
            hipLaunchKernelGGL(k_rows64, dim3((nu + 3) / 4), dim3(256), 0, s, h->d_mv_units.as<MvUnit>(), nu,
                               g->d_leaf_start.as<int>(), g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(),
                               g->d_pcol_idx.as<int>(), h->dense.as<double2>(), h->blocks.as<LRBlock>(),
                               h->U.as<double2>(), t, x, N, R, slot, h->mv_part.as<double2>());
        hipLaunchKernelGGL(k_rows64_reduce, dim3(g->nleaf), dim3(256), 0, s, h->d_mv_uptr.as<int>(),
                           g->d_leaf_start.as<int>(), h->mv_part.as<double2>(), x, N, R, slot, y);
    });
    ARH_HIP(hipGetLastError());
}

}  // namespace arh
