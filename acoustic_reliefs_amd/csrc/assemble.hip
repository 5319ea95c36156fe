// H-matrix assembly on gfx950: dense near-field panels and batched partially pivoted ACA.
#include <algorithm>
#include <numeric>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

constexpr int kAcaThreads = 256;
constexpr int kAcaWaves = kAcaThreads / ARH_WAVE;
constexpr int kMaxRankLds = 128;  // max ACA rank kept in LDS coefficient arrays
#ifndef ARH_ACA_BIGDIM
#define ARH_ACA_BIGDIM 512
#endif
constexpr int kAcaBigDim = ARH_ACA_BIGDIM;  // blocks with m + n at least this get a workgroup, smaller ones a wavefront

// ------------------------------------------------------------------ dense panels
// One thread per entry of a (row leaf, column chunk) task; entries of a column are contiguous,
// so the store of a task is one coalesced run of nc * m complex values (compute_block,
// integrate.cpp:87-100, column by column).
__global__ __launch_bounds__(256) void k_dense(const DenseTask *__restrict__ tasks, const int *__restrict__ leaf_start,
                                               const int *__restrict__ pcol_ptr, const int64_t *__restrict__ panel_off,
                                               const int *__restrict__ pcol_idx, const Tri *__restrict__ tri, double k,
                                               double2 *__restrict__ out, int *__restrict__ nanflag) {
    const DenseTask t = tasks[blockIdx.x];
    const int rs = leaf_start[t.leaf];
    const int m = leaf_start[t.leaf + 1] - rs;
    const int tid = threadIdx.x;
    if (tid >= t.nc * m) return;
    const int cc = tid / m, i = tid - cc * m;
    const int c = t.c0 + cc;
    const int j = pcol_idx[pcol_ptr[t.leaf] + c];
    const Tri p = tri[rs + i];
    const Tri q = tri[j];
    double2 v = bem_entry(k, p.gx, p.gy, p.gz, q);
    if (isnan(v.x) || isnan(v.y)) atomicOr(nanflag, 1);
    out[panel_off[t.leaf] + (int64_t)c * m + i] = v;
}

// ------------------------------------------------------------------ ACA cooperation groups
struct AcaShared {
    double2 coef[kMaxRankLds];
    double du[kMaxRankLds], dv[kMaxRankLds];
    double rv[kAcaWaves];
    int ri[kAcaWaves];
};

// The threads working on one ACA block: the whole 256-thread workgroup (WAVE false) or a single
// wavefront (WAVE true, no workgroup barrier: a wavefront's memory operations complete in order,
// so a fence at wavefront scope, which also stops compiler reordering, is its barrier).
template <bool WAVE>
struct AcaGroup {
    int tid, nthr, w, nw, lane;
    AcaShared &sh;
    __device__ __forceinline__ void sync() const {
        if (WAVE) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        } else {
            __syncthreads();
        }
    }
    // fixed-order sums: all threads get the same bits
    __device__ __forceinline__ double sum(double v) const {
        v = wave_sum(v);
        if (WAVE) return v;
        sync();
        if (lane == 0) sh.rv[w] = v;
        sync();
        double s = sh.rv[0];
#pragma unroll
        for (int q = 1; q < kAcaWaves; q++) s += sh.rv[q];
        return s;
    }
    // (value, index) max with ties to the lower index
    __device__ __forceinline__ void argmax(double &v, int &i) const {
        wave_argmax(v, i);
        if (WAVE) return;
        sync();
        if (lane == 0) sh.rv[w] = v, sh.ri[w] = i;
        sync();
        double bv = sh.rv[0];
        int bi = sh.ri[0];
#pragma unroll
        for (int q = 1; q < kAcaWaves; q++)
            if (sh.rv[q] > bv || (sh.rv[q] == bv && sh.ri[q] < bi)) bv = sh.rv[q], bi = sh.ri[q];
        v = bv;
        i = bi;
    }
};

// ------------------------------------------------------------------ ACA of one block
// compute_ACA_block (aca.cpp:38-82): start at row 0; residual row, first-max pivot column,
// normalised row (unnormalised if the pivot is 0), residual column, non-conjugated norm
// estimate, stop when sqrt(|u|^2|v|^2 / err) <= eps discarding the current cross, next row =
// first max of |u| over unused rows; rank limit min(max_rank, m-1, n-1).
// Crosses whose u or v is exactly zero add nothing to any later quantity, so they are counted
// (rank_ref) but not stored (rank).
template <bool WAVE>
__device__ __forceinline__ void aca_block(const AcaGroup<WAVE> &G, LRBlock *__restrict__ blocks, int bid,
                                          const Tri *__restrict__ tri, double k, double eps, double2 *__restrict__ Us,
                                          double2 *__restrict__ Vs, unsigned char *__restrict__ flags) {
    AcaShared &sh = G.sh;
    const LRBlock B = blocks[bid];
    const int rs = B.rs, cs = B.cs, m = B.re - B.rs, n = B.ce - B.cs, lim = B.lim;
    const int tid = G.tid, nthr = G.nthr, lane = G.lane;
    if (lim <= 0) {
        if (tid == 0) blocks[bid].rank = 0, blocks[bid].rank_ref = 0;
        return;
    }
    double2 *U = Us + B.us_off;
    double2 *V = Vs + B.vs_off;
    unsigned char *used = flags + B.flag_off;
    for (int q = tid; q < m; q += nthr) used[q] = (q == 0) ? 1 : 0;
    int i = 0, it = 0, sr = 0;
    double err = 0.0;
    G.sync();
    while (it < lim) {
        // coefficients U(i, p), p < sr
        for (int p = tid; p < sr; p += nthr) sh.coef[p] = U[(int64_t)p * m + i];
        G.sync();
        // residual row into V row `sr`, with the first-max pivot search
        const Tri ti = tri[rs + i];
        double2 *vrow = V + (int64_t)sr * n;
        double best = -1.0;
        int bi = 0x7fffffff;
        for (int c = tid; c < n; c += nthr) {
            double2 v = bem_entry(k, ti.gx, ti.gy, ti.gz, tri[cs + c]);
            for (int p = 0; p < sr; p++) {
                double2 vp = V[(int64_t)p * n + c];
                double2 up = sh.coef[p];
                v.x -= up.x * vp.x - up.y * vp.y;
                v.y -= up.x * vp.y + up.y * vp.x;
            }
            vrow[c] = v;
            double a = cnorm2(v);
            if (a > best) best = a, bi = c;
        }
        G.argmax(best, bi);
        const int j = (best < 0.0) ? 0 : bi;
        G.sync();
        const double2 piv = vrow[j];
        // every thread holds the pivot before its owner normalises it in place
        G.sync();
        const bool nz = (piv.x != 0.0 || piv.y != 0.0);
        double vn = 0.0;
        for (int c = tid; c < n; c += nthr) {
            double2 v = vrow[c];
            if (nz) v = cdiv(v, piv);
            vrow[c] = v;
            vn += cnorm2(v);
        }
        vn = G.sum(vn);
        for (int p = tid; p < sr; p += nthr) sh.coef[p] = V[(int64_t)p * n + j];
        G.sync();
        // residual column into U column `sr`, with the masked first-max search for the next row
        const Tri tj = tri[cs + j];
        double2 *ucol = U + (int64_t)sr * m;
        double un = 0.0, bu = -1.0;
        int bq = 0x7fffffff;
        for (int q = tid; q < m; q += nthr) {
            const Tri tq = tri[rs + q];
            double2 u = bem_entry(k, tq.gx, tq.gy, tq.gz, tj);
            for (int p = 0; p < sr; p++) {
                double2 up = U[(int64_t)p * m + q];
                double2 vp = sh.coef[p];
                u.x -= up.x * vp.x - up.y * vp.y;
                u.y -= up.x * vp.y + up.y * vp.x;
            }
            ucol[q] = u;
            double a = cnorm2(u);
            un += a;
            if (!used[q] && a > bu) bu = a, bq = q;
        }
        un = G.sum(un);
        G.argmax(bu, bq);
        G.sync();
        // |u_p^T u| and |v v_p^T| for stored crosses p < sr (aca.cpp:64-69): wave w takes p = w mod nw
        for (int p = G.w; p < sr; p += G.nw) {
            double2 su = make_double2(0.0, 0.0), sv = make_double2(0.0, 0.0);
            const double2 *up = U + (int64_t)p * m;
            const double2 *vp = V + (int64_t)p * n;
            for (int q = lane; q < m; q += ARH_WAVE) cfma(su, ucol[q], up[q]);
            for (int c = lane; c < n; c += ARH_WAVE) cfma(sv, vrow[c], vp[c]);
            su = wave_sum2(su);
            sv = wave_sum2(sv);
            if (lane == 0) sh.du[p] = cabs_(su), sh.dv[p] = cabs_(sv);
        }
        G.sync();
        double dots = 0.0;
        for (int p = 0; p < sr; p++) dots += sh.du[p] * sh.dv[p];
        err += un * vn + 2.0 * dots;
        if (sqrt(un * vn / err) <= eps) break;
        if (un * vn > 0.0) sr++;
        i = (bu < 0.0) ? 0 : bq;
        G.sync();
        if (tid == 0) used[i] = 1;
        it++;
        G.sync();
    }
    if (tid == 0) blocks[bid].rank = sr, blocks[bid].rank_ref = it;
}

// Workgroups [0, nbig) each take one big block (order[0..nbig)); the others take four small
// blocks, one per wavefront.
__global__ __launch_bounds__(kAcaThreads) void k_aca(LRBlock *__restrict__ blocks, const int *__restrict__ order,
                                                     int nbig, int na, const Tri *__restrict__ tri, double k,
                                                     double eps, double2 *__restrict__ Us, double2 *__restrict__ Vs,
                                                     unsigned char *__restrict__ flags) {
    __shared__ AcaShared sh[kAcaWaves];
    const int lane = threadIdx.x % ARH_WAVE;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / ARH_WAVE));
    if ((int)blockIdx.x < nbig) {
        const AcaGroup<false> G{(int)threadIdx.x, kAcaThreads, w, kAcaWaves, lane, sh[0]};
        aca_block<false>(G, blocks, order[blockIdx.x], tri, k, eps, Us, Vs, flags);
        return;
    }
    const int u = nbig + ((int)blockIdx.x - nbig) * kAcaWaves + w;
    if (u >= na) return;
    const AcaGroup<true> G{lane, ARH_WAVE, 0, 1, lane, sh[w]};
    aca_block<true>(G, blocks, order[u], tri, k, eps, Us, Vs, flags);
}

// ------------------------------------------------------------------ compaction
// U scratch (m x lim) -> leaf-segmented packed U: for each row leaf [ls, le) of the block, an
// (le-ls) x rank column-major chunk at u_off + (ls - rs) * rank. V scratch -> rank x n row-major.
__global__ __launch_bounds__(256) void k_pack(const LRBlock *__restrict__ blocks, const int *__restrict__ active,
                                              const int *__restrict__ leaf_start, const double2 *__restrict__ Us,
                                              const double2 *__restrict__ Vs, double2 *__restrict__ U,
                                              double2 *__restrict__ V) {
    const LRBlock B = blocks[active[blockIdx.x]];
    const int m = B.re - B.rs, n = B.ce - B.cs, r = B.rank;
    const double2 *us = Us + B.us_off;
    const double2 *vs = Vs + B.vs_off;
    double2 *u = U + B.u_off;
    double2 *v = V + B.v_off;
    const int64_t nv = (int64_t)r * n;
    for (int64_t e = threadIdx.x; e < nv; e += blockDim.x) v[e] = vs[e];
    for (int l = B.l0; leaf_start[l] < B.re; l++) {
        const int ls = leaf_start[l] - B.rs, ms = leaf_start[l + 1] - leaf_start[l];
        double2 *seg = u + (int64_t)ls * r;
        for (int e = threadIdx.x; e < ms * r; e += blockDim.x) {
            const int p = e / ms, i = e - p * ms;
            seg[e] = us[(int64_t)p * m + ls + i];
        }
    }
}

// FP32 copy of V, transposed per block to column-major (n x r, V32[v_off + c r + p]): the MFMA
// product reads 16 consecutive rank rows of one column.
__global__ __launch_bounds__(256) void k_v_to_c64_cm(const LRBlock *__restrict__ blocks, const int *__restrict__ active,
                                                     const double2 *__restrict__ V, float2 *__restrict__ V32) {
    const LRBlock B = blocks[active[blockIdx.x]];
    const int n = B.ce - B.cs, r = B.rank;
    const int64_t tot = (int64_t)r * n;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const int c = (int)(e / r), p = (int)(e - (int64_t)c * r);
        const double2 v = V[B.v_off + (int64_t)p * n + c];
        V32[B.v_off + e] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ void k_to_c64(const double2 *__restrict__ a, float2 *__restrict__ b, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double2 v = a[e];
        b[e] = make_float2((float)v.x, (float)v.y);
    }
}

// ------------------------------------------------------------------ host drivers
void assemble_dense(arh_hmat *h) {
    arh_geom *g = h->g;
    hipStream_t s = g->ctx->stream;
    h->dense.alloc(sizeof(double2) * std::max<int64_t>(1, g->dense_entries));
    DBuf flag;
    flag.alloc(sizeof(int));
    ARH_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    if (!g->dtasks.empty()) {
        hipLaunchKernelGGL(k_dense, dim3((unsigned)g->dtasks.size()), dim3(256), 0, s, g->d_dtasks.as<DenseTask>(),
                           g->d_leaf_start.as<int>(), g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(),
                           g->d_pcol_idx.as<int>(), g->tri.as<Tri>(), h->k, h->dense.as<double2>(), flag.as<int>());
        ARH_HIP(hipGetLastError());
    }
    int nan = 0;
    ARH_HIP(hipMemcpyAsync(&nan, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ARH_HIP(hipStreamSynchronize(s));
    ARH_CHECK(!nan, ARH_ENAN, "NaN in a dense near-field block (a collocation point coincides with a quadrature point)");
}

static void ensure_aca_scratch(arh_geom *g, int max_rank) {
    if (g->scratch_rank == max_rank) return;
    const int na = g->na;
    g->lim.assign(na, 0);
    g->us_off.assign(na, 0);
    g->vs_off.assign(na, 0);
    g->flag_off.assign(na, 0);
    int64_t uo = 0, vo = 0, fo = 0;
    for (int a = 0; a < na; a++) {
        int m = g->lowrank[4 * a + 1] - g->lowrank[4 * a], n = g->lowrank[4 * a + 3] - g->lowrank[4 * a + 2];
        int lim = std::min(max_rank, std::min(m - 1, n - 1));
        lim = std::max(lim, 0);
        g->lim[a] = lim;
        g->us_off[a] = uo, g->vs_off[a] = vo, g->flag_off[a] = fo;
        uo += (int64_t)m * lim, vo += (int64_t)lim * n, fo += m;
    }
    g->aca_U.alloc(sizeof(double2) * std::max<int64_t>(1, uo));
    g->aca_V.alloc(sizeof(double2) * std::max<int64_t>(1, vo));
    g->aca_flags.alloc(std::max<int64_t>(16, fo));
    g->aca_u_total = uo, g->aca_v_total = vo, g->aca_f_total = fo;
    g->scratch_rank = max_rank;
}

void assemble_aca(arh_hmat *h) {
    arh_geom *g = h->g;
    hipStream_t s = g->ctx->stream;
    const int na = g->na;
    h->hblocks.assign(na, LRBlock{});
    h->active_lr.clear();
    h->u_total = h->v_total = h->t_total = 0;
    if (na == 0) return;
    ARH_CHECK(h->max_rank <= kMaxRankLds, ARH_EINVAL, "max_rank above the supported 128");
    ensure_aca_scratch(g, h->max_rank);
    for (int a = 0; a < na; a++) {
        LRBlock &B = h->hblocks[a];
        B.rs = g->lowrank[4 * a], B.re = g->lowrank[4 * a + 1], B.cs = g->lowrank[4 * a + 2];
        B.ce = g->lowrank[4 * a + 3];
        B.lim = g->lim[a];
        B.l0 = g->leaf_of[B.rs];
        B.us_off = g->us_off[a], B.vs_off = g->vs_off[a], B.flag_off = g->flag_off[a];
    }
    std::vector<int> order;  // kept until the synchronisation below (the upload may read it late)
    if (g->aca_order_rank != h->max_rank) {
        order.resize(na);
        for (int a = 0; a < na; a++) order[a] = a;
        // largest work first: the ACA cost grows like lim^2 (m + n)
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            const LRBlock &X = h->hblocks[x], &Y = h->hblocks[y];
            double wx = (double)X.lim * X.lim * (X.re - X.rs + X.ce - X.cs) + (X.re - X.rs + X.ce - X.cs);
            double wy = (double)Y.lim * Y.lim * (Y.re - Y.rs + Y.ce - Y.cs) + (Y.re - Y.rs + Y.ce - Y.cs);
            return wx > wy;
        });
        // blocks with m + n >= kAcaBigDim get a workgroup each (they lead the order, largest first);
        // the rest one wavefront each
        std::stable_partition(order.begin(), order.end(), [&](int x) {
            const LRBlock &X = h->hblocks[x];
            return X.re - X.rs + X.ce - X.cs >= kAcaBigDim;
        });
        int nbig = 0;
        while (nbig < na && h->hblocks[order[nbig]].re - h->hblocks[order[nbig]].rs + h->hblocks[order[nbig]].ce -
                                    h->hblocks[order[nbig]].cs >=
                                kAcaBigDim)
            nbig++;
        upload(g->d_aca_order, order, s);
        g->aca_nbig = nbig;
        g->aca_order_rank = h->max_rank;
    }
    const int nbig = g->aca_nbig;
    upload(h->blocks, h->hblocks, s);
    const int grid = nbig + (na - nbig + kAcaWaves - 1) / kAcaWaves;
    hipLaunchKernelGGL(k_aca, dim3(grid), dim3(kAcaThreads), 0, s, h->blocks.as<LRBlock>(), g->d_aca_order.as<int>(),
                       nbig, na, g->tri.as<Tri>(), h->k, h->aca_tol, g->aca_U.as<double2>(), g->aca_V.as<double2>(),
                       g->aca_flags.as<unsigned char>());
    ARH_HIP(hipGetLastError());
    ARH_HIP(hipMemcpyAsync(h->hblocks.data(), h->blocks.p, sizeof(LRBlock) * na, hipMemcpyDeviceToHost, s));
    ARH_HIP(hipStreamSynchronize(s));
    int64_t uo = 0, vo = 0, to = 0;
    for (int a = 0; a < na; a++) {
        LRBlock &B = h->hblocks[a];
        B.u_off = uo, B.v_off = vo, B.t_off = to;
        uo += (int64_t)B.rank * (B.re - B.rs);
        vo += (int64_t)B.rank * (B.ce - B.cs);
        to += B.rank;
        if (B.rank > 0) h->active_lr.push_back(a);
    }
    h->u_total = uo, h->v_total = vo, h->t_total = to;
    upload(h->blocks, h->hblocks, s);
    upload(h->d_active_lr, h->active_lr, s);
    h->U.alloc(sizeof(double2) * std::max<int64_t>(1, uo));
    h->V.alloc(sizeof(double2) * std::max<int64_t>(1, vo));
    if (!h->active_lr.empty()) {
        hipLaunchKernelGGL(k_pack, dim3((unsigned)h->active_lr.size()), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                           h->d_active_lr.as<int>(), g->d_leaf_start.as<int>(), g->aca_U.as<double2>(),
                           g->aca_V.as<double2>(), h->U.as<double2>(), h->V.as<double2>());
        ARH_HIP(hipGetLastError());
    }
    ARH_HIP(hipStreamSynchronize(s));
    // the scratch goes back to the device memory cache (the next assembly takes it from there), so an
    // idle geometry holds no ACA scratch
    g->aca_U.reset();
    g->aca_V.reset();
    g->aca_flags.reset();
    g->scratch_rank = -1;
}

// FP32 copy for the generic (non-MFMA) product: dense panels and U elementwise, V column-major
// per block. Built on first use.
void make_generic_fp32(arh_hmat *h) {
    if (h->generic32) return;
    arh_geom *g = h->g;
    hipStream_t s = g->ctx->stream;
    auto conv = [&](DBuf &src, DBuf &dst, int64_t n) {
        dst.alloc(sizeof(float2) * std::max<int64_t>(1, n));
        if (n > 0) {
            int blocks = (int)std::min<int64_t>(4096, (n + 255) / 256);
            hipLaunchKernelGGL(k_to_c64, dim3(blocks), dim3(256), 0, s, src.as<double2>(), dst.as<float2>(), n);
            ARH_HIP(hipGetLastError());
        }
    };
    conv(h->dense, h->dense32, g->dense_entries);
    conv(h->U, h->U32, h->u_total);
    h->V32.alloc(sizeof(float2) * std::max<int64_t>(1, h->v_total));
    if (!h->active_lr.empty()) {
        hipLaunchKernelGGL(k_v_to_c64_cm, dim3((unsigned)h->active_lr.size()), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                           h->d_active_lr.as<int>(), h->V.as<double2>(), h->V32.as<float2>());
        ARH_HIP(hipGetLastError());
    }
    ARH_HIP(hipStreamSynchronize(s));
    h->generic32 = true;
}

// FP32 copy at assembly: the MFMA layout when the leaves fit it, else the generic one.
void make_fp32_copy(arh_hmat *h) {
    if (h->g->max_leaf <= mfma_max_leaf())
        mfma_prepare(h, 0, 0);
    else
        make_generic_fp32(h);
}

}  // namespace arh
