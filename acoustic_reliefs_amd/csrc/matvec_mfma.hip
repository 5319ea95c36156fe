// Multi-rhs complex-FP32 H-matrix product on gfx950 MFMA (v_mfma_f32_16x16x4_f32).
//
// A complex product Y = A X is one real product with A read as interleaved (re, im) columns:
//   Y' = A' X',  A'[:, 2c] = Re A[:, c], A'[:, 2c+1] = Im A[:, c],
//   X'[2c, 2j] = Re X[c, j], X'[2c, 2j+1] = Im X[c, j], X'[2c+1, 2j] = -Im X[c, j], X'[2c+1, 2j+1] = Re X[c, j],
// so Y'[:, 2j] = Re Y[:, j] and Y'[:, 2j+1] = Im Y[:, j]. Each lane loads one complex A entry
// (a float2) for 4 consecutive complex columns (the K = 4 of the instruction) and feeds .x to one
// MFMA (with the Re rows of X') and .y to a second (with the Im rows). The output columns
// 0..2R-1 are the interleaved re/im parts of the R right-hand sides, padded to NT tiles of 16.
//
// Operands: the rhs block is first packed row-major in FP32 (xf[i * RP + j], RP = 8 NT), so the
// 4 columns of one MFMA step read contiguous rows; intermediates t = V x are written in the same
// row-major layout. Accumulation is FP32; y = -x/2 + A x is formed in FP64.
//
// Lane layout of v_mfma_f32_16x16x4_f32 (64 lanes): A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i / 4), register i % 4.
#include <algorithm>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// xf[i * RP + j] = (float) x[j * N + i] for j < R, 0 for R <= j < RP
__global__ void k_pack_xf(const double2 *__restrict__ x, int N, int R, int RP, float2 *__restrict__ xf) {
    const int64_t total = (int64_t)N * RP;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / RP), j = (int)(e % RP);
        float2 v = make_float2(0.f, 0.f);
        if (j < R) {
            double2 d = x[(int64_t)j * N + i];
            v = make_float2((float)d.x, (float)d.y);
        }
        xf[e] = v;
    }
}

// one B operand pair from a packed complex rhs row: lane's output column n = 16 nt + i16 -> rhs
// n / 2, part n % 2
__device__ __forceinline__ void b_pair(float2 xv, int s, float &br, float &bi) {
    br = s ? xv.y : xv.x;
    bi = s ? xv.x : -xv.y;
}

// ---------------------------------------------------------------- phase 1: t_b = V_b x[cs:ce]
// one wavefront per (low-rank block, 16-row tile of its rank); V row-major (r x n).
template <int NT>
__global__ __launch_bounds__(256) void k_lr_mfma(const LRBlock *__restrict__ blocks, const int2 *__restrict__ units,
                                                 int nunits, const float2 *__restrict__ V,
                                                 const float2 *__restrict__ xf, float *__restrict__ t32) {
    constexpr int RP = 8 * NT;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int u = blockIdx.x * 4 + w;
    if (u >= nunits) return;
    const int2 un = units[u];
    const LRBlock B = blocks[un.x];
    const int r = B.rank, n = B.ce - B.cs;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    const int p = un.y * 16 + i16;
    const bool prow = p < r;
    const float2 *vrow = V + B.v_off + (int64_t)(prow ? p : 0) * n;
    const float2 *xb = xf + (int64_t)B.cs * RP + (i16 >> 1);
    f4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < n; c0 += 4) {
        const int c = c0 + kk;
        const bool ok = c < n;
        const float2 a = (ok && prow) ? vrow[c] : make_float2(0.f, 0.f);
        const float2 *xr = xb + (int64_t)(ok ? c : 0) * RP;
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const float2 xv = ok ? xr[nt * 8] : make_float2(0.f, 0.f);
            float br, bi;
            b_pair(xv, s, br, bi);
            acc[nt] = mfma4(a.x, br, acc[nt]);
            acc[nt] = mfma4(a.y, bi, acc[nt]);
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = un.y * 16 + kk * 4 + v;
            if (row < r) t32[(B.t_off + row) * (2 * RP) + nt * 16 + i16] = acc[nt][v];
        }
}

// ---------------------------------------------------------------- phase 2: one workgroup (8 waves) per row leaf
// columns of the leaf: its dense panel, then the leaf's segment of every low-rank block covering
// it; 4-column groups are dealt round-robin to the waves; the 8 partial tiles are summed in a
// fixed order in LDS.
constexpr int kRowWaves = 8;

template <int NT, int MAXRT>
__global__ __launch_bounds__(512) void k_rows_mfma(const int *__restrict__ leaf_start, const int *__restrict__ pcol_ptr,
                                                   const int64_t *__restrict__ panel_off,
                                                   const int *__restrict__ pcol_idx, const float2 *__restrict__ panel,
                                                   const int *__restrict__ lr_ptr, const int *__restrict__ lr_blk,
                                                   const LRBlock *__restrict__ blocks, const float2 *__restrict__ U,
                                                   const float2 *__restrict__ t32, const float2 *__restrict__ xf,
                                                   const double *__restrict__ x, int N, int R,
                                                   double *__restrict__ y) {
    constexpr int RP = 8 * NT;
    __shared__ float red[kRowWaves / 2][MAXRT][NT][4][64];
    const int l = blockIdx.x;
    const int rs = leaf_start[l], m = leaf_start[l + 1] - rs;
    const int nrt = (m + 15) >> 4;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    f4 acc[MAXRT][NT];
#pragma unroll
    for (int rt = 0; rt < MAXRT; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[rt][nt] = f4{0.f, 0.f, 0.f, 0.f};

    auto step = [&](const float2 *acol, bool ok, const float2 *brow) {
        float br[NT], bi[NT];
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const float2 xv = ok ? brow[nt * 8] : make_float2(0.f, 0.f);
            b_pair(xv, s, br[nt], bi[nt]);
        }
#pragma unroll
        for (int rt = 0; rt < MAXRT; rt++) {
            if (rt < nrt) {
                const int i = rt * 16 + i16;
                const float2 a = (ok && i < m) ? acol[i] : make_float2(0.f, 0.f);
#pragma unroll
                for (int nt = 0; nt < NT; nt++) {
                    acc[rt][nt] = mfma4(a.x, br[nt], acc[rt][nt]);
                    acc[rt][nt] = mfma4(a.y, bi[nt], acc[rt][nt]);
                }
            }
        }
    };

    int g = 0;  // running 4-column group index over the leaf's columns
    {
        const int ncol = pcol_ptr[l + 1] - pcol_ptr[l];
        const float2 *P = panel + panel_off[l];
        const int *cidx = pcol_idx + pcol_ptr[l];
        const int ngr = (ncol + 3) >> 2;
        for (int q = (w - g % kRowWaves + kRowWaves) % kRowWaves; q < ngr; q += kRowWaves) {
            const int c = q * 4 + kk;
            const bool ok = c < ncol;
            const int col = ok ? cidx[c] : 0;
            step(P + (int64_t)(ok ? c : 0) * m, ok, xf + (int64_t)col * RP + (i16 >> 1));
        }
        g += ngr;
    }
    for (int e = lr_ptr[l]; e < lr_ptr[l + 1]; e++) {
        const LRBlock &B = blocks[lr_blk[e]];
        const int r = B.rank;
        const float2 *seg = U + B.u_off + (int64_t)(rs - B.rs) * r;
        const float2 *tb = t32 + B.t_off * RP + (i16 >> 1);
        const int ngr = (r + 3) >> 2;
        for (int q = (w - g % kRowWaves + kRowWaves) % kRowWaves; q < ngr; q += kRowWaves) {
            const int p = q * 4 + kk;
            const bool ok = p < r;
            step(seg + (int64_t)(ok ? p : 0) * m, ok, tb + (int64_t)(ok ? p : 0) * RP);
        }
        g += ngr;
    }

    // fixed-order tree reduction of the 8 wave partials: (w, w+4), then (w, w+2), then (0, 1)
    for (int half = kRowWaves / 2; half >= 1; half >>= 1) {
        __syncthreads();
        if (w >= half && w < 2 * half)
#pragma unroll
            for (int rt = 0; rt < MAXRT; rt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) red[w - half][rt][nt][v][lane] = acc[rt][nt][v];
        __syncthreads();
        if (w < half)
#pragma unroll
            for (int rt = 0; rt < MAXRT; rt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[rt][nt][v] += red[w][rt][nt][v][lane];
    }
    if (w == 0) {
#pragma unroll
        for (int rt = 0; rt < MAXRT; rt++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = rt * 16 + kk * 4 + v;
                    const int n = nt * 16 + i16;
                    const int j = n >> 1;
                    if (row < m && j < R) {
                        const int64_t idx = ((int64_t)j * N + rs + row) * 2 + (n & 1);
                        y[idx] = -0.5 * x[idx] + (double)acc[rt][nt][v];
                    }
                }
    }
}

// ---------------------------------------------------------------- host dispatch
template <int NT>
static void launch_nt(arh_hmat *h, int R, const double2 *x, double2 *y, hipStream_t s, int maxrt) {
    arh_geom *g = h->g;
    Profiler &pf = g->ctx->prof;
    const int N = g->N;
    constexpr int RP = 8 * NT;
    h->xf32.alloc(sizeof(float2) * (size_t)N * RP);
    h->t32.alloc(sizeof(float2) * std::max<int64_t>(1, h->t_total) * RP);
    hipLaunchKernelGGL(k_pack_xf, dim3((unsigned)std::min<int64_t>(2048, ((int64_t)N * RP + 255) / 256)), dim3(256), 0,
                       s, x, N, R, RP, h->xf32.as<float2>());
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
    const int nu = (int)h->p1units.size();
    if (nu)
        timed(PROF_LR32, 8.0 * h->v_total + 16.0 * R * (N + h->t_total), 8.0 * R * h->v_total, [&] {
            hipLaunchKernelGGL(k_lr_mfma<NT>, dim3((nu + 3) / 4), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                               h->d_p1units.as<int2>(), nu, h->V32.as<float2>(), h->xf32.as<float2>(),
                               h->t32.as<float>());
        });
    const int64_t ent = g->dense_entries + h->u_total;
    timed(PROF_ROWS32, 8.0 * ent + 16.0 * R * (2.0 * N + h->t_total), 8.0 * R * ent, [&] {
        if (maxrt <= 2)
            hipLaunchKernelGGL((k_rows_mfma<NT, 2>), dim3(g->nleaf), dim3(512), 0, s, g->d_leaf_start.as<int>(),
                               g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(), g->d_pcol_idx.as<int>(),
                               h->dense32.as<float2>(), g->d_lrR_ptr.as<int>(), g->d_lrR_blk.as<int>(),
                               h->blocks.as<LRBlock>(), h->U32.as<float2>(), h->t32.as<float2>(),
                               h->xf32.as<float2>(), (const double *)x, N, R, (double *)y);
        else
            hipLaunchKernelGGL((k_rows_mfma<NT, 4>), dim3(g->nleaf), dim3(512), 0, s, g->d_leaf_start.as<int>(),
                               g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(), g->d_pcol_idx.as<int>(),
                               h->dense32.as<float2>(), g->d_lrR_ptr.as<int>(), g->d_lrR_blk.as<int>(),
                               h->blocks.as<LRBlock>(), h->U32.as<float2>(), h->t32.as<float2>(),
                               h->xf32.as<float2>(), (const double *)x, N, R, (double *)y);
    });
    ARH_HIP(hipGetLastError());
}

int mfma_max_leaf() { return 64; }

// y = (M - I/2) x on the complex-FP32 copy, right-hand sides in chunks of up to 40.
void hmat_apply_mfma(arh_hmat *h, int R, const double2 *x, double2 *y, hipStream_t s) {
    arh_geom *g = h->g;
    int maxm = 0;
    for (int l = 0; l < g->nleaf; l++) maxm = std::max(maxm, g->leaf_start[l + 1] - g->leaf_start[l]);
    ARH_CHECK(maxm <= mfma_max_leaf(), ARH_EINVAL, "MFMA product supports leaves of at most 64 elements");
    const int maxrt = (maxm + 15) / 16;
    if (h->p1units.empty() && !h->active_lr.empty()) {
        for (int a : h->active_lr)
            for (int rt = 0; rt * 16 < h->hblocks[a].rank; rt++) h->p1units.push_back(int2{a, rt});
        upload(h->d_p1units, h->p1units, s);
    }
    const int64_t N = g->N;
    for (int j0 = 0; j0 < R; j0 += 40) {
        const int Rc = std::min(40, R - j0);
        const double2 *xc = x + j0 * N;
        double2 *yc = y + j0 * N;
        switch ((Rc + 7) / 8) {
        case 1: launch_nt<1>(h, Rc, xc, yc, s, maxrt); break;
        case 2: launch_nt<2>(h, Rc, xc, yc, s, maxrt); break;
        case 3: launch_nt<3>(h, Rc, xc, yc, s, maxrt); break;
        case 4: launch_nt<4>(h, Rc, xc, yc, s, maxrt); break;
        default: launch_nt<5>(h, Rc, xc, yc, s, maxrt); break;
        }
    }
}

}  // namespace arh
