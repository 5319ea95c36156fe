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
// Work is cut into "groups" of up to 4 consecutive complex columns of one operand segment. Every
// wavefront keeps a ring of kDepth groups in flight (A entries + rhs rows loaded ahead), which
// hides the HBM/L2 latency behind the MFMAs. Group descriptors are per-leaf flat lists built once
// per assembly (uniform per wavefront: scalar loads).
//
// Operands: the rhs block is packed row-major in FP32 (xf[i * RP + j], RP = 8 NT); intermediates
// t = V x are written in the same layout; the FP32 copy of V is column-major per block. FP32
// accumulation; y = -x/2 + A x is formed in FP64.
//
// Lane layout of v_mfma_f32_16x16x4_f32 (64 lanes): A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i / 4), register i % 4.
#include <algorithm>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

constexpr int kDepth = 4;      // groups in flight per wavefront
constexpr int kWaves = 8;      // wavefronts of the leaf / big-block workgroups
constexpr int kBigCols = 256;  // low-rank blocks with at least this many columns get a whole workgroup

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

// one B operand pair from a packed complex rhs row: output column n = 16 nt + i16 -> rhs n / 2, part n % 2
__device__ __forceinline__ void b_pair(float2 xv, int s, float &br, float &bi) {
    br = s ? xv.y : xv.x;
    bi = s ? xv.x : -xv.y;
}

// A group in registers: A entries of up to NRT 16-row tiles, rhs rows for NT output tiles.
template <int NT, int NRT>
struct Grp {
    float2 a[NRT];
    float2 xv[NT];
};

template <int NT, int NRT>
__device__ __forceinline__ void grp_mma(f4 (&acc)[NRT][NT], const Grp<NT, NRT> &g, int s, int nrt) {
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
        float br, bi;
        b_pair(g.xv[nt], s, br, bi);
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) {
            if (rt < nrt) {
                acc[rt][nt] = mfma4(g.a[rt].x, br, acc[rt][nt]);
                acc[rt][nt] = mfma4(g.a[rt].y, bi, acc[rt][nt]);
            }
        }
    }
}

// Load column kk of a group: A column at acol (rows i < arows valid), rhs row at brow.
template <int NT, int NRT>
__device__ __forceinline__ void grp_load(Grp<NT, NRT> &g, bool ok, const float2 *acol, int arows, int nrt, int i16,
                                         const float2 *brow) {
    const float2 z = make_float2(0.f, 0.f);
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
        const int i = rt * 16 + i16;
        g.a[rt] = (ok && rt < nrt && i < arows) ? acol[i] : z;
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++) g.xv[nt] = ok ? brow[nt * 8] : z;
}

// Fixed-order tree reduction of the kWaves wave partials into wave 0: (w, w+4), (w, w+2), (0, 1).
template <int NT, int NRT>
__device__ __forceinline__ void reduce_waves(f4 (&acc)[NRT][NT], float (*red)[NRT][NT][4][64], int w, int lane) {
    for (int half = kWaves / 2; half >= 1; half >>= 1) {
        __syncthreads();
        if (w >= half && w < 2 * half)
#pragma unroll
            for (int rt = 0; rt < NRT; rt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) red[w - half][rt][nt][v][lane] = acc[rt][nt][v];
        __syncthreads();
        if (w < half)
#pragma unroll
            for (int rt = 0; rt < NRT; rt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[rt][nt][v] += red[w][rt][nt][v][lane];
    }
}

// ---------------------------------------------------------------- phase 1: t_b = V_b x[cs:ce]
// A unit is (low-rank block, 32 rank rows). The FP32 V copy is column-major per block (V[c r + p]).
// Small blocks: one wavefront per unit; big blocks (>= kBigCols columns): a workgroup of kWaves
// wavefronts splits the columns and reduces in LDS.
template <int NT>
__device__ __forceinline__ void lr_unit(f4 (&acc)[2][NT], const LRBlock &B, int base, const float2 *__restrict__ V,
                                        const float2 *__restrict__ xf, int q0, int qstride, int lane) {
    constexpr int RP = 8 * NT;
    const int r = B.rank, n = B.ce - B.cs;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    const int arows = min(32, r - base);
    const int nrt = arows > 16 ? 2 : 1;
    const float2 *A0 = V + B.v_off + base;
    const float2 *B0 = xf + (int64_t)B.cs * RP + (i16 >> 1);
    const int ngr = (n + 3) >> 2;
    auto ld = [&](int q, Grp<NT, 2> &g) {
        const int c = q * 4 + kk;
        const bool ok = c < n;
        const int cc = ok ? c : 0;
        grp_load<NT, 2>(g, ok, A0 + (int64_t)cc * r, arows, nrt, i16, B0 + (int64_t)cc * RP);
    };
    Grp<NT, 2> ring[kDepth];
#pragma unroll
    for (int j = 0; j < kDepth; j++)
        if (q0 + j * qstride < ngr) ld(q0 + j * qstride, ring[j]);
    for (int qb = q0; qb < ngr; qb += kDepth * qstride) {
#pragma unroll
        for (int j = 0; j < kDepth; j++) {
            const int q = qb + j * qstride;
            if (q < ngr) {
                grp_mma<NT, 2>(acc, ring[j], s, nrt);
                const int qn = q + kDepth * qstride;
                if (qn < ngr) ld(qn, ring[j]);
            }
        }
    }
}

template <int NT>
__device__ __forceinline__ void lr_store(const f4 (&acc)[2][NT], const LRBlock &B, int base, float *__restrict__ t32,
                                         int lane) {
    constexpr int RP = 8 * NT;
    const int i16 = lane & 15, kk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = base + rt * 16 + kk * 4 + v;
                if (row < B.rank) t32[(B.t_off + row) * (2 * RP) + nt * 16 + i16] = acc[rt][nt][v];
            }
}

// Units [0, nbig) are big-block units (the whole workgroup on one unit, LDS reduction); the rest are
// small-block units, one per wavefront.
template <int NT>
__global__ __launch_bounds__(512) void k_lr_mfma(const LRBlock *__restrict__ blocks, const int2 *__restrict__ units,
                                                 int nbig, int nunits, const float2 *__restrict__ V,
                                                 const float2 *__restrict__ xf, float *__restrict__ t32) {
    __shared__ float red[kWaves / 2][2][NT][4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    f4 acc[2][NT];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[rt][nt] = f4{0.f, 0.f, 0.f, 0.f};
    if ((int)blockIdx.x < nbig) {
        const int2 un = units[blockIdx.x];
        const LRBlock B = blocks[un.x];
        lr_unit<NT>(acc, B, un.y * 32, V, xf, w, kWaves, lane);
        reduce_waves<NT, 2>(acc, red, w, lane);
        if (w == 0) lr_store<NT>(acc, B, un.y * 32, t32, lane);
        return;
    }
    const int u = nbig + ((int)blockIdx.x - nbig) * kWaves + w;
    if (u >= nunits) return;
    const int2 un = units[u];
    const LRBlock B = blocks[un.x];
    lr_unit<NT>(acc, B, un.y * 32, V, xf, 0, 1, lane);
    lr_store<NT>(acc, B, un.y * 32, t32, lane);
}

// ---------------------------------------------------------------- phase 2: one workgroup (8 waves) per row leaf
// The leaf's columns are its dense blocks (panel columns, rhs rows of consecutive elements) and its
// segment of every low-rank block covering it (U columns, rhs rows = the block's t rows), cut into
// groups; wave w takes groups w, w + 8, ...; the partial tiles are summed in a fixed order in LDS.
template <int NT, int MAXRT>
__global__ __launch_bounds__(512) void k_rows_mfma(const int *__restrict__ leaf_start, const int *__restrict__ gptr,
                                                   const ColGroup *__restrict__ groups,
                                                   const float2 *__restrict__ panel, const float2 *__restrict__ U,
                                                   const float2 *__restrict__ t32, const float2 *__restrict__ xf,
                                                   const double *__restrict__ x, int N, int R,
                                                   double *__restrict__ y) {
    constexpr int RP = 8 * NT;
    __shared__ float red[kWaves / 2][MAXRT][NT][4][64];
    const int l = blockIdx.x;
    const int rs = leaf_start[l], m = leaf_start[l + 1] - rs;
    const int nrt = (m + 15) >> 4;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    const int g0 = gptr[l], ngr = gptr[l + 1] - g0;
    f4 acc[MAXRT][NT];
#pragma unroll
    for (int rt = 0; rt < MAXRT; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[rt][nt] = f4{0.f, 0.f, 0.f, 0.f};
    auto ld = [&](int q, Grp<NT, MAXRT> &g) {
        const ColGroup d = groups[g0 + q];
        const bool ok = kk < d.ncols;
        const int c = ok ? kk : 0;
        const float2 *acol = (d.kind ? U : panel) + d.a_off + (int64_t)c * m;
        const float2 *brow = (d.kind ? t32 : xf) + (int64_t)(d.b_row + c) * RP + (i16 >> 1);
        grp_load<NT, MAXRT>(g, ok, acol, m, nrt, i16, brow);
    };
    Grp<NT, MAXRT> ring[kDepth];
#pragma unroll
    for (int j = 0; j < kDepth; j++)
        if (w + j * kWaves < ngr) ld(w + j * kWaves, ring[j]);
    for (int qb = w; qb < ngr; qb += kDepth * kWaves) {
#pragma unroll
        for (int j = 0; j < kDepth; j++) {
            const int q = qb + j * kWaves;
            if (q < ngr) {
                grp_mma<NT, MAXRT>(acc, ring[j], s, nrt);
                const int qn = q + kDepth * kWaves;
                if (qn < ngr) ld(qn, ring[j]);
            }
        }
    }
    reduce_waves<NT, MAXRT>(acc, red, w, lane);
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

// ---------------------------------------------------------------- host
// Per-assembly work lists: phase-1 units (small / big blocks) and the per-leaf column groups.
static void build_mfma_lists(arh_hmat *h, hipStream_t s) {
    arh_geom *g = h->g;
    // phase 1: big-block units first, each list by decreasing column count (longest first)
    std::vector<int2> big, small;
    for (int a : h->active_lr) {
        const LRBlock &B = h->hblocks[a];
        for (int rt = 0; rt * 32 < B.rank; rt++) (B.ce - B.cs >= kBigCols ? big : small).push_back(int2{a, rt});
    }
    auto by_cols = [&](const int2 &p, const int2 &q) {
        const int np = h->hblocks[p.x].ce - h->hblocks[p.x].cs, nq = h->hblocks[q.x].ce - h->hblocks[q.x].cs;
        return np != nq ? np > nq : (p.x != q.x ? p.x < q.x : p.y < q.y);
    };
    std::sort(big.begin(), big.end(), by_cols);
    std::sort(small.begin(), small.end(), by_cols);
    h->p1big = (int)big.size();
    h->p1units = big;
    h->p1units.insert(h->p1units.end(), small.begin(), small.end());
    // phase 2: per row leaf, its dense blocks in panel order, then its low-rank segments
    std::vector<std::vector<int>> rows(g->nleaf);
    for (int b = 0; b < g->nd; b++) rows[g->dblk_leafr[b]].push_back(b);
    h->cgroups.clear();
    h->cg_ptr.assign(g->nleaf + 1, 0);
    for (int l = 0; l < g->nleaf; l++) {
        const int rs = g->leaf_start[l], m = g->leaf_start[l + 1] - rs;
        std::sort(rows[l].begin(), rows[l].end(), [&](int p, int q) { return g->dblk_col[p] < g->dblk_col[q]; });
        for (int b : rows[l]) {
            const int pc = g->dblk_col[b], cs = g->dense[4 * b + 2], n = g->dense[4 * b + 3] - cs;
            ARH_CHECK(g->dense[4 * b] == rs && g->dense[4 * b + 1] - rs == m, ARH_EINVAL, "dense block is not a leaf row");
            ARH_CHECK(pc >= 0 && pc + n <= g->pcol_ptr[l + 1] - g->pcol_ptr[l], ARH_EINVAL, "panel column out of range");
            for (int c0 = 0; c0 < n; c0 += 4)
                h->cgroups.push_back(
                    ColGroup{g->panel_off[l] + (int64_t)(pc + c0) * m, cs + c0, (short)std::min(4, n - c0), (short)0});
        }
        for (int e = g->lrR_ptr[l]; e < g->lrR_ptr[l + 1]; e++) {
            const LRBlock &B = h->hblocks[g->lrR_blk[e]];
            const int r = B.rank;
            ARH_CHECK(rs >= B.rs && rs + m <= B.re, ARH_EINVAL, "low-rank block does not cover its row leaf");
            ARH_CHECK(B.t_off + r <= h->t_total && B.t_off + r < INT32_MAX, ARH_EINVAL, "intermediate offset out of range");
            for (int p0 = 0; p0 < r; p0 += 4)
                h->cgroups.push_back(ColGroup{B.u_off + (int64_t)(rs - B.rs) * r + (int64_t)p0 * m, (int)(B.t_off + p0),
                                              (short)std::min(4, r - p0), (short)1});
        }
        h->cg_ptr[l + 1] = (int)h->cgroups.size();
    }
    upload(h->d_p1units, h->p1units, s);
    upload(h->d_cgroups, h->cgroups, s);
    upload(h->d_cg_ptr, h->cg_ptr, s);
    h->mfma_ready = true;
}

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
    const int nu = (int)h->p1units.size(), nb = h->p1big;
    if (nu)
        timed(PROF_LR32, 8.0 * h->v_total + 16.0 * R * (N + h->t_total), 8.0 * R * h->v_total, [&] {
            hipLaunchKernelGGL(k_lr_mfma<NT>, dim3(nb + (nu - nb + kWaves - 1) / kWaves), dim3(512), 0, s,
                               h->blocks.as<LRBlock>(), h->d_p1units.as<int2>(), nb, nu, h->V32.as<float2>(),
                               h->xf32.as<float2>(), h->t32.as<float>());
        });
    const int64_t ent = g->dense_entries + h->u_total;
    timed(PROF_ROWS32, 8.0 * ent + 16.0 * R * (2.0 * N + h->t_total), 8.0 * R * ent, [&] {
        if (maxrt <= 2)
            hipLaunchKernelGGL((k_rows_mfma<NT, 2>), dim3(g->nleaf), dim3(512), 0, s, g->d_leaf_start.as<int>(),
                               h->d_cg_ptr.as<int>(), h->d_cgroups.as<ColGroup>(), h->dense32.as<float2>(),
                               h->U32.as<float2>(), h->t32.as<float2>(), h->xf32.as<float2>(), (const double *)x, N, R,
                               (double *)y);
        else
            hipLaunchKernelGGL((k_rows_mfma<NT, 4>), dim3(g->nleaf), dim3(512), 0, s, g->d_leaf_start.as<int>(),
                               h->d_cg_ptr.as<int>(), h->d_cgroups.as<ColGroup>(), h->dense32.as<float2>(),
                               h->U32.as<float2>(), h->t32.as<float2>(), h->xf32.as<float2>(), (const double *)x, N, R,
                               (double *)y);
    });
    ARH_HIP(hipGetLastError());
}

int mfma_max_leaf() { return 64; }

// y = (M - I/2) x on the complex-FP32 copy, right-hand sides in chunks of up to 40.
void hmat_apply_mfma(arh_hmat *h, int R, const double2 *x, double2 *y, hipStream_t s) {
    arh_geom *g = h->g;
    ARH_CHECK(g->max_leaf <= mfma_max_leaf(), ARH_EINVAL, "MFMA product supports leaves of at most 64 elements");
    const int maxrt = (g->max_leaf + 15) / 16;
    if (!h->mfma_ready) build_mfma_lists(h, s);
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
