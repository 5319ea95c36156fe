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
// The FP32 operator has its own layout, built once per assembly:
//  * V stacks: the V factors of all low-rank blocks sharing a column cluster, stacked row-wise and
//    stored column-major, so phase 1 (t = V x) runs on full 16-row tiles even when ranks are small
//    and reads each x row once per stack instead of once per block;
//  * fused row-leaf panels: per row leaf, its dense columns followed by the U columns of every
//    low-rank block covering it (zero-padded to a multiple of 4 columns), with one source row per
//    column into the packed [x; t] buffer, so phase 2 (y = D x + U t) is one uniform column sweep.
// The rhs block and the intermediates share one row-major FP32 buffer xt (row i: RP complex).
//
// Work is cut into "groups" of 4 consecutive columns. Each wavefront loads the next kDepth groups
// (A entries + xt rows) before multiplying the current kDepth, which hides the memory latency
// behind the MFMAs. FP32 accumulation; y = -x/2 + A x is formed in FP64.
//
// Lane layout of v_mfma_f32_16x16x4_f32 (64 lanes): A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i / 4), register i % 4.
#include <algorithm>
#include <map>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

#ifndef ARH_MV_DEPTH
#define ARH_MV_DEPTH 2
#endif
#ifndef ARH_MV_WAVES
#define ARH_MV_WAVES 8
#endif
#ifndef ARH_MV_CHUNKED
#define ARH_MV_CHUNKED 1
#endif
#ifndef ARH_MV_XCD
#define ARH_MV_XCD 1
#endif
constexpr int kDepth = ARH_MV_DEPTH;  // groups per pipeline stage of a wavefront
constexpr int kWaves = ARH_MV_WAVES;  // wavefronts of the row-task / big-stack workgroups (power of 2)
constexpr bool kChunked = ARH_MV_CHUNKED;  // a workgroup's waves take contiguous group ranges (else strided)
constexpr bool kXcdOrder = ARH_MV_XCD;     // row tasks ordered for the per-die L2s
constexpr int kXcds = 8;                   // dies (XCDs) of an MI355X, each with its own L2
constexpr int kBigCols = 256;     // stacks with at least this many columns get a whole workgroup per unit
constexpr int kPartRows = 64;     // rows of a partial tile (the largest leaf)
constexpr int kTaskSlots = 1024;  // row tasks aimed for: a few per compute unit
constexpr int kMaxRC = 40;        // right-hand sides per pass (RP <= 40)

// Right-hand side j of a pass is column cols[j0 + j] of x / y (j0 + j without a list).
__device__ __forceinline__ int64_t col_of(const int *cols, int j0, int j) { return cols ? cols[j0 + j] : j0 + j; }

// A pass is skipped when the gate (the solver's count of active right-hand sides) is 0.
__device__ __forceinline__ bool gated(const int *gate) { return gate && *gate == 0; }

// xt[i * RP + j] = (float) x[col j][i] for j < R, 0 for R <= j < RP (rows i < N)
__global__ void k_pack_xf(const double2 *__restrict__ x, int N, int R, int RP, float2 *__restrict__ xf,
                          const int *__restrict__ cols, int j0, const int *__restrict__ gate) {
    if (gated(gate)) return;
    const int64_t total = (int64_t)N * RP;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / RP), j = (int)(e % RP);
        float2 v = make_float2(0.f, 0.f);
        if (j < R) {
            double2 d = x[col_of(cols, j0, j) * N + i];
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

// A group in registers: A entries of NRT 16-row tiles, rhs rows for NT output tiles.
template <int NT, int NRT>
struct Grp {
    float2 a[NRT];
    float2 xv[NT];
    float okf;  // 0 for a padding column (masked groups only)
};

// acc += group: all real-part products first, then all imaginary-part products, so consecutive
// MFMAs never write the same accumulator.
template <int NT, int NRT, bool MASK>
__device__ __forceinline__ void grp_mma(f4 (&acc)[NRT][NT], const Grp<NT, NRT> &g, int s) {
    float br[NT], bi[NT], ar[NRT], ai[NRT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) b_pair(g.xv[nt], s, br[nt], bi[nt]);
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
        ar[rt] = MASK ? g.a[rt].x * g.okf : g.a[rt].x;
        ai[rt] = MASK ? g.a[rt].y * g.okf : g.a[rt].y;
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) acc[rt][nt] = mfma4(ar[rt], br[nt], acc[rt][nt]);
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) acc[rt][nt] = mfma4(ai[rt], bi[nt], acc[rt][nt]);
}

// Load one column of a group: A column at acol (rows past amax clamped: their outputs are
// dropped), xt row at brow. No load is conditional, and nothing waits on a load before the group
// is consumed (a padding mask is applied then).
template <int NT, int NRT>
__device__ __forceinline__ void grp_load(Grp<NT, NRT> &g, float okf, const float2 *acol, int amax, int i16,
                                         const float2 *brow) {
    g.okf = okf;
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) g.a[rt] = acol[min(rt * 16 + i16, amax)];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) g.xv[nt] = brow[nt * 8];
}

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

template <int NT, int NRT>
__device__ __forceinline__ void zero_acc(f4 (&acc)[NRT][NT]) {
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[rt][nt] = f4{0.f, 0.f, 0.f, 0.f};
}

// Groups q0, q0 + qstride, ... < qend, kDepth at a time, double-buffered: the next kDepth groups
// are loaded before the current ones are multiplied, so the only wait on loads comes after kDepth
// groups of products. A group's uniform metadata (pre: scalar loads) is fetched one stage before
// its data. Every load is issued unconditionally (indices past the end re-read the last group,
// whose values are then not used), which keeps the count of loads in flight fixed.
template <int NT, int NRT, bool MASK, class Meta, class Pre, class Load>
__device__ __forceinline__ void ring_run(f4 (&acc)[NRT][NT], int q0, int qstride, int qend, int s, Pre &&pre,
                                         Load &&ld) {
    if (q0 >= qend) return;
    const int qlast = qend - 1;
    Grp<NT, NRT> cur[kDepth], nxt[kDepth];
    Meta mc[kDepth], mn[kDepth];
#pragma unroll
    for (int j = 0; j < kDepth; j++) mc[j] = pre(min(q0 + j * qstride, qlast));
#pragma unroll
    for (int j = 0; j < kDepth; j++) ld(min(q0 + j * qstride, qlast), mc[j], cur[j]);
#pragma unroll
    for (int j = 0; j < kDepth; j++) mn[j] = pre(min(q0 + (kDepth + j) * qstride, qlast));
    for (int qb = q0; qb < qend; qb += kDepth * qstride) {
#pragma unroll
        for (int j = 0; j < kDepth; j++) ld(min(qb + (kDepth + j) * qstride, qlast), mn[j], nxt[j]);
#pragma unroll
        for (int j = 0; j < kDepth; j++) mn[j] = pre(min(qb + (2 * kDepth + j) * qstride, qlast));
#pragma unroll
        for (int j = 0; j < kDepth; j++)
            if (qb + j * qstride < qend) grp_mma<NT, NRT, MASK>(acc, cur[j], s);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kDepth; j++) cur[j] = nxt[j];
    }
}

// The groups of [0, n) that wave w of a kWaves workgroup takes: (first, stride, end).
__device__ __forceinline__ int3 wave_share(int n, int w) {
    if (kChunked) {
        const int chunk = (n + kWaves - 1) / kWaves;
        return make_int3(w * chunk, 1, min(n, (w + 1) * chunk));
    }
    return make_int3(w, kWaves, n);
}

// Fixed-order tree reduction of the kWaves wave partials into wave 0: (w, w+4), (w, w+2), (0, 1).
// red holds kWaves / 2 slots of NRT * NT * 256 floats.
template <int NT, int NRT>
__device__ __forceinline__ void reduce_waves(f4 (&acc)[NRT][NT], float *red, int w, int lane) {
    constexpr int S = NRT * NT * 256;
#pragma unroll
    for (int half = kWaves / 2; half >= 1; half >>= 1) {
        __syncthreads();
        if (w >= half && w < 2 * half) {
            float *r = red + (w - half) * S;
#pragma unroll
            for (int rt = 0; rt < NRT; rt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) r[(rt * NT + nt) * 256 + v * 64 + lane] = acc[rt][nt][v];
        }
        __syncthreads();
        if (w < half) {
            const float *r = red + w * S;
#pragma unroll
            for (int rt = 0; rt < NRT; rt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[rt][nt][v] += r[(rt * NT + nt) * 256 + v * 64 + lane];
        }
    }
}

// Write a wave's result tile as rows of 2 RP floats: tile[row * 2 RP + n] (row < rows).
template <int NT, int NRT>
__device__ __forceinline__ void store_tile(const f4 (&acc)[NRT][NT], float *tile, int rows, int lane) {
    constexpr int W = 16 * NT;
    const int i16 = lane & 15, kk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rt * 16 + kk * 4 + v;
                if (row < rows) tile[row * W + nt * 16 + i16] = acc[rt][nt][v];
            }
}

// ---------------------------------------------------------------- phase 1: t_s = V_s x[cs:ce]
// A unit is (stack, 32 stacked rows). Units [0, nbig) belong to stacks with at least kBigCols
// columns: the whole workgroup takes the unit and reduces in LDS. The rest are taken one per
// wavefront. The results go to xt rows t_row + p.
template <int NT, int NRT>
__device__ __forceinline__ void lr_run(const VStack &S, int base, const float2 *__restrict__ V,
                                       float2 *__restrict__ xt, float *red, bool big, int w, int lane) {
    constexpr int RP = 8 * NT;
    const int r = S.rtot, n = S.n;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    const int amax = min(32, r - base) - 1;
    const float2 *A0 = V + S.v_off + base;
    const float2 *B0 = xt + (int64_t)S.cs * RP + (i16 >> 1);
    f4 acc[NRT][NT];
    zero_acc<NT, NRT>(acc);
    const int ngr = (n + 3) >> 2;
    const int3 sh = big ? wave_share(ngr, w) : make_int3(0, 1, ngr);
    ring_run<NT, NRT, true, int>(
        acc, sh.x, sh.y, sh.z, s, [](int) { return 0; },
        [&](int q, int, Grp<NT, NRT> &g) {
            const int c = q * 4 + kk;
            const bool ok = c < n;
            const int cc = ok ? c : q * 4;
            grp_load<NT, NRT>(g, ok ? 1.f : 0.f, A0 + (int64_t)cc * r, amax, i16, B0 + (int64_t)cc * RP);
        });
    if (big) {
        reduce_waves<NT, NRT>(acc, red, w, lane);
        if (w != 0) return;
    }
    float *out = reinterpret_cast<float *>(xt + (S.t_row + base) * RP);
    store_tile<NT, NRT>(acc, out, r - base, lane);
}

template <int NT>
__global__ __launch_bounds__(64 * kWaves) void k_lr_mfma(const VStack *__restrict__ stacks, const int2 *__restrict__ units,
                                                 int nbig, int nunits, const float2 *__restrict__ V,
                                                 float2 *__restrict__ xt, const int *__restrict__ gate) {
    __shared__ float red[(kWaves / 2) * 2 * NT * 256];
    if (gated(gate)) return;
    const int w = wave_id(), lane = threadIdx.x & 63;
    const bool big = (int)blockIdx.x < nbig;
    const int u = big ? (int)blockIdx.x : nbig + ((int)blockIdx.x - nbig) * kWaves + w;
    if (u >= nunits) return;  // small units only: big workgroups always have a unit
    const int2 un = units[u];
    const VStack S = stacks[un.x];
    const int base = un.y * 32;
    if (S.rtot - base > 16)
        lr_run<NT, 2>(S, base, V, xt, red, big, w, lane);
    else
        lr_run<NT, 1>(S, base, V, xt, red, big, w, lane);
}

// ---------------------------------------------------------------- phase 2: row-leaf tasks
// A task is a range of one row leaf's panel groups, taken by one workgroup: each wave takes a share
// of the groups (wave_share) and the partial tiles are summed in a fixed order in LDS. A leaf with a single task
// writes y = -x/2 + acc itself; the tasks of a split leaf write partial tiles that k_rows_sum adds
// in task order. Output goes through LDS so that y rows are written contiguously.
template <int NT, int NRT>
__device__ __forceinline__ void rows_run(const RowTask &T, int rs, int m, const LeafPanel &L,
                                         const int *__restrict__ psrc, const float2 *__restrict__ panel,
                                         const float2 *__restrict__ xt, const double *__restrict__ x, int N, int R,
                                         double *__restrict__ y, const int *__restrict__ cols, int j0,
                                         float *__restrict__ part, float *red, int w, int lane) {
    constexpr int RP = 8 * NT;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    f4 acc[NRT][NT];
    zero_acc<NT, NRT>(acc);
    const float2 *A0 = panel + L.off + (int64_t)kk * m;
    const int4 *src4 = reinterpret_cast<const int4 *>(psrc + L.col0) + T.g0;
    const int3 sh = wave_share(T.g1 - T.g0, w);
    ring_run<NT, NRT, false, int4>(
        acc, sh.x, sh.y, sh.z, s, [&](int q) { return src4[q]; },  // the group's 4 source rows (scalar load)
        [&](int q, const int4 &sr, Grp<NT, NRT> &g) {
            const int row = kk == 0 ? sr.x : kk == 1 ? sr.y : kk == 2 ? sr.z : sr.w;
            grp_load<NT, NRT>(g, 1.f, A0 + (int64_t)(T.g0 + q) * 4 * m, m - 1, i16,
                              xt + (int64_t)row * RP + (i16 >> 1));
        });
    reduce_waves<NT, NRT>(acc, red, w, lane);
    // wave 0 holds the tile; stage it in LDS (slot 0 of red, which only wave 0 has read last)
    if (w == 0) store_tile<NT, NRT>(acc, red, m, lane);
    __syncthreads();
    if (T.slot >= 0) {
        float *dst = part + (int64_t)T.slot * kPartRows * (2 * RP);
        for (int e = threadIdx.x; e < m * 2 * RP; e += blockDim.x) dst[e] = red[e];
    } else {
        for (int e = threadIdx.x; e < 2 * m * R; e += blockDim.x) {
            const int j = e / (2 * m), q = e - j * 2 * m, row = q >> 1, p = q & 1;
            const int64_t idx = (col_of(cols, j0, j) * N + rs + row) * 2 + p;
            y[idx] = -0.5 * x[idx] + (double)red[row * 2 * RP + 2 * j + p];
        }
    }
}

template <int NT, int MAXRT>
__global__ __launch_bounds__(64 * kWaves) void k_rows_mfma(const RowTask *__restrict__ tasks, const int *__restrict__ leaf_start,
                                                   const LeafPanel *__restrict__ lpanel, const int *__restrict__ psrc,
                                                   const float2 *__restrict__ panel, const float2 *__restrict__ xt,
                                                   const double *__restrict__ x, int N, int R, double *__restrict__ y,
                                                   const int *__restrict__ cols, int j0, const int *__restrict__ gate,
                                                   float *__restrict__ part) {
    __shared__ float red[(kWaves / 2) * MAXRT * NT * 256];
    if (gated(gate)) return;
    const int w = wave_id(), lane = threadIdx.x & 63;
    const RowTask T = tasks[blockIdx.x];
    if (T.leaf < 0) return;  // padding task
    const int rs = leaf_start[T.leaf], m = leaf_start[T.leaf + 1] - rs;
    const LeafPanel L = lpanel[T.leaf];
    const int nrt = (m + 15) >> 4;
    if (nrt <= 1)
        rows_run<NT, 1>(T, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, red, w, lane);
    else if (nrt == 2 || MAXRT == 2)
        rows_run<NT, 2>(T, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, red, w, lane);
    else if constexpr (MAXRT >= 4) {
        if (nrt == 3)
            rows_run<NT, 3>(T, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, red, w, lane);
        else
            rows_run<NT, 4>(T, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, red, w, lane);
    }
}

// y = -x/2 + the task partials of each split leaf, summed in task order: partial rows are read
// contiguously into LDS, y rows are written contiguously from it.
template <int NT>
__global__ __launch_bounds__(256) void k_rows_sum(const int4 *__restrict__ split, const int *__restrict__ leaf_start,
                                                  const float *__restrict__ part, const double *__restrict__ x, int N,
                                                  int R, double *__restrict__ y, const int *__restrict__ cols, int j0,
                                                  const int *__restrict__ gate) {
    constexpr int W = 16 * NT;
    __shared__ float tile[kPartRows * W];
    if (gated(gate)) return;
    const int4 S = split[blockIdx.x];
    const int rs = leaf_start[S.x], m = leaf_start[S.x + 1] - rs;
    const float *src = part + (int64_t)S.y * kPartRows * W;
    for (int e = threadIdx.x; e < m * W; e += blockDim.x) {
        float acc = 0.f;
        for (int k = 0; k < S.z; k++) acc += src[(int64_t)k * kPartRows * W + e];
        tile[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * m * R; e += blockDim.x) {
        const int j = e / (2 * m), q = e - j * 2 * m, row = q >> 1, p = q & 1;
        const int64_t idx = (col_of(cols, j0, j) * N + rs + row) * 2 + p;
        y[idx] = -0.5 * x[idx] + (double)tile[row * W + 2 * j + p];
    }
}

// ---------------------------------------------------------------- layout (once per assembly)
// Stacked V: column c of stack entry (block b at stack row o) <- V_b[p][c] (FP64 row-major).
struct VMap {
    int blk, rtot;
    int64_t dst;  // Vs offset of (c = 0, p = 0) of the block
};

__global__ __launch_bounds__(256) void k_stack_v(const LRBlock *__restrict__ blocks, const VMap *__restrict__ maps,
                                                 const double2 *__restrict__ V, float2 *__restrict__ Vs) {
    const VMap M = maps[blockIdx.x];
    const LRBlock B = blocks[M.blk];
    const int n = B.ce - B.cs, r = B.rank;
    for (int64_t e = threadIdx.x; e < (int64_t)r * n; e += blockDim.x) {
        const int c = (int)(e / r), p = (int)(e - (int64_t)c * r);
        const double2 v = V[B.v_off + (int64_t)p * n + c];
        Vs[M.dst + (int64_t)c * M.rtot + p] = make_float2((float)v.x, (float)v.y);
    }
}

// Fused panels: panel column c of leaf l <- FP64 column at csrc[col0 + c] (>= 0: dense panel
// offset, <= -2: U offset -(v + 2), -1: zero column).
__global__ __launch_bounds__(256) void k_fuse_panel(const int *__restrict__ leaf_start,
                                                    const LeafPanel *__restrict__ lpanel,
                                                    const int64_t *__restrict__ csrc, const double2 *__restrict__ dense,
                                                    const double2 *__restrict__ U, float2 *__restrict__ panel) {
    const int l = blockIdx.x;
    const int m = leaf_start[l + 1] - leaf_start[l];
    const LeafPanel L = lpanel[l];
    const int64_t tot = (int64_t)L.ngroups * 4 * m;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const int c = (int)(e / m), i = (int)(e - (int64_t)c * m);
        const int64_t so = csrc[L.col0 + c];
        double2 v = make_double2(0.0, 0.0);
        if (so >= 0)
            v = dense[so + i];
        else if (so <= -2)
            v = U[-(so + 2) + i];
        panel[L.off + e] = make_float2((float)v.x, (float)v.y);
    }
}

void mfma_prepare(arh_hmat *h) {
    arh_geom *g = h->g;
    hipStream_t s = g->ctx->stream;
    const int N = g->N;
    ARH_CHECK(g->max_leaf <= kPartRows, ARH_EINVAL, "MFMA product supports leaves of at most 64 elements");
    // V stacks: active low-rank blocks grouped by column cluster, in order of first appearance
    std::map<std::pair<int, int>, int> sid;
    std::vector<std::vector<int>> members;
    for (int a : h->active_lr) {
        const LRBlock &B = h->hblocks[a];
        auto it = sid.emplace(std::make_pair(B.cs, B.ce), (int)members.size());
        if (it.second) members.emplace_back();
        members[it.first->second].push_back(a);
    }
    h->stacks.clear();
    std::vector<VMap> maps;
    std::vector<int64_t> t_row(h->hblocks.size(), -1);
    int64_t voff = 0, trow = N;
    for (auto &mem : members) {
        const LRBlock &B0 = h->hblocks[mem[0]];
        VStack S{B0.cs, B0.ce - B0.cs, 0, 0, voff, trow};
        for (int a : mem) S.rtot += h->hblocks[a].rank;
        int o = 0;
        for (int a : mem) {
            maps.push_back(VMap{a, S.rtot, voff + o});
            t_row[a] = trow + o;
            o += h->hblocks[a].rank;
        }
        h->stacks.push_back(S);
        voff += (int64_t)S.n * S.rtot;
        trow += S.rtot;
    }
    h->vs32_total = voff;
    ARH_CHECK(trow - N == h->t_total && trow < INT32_MAX, ARH_EINVAL, "stacked intermediates do not match the ranks");
    // phase-1 units: big stacks first, each list by decreasing column count
    std::vector<int2> big, small;
    for (int k = 0; k < (int)h->stacks.size(); k++)
        for (int rt = 0; rt * 32 < h->stacks[k].rtot; rt++) (h->stacks[k].n >= kBigCols ? big : small).push_back(int2{k, rt});
    auto by_cols = [&](const int2 &p, const int2 &q) {
        const int np = h->stacks[p.x].n, nq = h->stacks[q.x].n;
        return np != nq ? np > nq : (p.x != q.x ? p.x < q.x : p.y < q.y);
    };
    std::sort(big.begin(), big.end(), by_cols);
    std::sort(small.begin(), small.end(), by_cols);
    h->p1big = (int)big.size();
    h->p1units = big;
    h->p1units.insert(h->p1units.end(), small.begin(), small.end());
    // fused panels: per row leaf, the dense panel columns, then the U columns of its low-rank blocks
    h->lpanel.assign(g->nleaf, LeafPanel{0, 0, 0});
    std::vector<int> psrc;
    std::vector<int64_t> csrc;
    int64_t poff = 0;
    for (int l = 0; l < g->nleaf; l++) {
        const int rs = g->leaf_start[l], m = g->leaf_start[l + 1] - rs;
        const int col0 = (int)psrc.size();
        const int nd = g->pcol_ptr[l + 1] - g->pcol_ptr[l];
        for (int c = 0; c < nd; c++) {
            psrc.push_back(g->pcol_idx[g->pcol_ptr[l] + c]);
            csrc.push_back(g->panel_off[l] + (int64_t)c * m);
        }
        for (int e = g->lrR_ptr[l]; e < g->lrR_ptr[l + 1]; e++) {
            const int b = g->lrR_blk[e];
            const LRBlock &B = h->hblocks[b];
            ARH_CHECK(rs >= B.rs && rs + m <= B.re, ARH_EINVAL, "low-rank block does not cover its row leaf");
            for (int p = 0; p < B.rank; p++) {
                psrc.push_back((int)(t_row[b] + p));
                csrc.push_back(-(B.u_off + (int64_t)(rs - B.rs) * B.rank + (int64_t)p * m) - 2);
            }
        }
        while ((psrc.size() - col0) % 4) psrc.push_back(0), csrc.push_back(-1);
        const int ncol = (int)psrc.size() - col0;
        h->lpanel[l] = LeafPanel{poff, col0, ncol / 4};
        poff += (int64_t)ncol * m;
    }
    h->panel32_total = poff;
    for (int v : psrc) ARH_CHECK(v >= 0 && v < N + h->t_total, ARH_EINVAL, "panel source row out of range");
    // row tasks: leaves with more than tg groups are split evenly; largest tasks first
    int64_t gt = 0;
    for (auto &L : h->lpanel) gt += L.ngroups;
    int tg = (int)std::max<int64_t>(8 * kWaves, (gt + kTaskSlots - 1) / kTaskSlots);
    tg = (tg + kWaves - 1) / kWaves * kWaves;
    h->rtasks.clear();
    h->rsplit.clear();
    h->nslots = 0;
    for (int l = 0; l < g->nleaf; l++) {
        const int n = h->lpanel[l].ngroups;
        const int nt = std::max(1, (n + tg - 1) / tg);
        if (nt > 1) h->rsplit.push_back(int4{l, h->nslots, nt, 0});
        for (int t = 0; t < nt; t++)
            h->rtasks.push_back(RowTask{l, (int)((int64_t)n * t / nt), (int)((int64_t)n * (t + 1) / nt),
                                        nt > 1 ? h->nslots + t : -1});
        if (nt > 1) h->nslots += nt;
    }
    // XCD-aware order: workgroups are dispatched round-robin over the kXcds dies, each with its own
    // L2, so die x gets a contiguous run of leaves (in tree order, which neighbouring leaves share
    // most of their low-rank blocks and hence their intermediates rows) with about 1/kXcds of the
    // groups; padding tasks (leaf -1) even out the runs.
    if (kXcdOrder) {
        std::vector<std::vector<RowTask>> run(kXcds);
        int64_t acc = 0;
        for (const RowTask &T : h->rtasks) {
            const int x = (int)std::min<int64_t>(kXcds - 1, acc * kXcds / std::max<int64_t>(1, gt));
            run[x].push_back(T);
            acc += T.g1 - T.g0;
        }
        size_t len = 0;
        for (auto &r : run) len = std::max(len, r.size());
        h->rtasks.assign(len * kXcds, RowTask{-1, 0, 0, -1});
        for (int x = 0; x < kXcds; x++)
            for (size_t i = 0; i < run[x].size(); i++) h->rtasks[i * kXcds + x] = run[x][i];
    } else {
        std::stable_sort(h->rtasks.begin(), h->rtasks.end(),
                         [](const RowTask &p, const RowTask &q) { return p.g1 - p.g0 > q.g1 - q.g0; });
    }
    // device copies and the FP32 data
    upload(h->d_stacks, h->stacks, s);
    upload(h->d_p1units, h->p1units, s);
    upload(h->d_lpanel, h->lpanel, s);
    upload(h->d_psrc, psrc, s);
    upload(h->d_rtasks, h->rtasks, s);
    upload(h->d_rsplit, h->rsplit, s);
    DBuf d_maps, d_csrc;
    upload(d_maps, maps, s);
    upload(d_csrc, csrc, s);
    h->vs32.alloc(sizeof(float2) * std::max<int64_t>(1, h->vs32_total));
    h->panel32.alloc(sizeof(float2) * std::max<int64_t>(1, h->panel32_total));
    if (!maps.empty())
        hipLaunchKernelGGL(k_stack_v, dim3((unsigned)maps.size()), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                           d_maps.as<VMap>(), h->V.as<double2>(), h->vs32.as<float2>());
    hipLaunchKernelGGL(k_fuse_panel, dim3(g->nleaf), dim3(256), 0, s, g->d_leaf_start.as<int>(),
                       h->d_lpanel.as<LeafPanel>(), d_csrc.as<int64_t>(), h->dense.as<double2>(), h->U.as<double2>(),
                       h->panel32.as<float2>());
    ARH_HIP(hipGetLastError());
    ARH_HIP(hipStreamSynchronize(s));  // d_maps / d_csrc are freed on return
}

template <int NT>
static void launch_nt(arh_hmat *h, int R, const double2 *x, double2 *y, const int *cols, int j0, const int *gate,
                      hipStream_t s, int maxrt) {
    arh_geom *g = h->g;
    Profiler &pf = g->ctx->prof;
    const int N = g->N;
    constexpr int RP = 8 * NT;
    h->xt32.alloc(sizeof(float2) * (size_t)(N + h->t_total) * RP);
    if (h->nslots) h->part32.alloc(sizeof(float) * (size_t)h->nslots * kPartRows * 2 * RP);
    float2 *xt = h->xt32.as<float2>();
    hipLaunchKernelGGL(k_pack_xf, dim3((unsigned)std::min<int64_t>(2048, ((int64_t)N * RP + 255) / 256)), dim3(256), 0,
                       s, x, N, R, RP, xt, cols, j0, gate);
    auto timed = [&](int kind, double bytes, double flops, auto &&launch) {
        if (!pf.on) {
            launch();
            return;
        }
        Profiler::Entry e{pf.get(), pf.get(), kind, bytes, flops, gate ? pf.cur_slot : -1};
        ARH_HIP(hipEventRecord(e.a, s));
        launch();
        ARH_HIP(hipEventRecord(e.b, s));
        pf.pending.push_back(e);
    };
    const int nu = (int)h->p1units.size(), nb = h->p1big;
    if (nu)
        timed(PROF_LR32, 8.0 * h->v_total + 16.0 * R * (N + h->t_total), 8.0 * R * h->v_total, [&] {
            hipLaunchKernelGGL(k_lr_mfma<NT>, dim3(nb + (nu - nb + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, s,
                               h->d_stacks.as<VStack>(), h->d_p1units.as<int2>(), nb, nu, h->vs32.as<float2>(), xt,
                               gate);
        });
    const int64_t ent = g->dense_entries + h->u_total;
    float *part = h->nslots ? h->part32.as<float>() : nullptr;
    const int ntask = (int)h->rtasks.size(), nsp = (int)h->rsplit.size();
    const double *xd = (const double *)x;
    double *yd = (double *)y;
    timed(PROF_ROWS32, 8.0 * ent + 16.0 * R * (2.0 * N + h->t_total), 8.0 * R * ent, [&] {
        if (maxrt <= 2)
            hipLaunchKernelGGL((k_rows_mfma<NT, 2>), dim3(ntask), dim3(64 * kWaves), 0, s, h->d_rtasks.as<RowTask>(),
                               g->d_leaf_start.as<int>(), h->d_lpanel.as<LeafPanel>(), h->d_psrc.as<int>(),
                               h->panel32.as<float2>(), (const float2 *)xt, xd, N, R, yd, cols, j0, gate, part);
        else
            hipLaunchKernelGGL((k_rows_mfma<NT, 4>), dim3(ntask), dim3(64 * kWaves), 0, s, h->d_rtasks.as<RowTask>(),
                               g->d_leaf_start.as<int>(), h->d_lpanel.as<LeafPanel>(), h->d_psrc.as<int>(),
                               h->panel32.as<float2>(), (const float2 *)xt, xd, N, R, yd, cols, j0, gate, part);
        if (nsp)
            hipLaunchKernelGGL(k_rows_sum<NT>, dim3(nsp), dim3(256), 0, s, h->d_rsplit.as<int4>(),
                               g->d_leaf_start.as<int>(), part, xd, N, R, yd, cols, j0, gate);
    });
    ARH_HIP(hipGetLastError());
}

int mfma_max_leaf() { return kPartRows; }

// y = (M - I/2) x on the MFMA FP32 layout, right-hand sides in passes of up to kMaxRC. With a
// column list (device array of ncols indices into the nrhs columns of x and y), only those columns
// are computed; the others of y are left untouched. A non-null gate skips every pass when *gate is 0.
void hmat_apply_mfma(arh_hmat *h, int R, const double2 *x, double2 *y, hipStream_t s, const int *cols, int ncols,
                     const int *gate) {
    arh_geom *g = h->g;
    ARH_CHECK(g->max_leaf <= mfma_max_leaf(), ARH_EINVAL, "MFMA product supports leaves of at most 64 elements");
    ARH_CHECK(h->panel32.p, ARH_ESTATE, "the H-matrix has no MFMA FP32 layout");
    const int maxrt = (g->max_leaf + 15) / 16;
    const int nc = cols ? ncols : R;
    ARH_CHECK(nc >= 0 && nc <= R, ARH_EINVAL, "column list longer than the number of right-hand sides");
    for (int j0 = 0; j0 < nc; j0 += kMaxRC) {
        const int Rc = std::min(kMaxRC, nc - j0);
        switch ((Rc + 7) / 8) {
        case 1: launch_nt<1>(h, Rc, x, y, cols, j0, gate, s, maxrt); break;
        case 2: launch_nt<2>(h, Rc, x, y, cols, j0, gate, s, maxrt); break;
        case 3: launch_nt<3>(h, Rc, x, y, cols, j0, gate, s, maxrt); break;
        case 4: launch_nt<4>(h, Rc, x, y, cols, j0, gate, s, maxrt); break;
        default: launch_nt<5>(h, Rc, x, y, cols, j0, gate, s, maxrt); break;
        }
    }
}

}  // namespace arh
