// Multi-rhs complex H-matrix product on gfx950 MFMA: v_mfma_f32_16x16x4_f32 on the FP32 copy of
// the operator (mixed precision) and v_mfma_f64_16x16x4_f64 on an FP64 copy (FP64 true residuals
// and FP64 multi-rhs solves), one templated engine.
//
// A complex product Y = A X is one real product with A read as interleaved (re, im) columns:
//   Y' = A' X',  A'[:, 2c] = Re A[:, c], A'[:, 2c+1] = Im A[:, c],
//   X'[2c, 2j] = Re X[c, j], X'[2c, 2j+1] = Im X[c, j], X'[2c+1, 2j] = -Im X[c, j], X'[2c+1, 2j+1] = Re X[c, j],
// so Y'[:, 2j] = Re Y[:, j] and Y'[:, 2j+1] = Im Y[:, j]. Each lane loads one complex A entry for
// 4 consecutive complex columns (the K = 4 of the instruction) and feeds the real part to one MFMA
// (with the Re rows of X') and the imaginary part to a second (with the Im rows). The output
// columns 0..2R-1 are the interleaved re/im parts of the R right-hand sides, padded to NT tiles of 16.
//
// The operator copy has its own layout, built once per assembly (per precision):
//  * V stacks: the V factors of all low-rank blocks sharing a column cluster, stacked row-wise and
//    stored column-major, so phase 1 (t = V x) runs on full 16-row tiles even when ranks are small
//    and reads each x row once per stack instead of once per block;
//  * fused row-leaf panels: per row leaf, its dense columns followed by the U columns of every
//    low-rank block covering it (zero-padded to a multiple of 4 columns), with one source row per
//    column into the packed [x; t] buffer, so phase 2 (y = D x + U t) is one uniform column sweep.
// The rhs block and the intermediates share one row-major buffer xt (row i: RP complex values).
//
// Work is cut into "groups" of 4 consecutive columns. Each wavefront loads the next stage of groups
// (A entries + xt rows) before multiplying the current one, which hides the memory latency
// behind the MFMAs. Accumulation in the operator's precision; y = -x/2 + A x is formed in FP64.
//
// Lane layouts (64 lanes, 16x16x4): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k; D[i][j] in
// lane j + 16 (i / 4), element i % 4 for f32, lane j + 16 (i % 4), element i / 4 for f64.
#include <algorithm>
#include <unordered_map>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

template <class T>
struct Mfma;
template <>
struct Mfma<float> {
    typedef float2 T2;
    typedef float V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 mma(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __host__ __device__ __forceinline__ T2 make(double x, double y) { return make_float2((float)x, (float)y); }
    // row (within the 16-row tile) of accumulator element v of lane group kk = lane / 16
    static __device__ __forceinline__ int row(int kk, int v) { return kk * 4 + v; }
};
template <>
struct Mfma<double> {
    typedef double2 T2;
    typedef double V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 mma(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __host__ __device__ __forceinline__ T2 make(double x, double y) { return make_double2(x, y); }
    static __device__ __forceinline__ int row(int kk, int v) { return v * 4 + kk; }
};

#ifndef ARH_MV_DEPTH32
#define ARH_MV_DEPTH32 2
#endif
#ifndef ARH_MV_DEPTH64
#define ARH_MV_DEPTH64 2
#endif
#ifndef ARH_MV_RDEPTH32
#define ARH_MV_RDEPTH32 3
#endif
#ifndef ARH_MV_RDEPTH64
#define ARH_MV_RDEPTH64 2
#endif
#ifndef ARH_MV_WAVES
#define ARH_MV_WAVES 8
#endif
#ifndef ARH_MV_RWAVES
#define ARH_MV_RWAVES 4
#endif
#ifndef ARH_MV_CHUNKED
#define ARH_MV_CHUNKED 1
#endif
#ifndef ARH_MV_XCD
#define ARH_MV_XCD 1
#endif
#ifndef ARH_MV_SLOTS
#define ARH_MV_SLOTS 2048
#endif
// Groups per pipeline stage of a wavefront, for the V-stack (lr) and the row-task kernels (FP64 groups hold
// twice the registers).
template <class T>
constexpr int lr_depth() {
    return sizeof(T) == 4 ? ARH_MV_DEPTH32 : ARH_MV_DEPTH64;
}
template <class T>
constexpr int row_depth() {
    return sizeof(T) == 4 ? ARH_MV_RDEPTH32 : ARH_MV_RDEPTH64;
}
// Wavefronts per workgroup (powers of 2). Row tasks take 4: a workgroup then holds one wavefront per SIMD,
// so the register budget, not the workgroup size, sets how many fit a compute unit.
constexpr int kLrWaves = ARH_MV_WAVES;    // V-stack units (big stacks: the whole workgroup)
constexpr int kRowWaves = ARH_MV_RWAVES;  // row tasks
constexpr bool kChunked = ARH_MV_CHUNKED;  // a workgroup's waves take contiguous group ranges (else strided)
constexpr bool kXcdOrder = ARH_MV_XCD;     // row tasks ordered for the per-die L2s
constexpr int kXcds = 8;                   // dies (XCDs) of an MI355X, each with its own L2
#ifndef ARH_MV_BIGCOLS
#define ARH_MV_BIGCOLS 256
#endif
constexpr int kBigCols = ARH_MV_BIGCOLS;  // stacks with at least this many columns get a whole workgroup per unit
constexpr int kPartRows = 64;     // rows of a partial tile for leaves of up to 64 elements
constexpr int kMaxLeaf = 128;     // the largest leaf the engine takes (8 row tiles; partial tiles of 128 rows)
constexpr int kTaskSlots = ARH_MV_SLOTS;  // row tasks aimed for: a few per compute unit
constexpr int kMaxRC = 40;        // right-hand sides per pass (RP <= 40)

// Right-hand side j of a pass is column cols[j0 + j] of x / y (j0 + j without a list).
__device__ __forceinline__ int64_t col_of(const int *cols, int j0, int j) { return cols ? cols[j0 + j] : j0 + j; }

// A pass is skipped when the gate (the solver's count of active right-hand sides) is 0.
__device__ __forceinline__ bool gated(const int *gate) { return gate && *gate == 0; }

// xt[i * RP + j] = x[col j][i] for j < R, 0 for R <= j < RP (rows i < N)
template <class T>
__global__ void k_pack_xf(const double2 *__restrict__ x, int N, int R, int RP, typename Mfma<T>::T2 *__restrict__ xf,
                          const int *__restrict__ cols, int j0, const int *__restrict__ gate) {
    if (gated(gate)) return;
    const int64_t total = (int64_t)N * RP;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / RP), j = (int)(e % RP);
        double2 d = make_double2(0.0, 0.0);
        if (j < R) d = x[col_of(cols, j0, j) * N + i];
        xf[e] = Mfma<T>::make(d.x, d.y);
    }
}

// The value of lane l ^ 1 (DPP quad permutation [1, 0, 3, 2]; every lane of the wavefront is active).
__device__ __forceinline__ float pair_swap(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_swap(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The B operands of output column n = 16 nt + i16 (rhs n / 2, part s = n % 2) from the one real
// value the lane holds, xb = component s of that rhs's xt entry (the lane pair n, n ^ 1 holds its
// real and imaginary parts): br = xb, bi = the partner's value, negated on the real-part lane.
template <class T>
__device__ __forceinline__ void b_pair(T xb, int s, T &br, T &bi) {
    const T nb = pair_swap(xb);
    br = xb;
    bi = s ? nb : -nb;
}

// A group in registers: A entries of NRT 16-row tiles, one real rhs value per output tile.
template <class T, int NT, int NRT>
struct Grp {
    typename Mfma<T>::T2 a[NRT];
    T xb[NT];
    T okf;  // 0 for a padding column (masked groups only)
};

template <class T, int NT, int NRT>
using Acc = typename Mfma<T>::V4[NRT][NT];

// acc += group: all real-part products first, then all imaginary-part products, so consecutive
// MFMAs never write the same accumulator.
template <class T, int NT, int NRT, bool MASK>
__device__ __forceinline__ void grp_mma(Acc<T, NT, NRT> &acc, const Grp<T, NT, NRT> &g, int s) {
    T br[NT], bi[NT], ar[NRT], ai[NRT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) b_pair(g.xb[nt], s, br[nt], bi[nt]);
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
        ar[rt] = MASK ? g.a[rt].x * g.okf : g.a[rt].x;
        ai[rt] = MASK ? g.a[rt].y * g.okf : g.a[rt].y;
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) acc[rt][nt] = Mfma<T>::mma(ar[rt], br[nt], acc[rt][nt]);
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) acc[rt][nt] = Mfma<T>::mma(ai[rt], bi[nt], acc[rt][nt]);
}

// Load one column of a group: A column at acol (rows past amax clamped: their outputs are dropped), xt row
// at brow (real values: the lane's output column is brow[16 nt]). No load is conditional, and nothing waits
// on a load before the group is consumed (a padding mask is applied then).
template <class T, int NT, int NRT>
__device__ __forceinline__ void grp_load(Grp<T, NT, NRT> &g, T okf, const typename Mfma<T>::T2 *acol, int amax,
                                         int i16, const T *brow) {
    g.okf = okf;
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) g.a[rt] = acol[min(rt * 16 + i16, amax)];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) g.xb[nt] = brow[nt * 16];
}

// Component kk (0..3) of v as bit selects (lane-dependent kk: no branches, no exec-mask changes).
__device__ __forceinline__ int pick4(const int4 &v, int kk) {
    const int m1 = -(kk & 1), m2 = -((kk >> 1) & 1);
    const int lo = (v.y & m1) | (v.x & ~m1), hi = (v.w & m1) | (v.z & ~m1);
    return (hi & m2) | (lo & ~m2);
}

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

template <class T, int NT, int NRT>
__device__ __forceinline__ void zero_acc(Acc<T, NT, NRT> &acc) {
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[rt][nt][v] = T(0);
}

// An empty use of a group's registers. Placed on the loop's exit path it makes the loaded values live
// there too, so the compiler cannot sink a stage's loads past the loop's branches to just before the
// products that read them (which would leave no load in flight while a wavefront multiplies).
template <class T, int NT, int NRT>
__device__ __forceinline__ void keep(const Grp<T, NT, NRT> &g) {
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) asm volatile("" ::"v"(g.a[rt].x), "v"(g.a[rt].y));
#pragma unroll
    for (int nt = 0; nt < NT; nt++) asm volatile("" ::"v"(g.xb[nt]));
}

// Groups q0, q0 + qstride, ... < qend, kDepth at a time, double-buffered as two register sets that
// swap roles every stage (stage A multiplies set a while set b loads, stage B the other way round), so
// no register copy has to wait for a load: the only wait is for the set about to be multiplied, one
// stage after its loads were issued. A group's uniform metadata (pre: scalar loads) is fetched two
// stages before its data. Every load is issued unconditionally (indices past the end re-read the last
// group, whose values are then not used), which keeps the count of loads in flight fixed.
template <class T, int NT, int NRT, bool MASK, int kDepth, class Meta, class Pre, class Load>
__device__ __forceinline__ void ring_run(Acc<T, NT, NRT> &acc, int q0, int qstride, int qend, int s, Pre &&pre,
                                         Load &&ld) {
    if (q0 >= qend) return;
    const int qlast = qend - 1;
    const int step = kDepth * qstride;
    Grp<T, NT, NRT> ga[kDepth], gb[kDepth];
    Meta ma[kDepth], mb[kDepth];
#pragma unroll
    for (int j = 0; j < kDepth; j++) ma[j] = pre(min(q0 + j * qstride, qlast));
#pragma unroll
    for (int j = 0; j < kDepth; j++) ld(min(q0 + j * qstride, qlast), ma[j], ga[j]);
#pragma unroll
    for (int j = 0; j < kDepth; j++) mb[j] = pre(min(q0 + step + j * qstride, qlast));
    for (int qb = q0; qb < qend; qb += 2 * step) {
        // stage A: groups qb.. from set a; set b loads groups qb + step..
#pragma unroll
        for (int j = 0; j < kDepth; j++) ld(min(qb + step + j * qstride, qlast), mb[j], gb[j]);
#pragma unroll
        for (int j = 0; j < kDepth; j++) ma[j] = pre(min(qb + 2 * step + j * qstride, qlast));
        __builtin_amdgcn_sched_barrier(0);  // the loads go out before the products that wait for the other set
#pragma unroll
        for (int j = 0; j < kDepth; j++)
            if (qb + j * qstride < qend)
                grp_mma<T, NT, NRT, MASK>(acc, ga[j], s);
            else
                keep(ga[j]);  // a skipped group's loads are waited for here, not at the next loop head
        __builtin_amdgcn_sched_barrier(0);
        if (qb + step >= qend) break;
        // stage B: groups qb + step.. from set b; set a loads groups qb + 2 step..
#pragma unroll
        for (int j = 0; j < kDepth; j++) ld(min(qb + 2 * step + j * qstride, qlast), ma[j], ga[j]);
#pragma unroll
        for (int j = 0; j < kDepth; j++) mb[j] = pre(min(qb + 3 * step + j * qstride, qlast));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kDepth; j++)
            if (qb + step + j * qstride < qend)
                grp_mma<T, NT, NRT, MASK>(acc, gb[j], s);
            else
                keep(gb[j]);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = 0; j < kDepth; j++) keep(ga[j]), keep(gb[j]);
}

// The groups of [0, n) that wave w of a W-wave workgroup takes: (first, stride, end).
template <int W>
__device__ __forceinline__ int3 wave_share(int n, int w) {
    if (kChunked) {
        const int chunk = (n + W - 1) / W;
        return make_int3(w * chunk, 1, min(n, (w + 1) * chunk));
    }
    return make_int3(w, W, n);
}

// LDS of a reduction / output staging: W / 2 slots of NT * 256 values (one row tile each), and room for
// PR x 16 NT values of an output tile.
template <class T, int NT, int W, int PR = kPartRows>
struct Lds {
    T v[std::max((W / 2) * NT * 256, PR * 16 * NT)];
};

// Fixed-order tree reduction of the W wave partials into wave 0, one row tile at a time:
// (w, w + W/2), ..., (0, 1).
template <int W, class T, int NT, int NRT>
__device__ __forceinline__ void reduce_waves(Acc<T, NT, NRT> &acc, T *red, int w, int lane) {
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
#pragma unroll
        for (int half = W / 2; half >= 1; half >>= 1) {
            __syncthreads();
            if (w >= half && w < 2 * half) {
                T *r = red + (w - half) * NT * 256;
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) r[nt * 256 + v * 64 + lane] = acc[rt][nt][v];
            }
            __syncthreads();
            if (w < half) {
                const T *r = red + w * NT * 256;
#pragma unroll
                for (int nt = 0; nt < NT; nt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[rt][nt][v] += r[nt * 256 + v * 64 + lane];
            }
        }
    }
}

// Write a wave's result tile as rows of 2 RP values: tile[row * 2 RP + n] (row < rows).
template <class T, int NT, int NRT>
__device__ __forceinline__ void store_tile(const Acc<T, NT, NRT> &acc, T *tile, int rows, int lane) {
    constexpr int W = 16 * NT;
    const int i16 = lane & 15, kk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rt * 16 + Mfma<T>::row(kk, v);
                if (row < rows) tile[row * W + nt * 16 + i16] = acc[rt][nt][v];
            }
}

// ---------------------------------------------------------------- phase 1: t_s = V_s x[cs:ce]
// A unit is (stack, 32 stacked rows). Units [0, nbig) belong to stacks with at least kBigCols
// columns: the whole workgroup takes the unit and reduces in LDS. The rest are taken one per
// wavefront. The results go to xt rows t_row + p.
template <class T, int NT, int NRT>
__device__ __forceinline__ void lr_run(const VStack &S, int base, const typename Mfma<T>::T2 *__restrict__ V,
                                       typename Mfma<T>::T2 *__restrict__ xt, T *red, bool big, int w, int lane) {
    typedef typename Mfma<T>::T2 T2;
    constexpr int RP = 8 * NT;
    const int r = S.rtot, n = S.n;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    const int amax = min(32, r - base) - 1;
    const T2 *A0 = V + S.v_off + base;
    const T *B0 = reinterpret_cast<const T *>(xt) + (int64_t)S.cs * 2 * RP + i16;
    Acc<T, NT, NRT> acc;
    zero_acc<T, NT, NRT>(acc);
    const int ngr = (n + 3) >> 2;
    const int3 sh = big ? wave_share<kLrWaves>(ngr, w) : make_int3(0, 1, ngr);
    ring_run<T, NT, NRT, true, lr_depth<T>(), int>(
        acc, sh.x, sh.y, sh.z, s, [](int) { return 0; },
        [&](int q, int, Grp<T, NT, NRT> &g) {
            const int c = q * 4 + kk;
            const bool ok = c < n;
            const int cc = ok ? c : q * 4;
            grp_load<T, NT, NRT>(g, ok ? T(1) : T(0), A0 + (int64_t)cc * r, amax, i16, B0 + (int64_t)cc * 2 * RP);
        });
    if (big) {
        reduce_waves<kLrWaves, T, NT, NRT>(acc, red, w, lane);
        if (w != 0) return;
    }
    T *out = reinterpret_cast<T *>(xt + (S.t_row + base) * RP);
    store_tile<T, NT, NRT>(acc, out, r - base, lane);
}

template <class T, int NT>
__global__ __launch_bounds__(64 * kLrWaves) void k_lr_mfma(const VStack *__restrict__ stacks,
                                                         const int2 *__restrict__ units, int nbig, int nunits,
                                                         const typename Mfma<T>::T2 *__restrict__ V,
                                                         typename Mfma<T>::T2 *__restrict__ xt,
                                                         const int *__restrict__ gate) {
    __shared__ Lds<T, NT, kLrWaves> red;
    if (gated(gate)) return;
    const int w = wave_id(), lane = threadIdx.x & 63;
    const bool big = (int)blockIdx.x < nbig;
    const int u = big ? (int)blockIdx.x : nbig + ((int)blockIdx.x - nbig) * kLrWaves + w;
    if (u >= nunits) return;  // small units only: big workgroups always have a unit
    const int2 un = units[u];
    const VStack S = stacks[un.x];
    const int base = un.y * 32;
    if (S.rtot - base > 16)
        lr_run<T, NT, 2>(S, base, V, xt, red.v, big, w, lane);
    else
        lr_run<T, NT, 1>(S, base, V, xt, red.v, big, w, lane);
}

// ---------------------------------------------------------------- phase 2: row-leaf tasks
// A task is a range of one row leaf's panel groups, taken by one workgroup: each wave takes a share
// of the groups (wave_share) and the partial tiles are summed in a fixed order in LDS. A leaf with
// a single task writes y = -x/2 + acc itself; the tasks of a split leaf write partial tiles that
// k_rows_sum adds in task order. Output goes through LDS so that y rows are written contiguously.
template <class T, int NT, int NRT>
__device__ __forceinline__ void rows_run(const RowTask &T_, int rs, int m, const LeafPanel &L,
                                         const int *__restrict__ psrc,
                                         const typename Mfma<T>::T2 *__restrict__ panel,
                                         const typename Mfma<T>::T2 *__restrict__ xt, const double *__restrict__ x,
                                         int N, int R, double *__restrict__ y, const int *__restrict__ cols, int j0,
                                         T *__restrict__ part, int prow, T *red, int w, int lane) {
    typedef typename Mfma<T>::T2 T2;
    constexpr int RP = 8 * NT;
    const int i16 = lane & 15, kk = lane >> 4, s = i16 & 1;
    Acc<T, NT, NRT> acc;
    zero_acc<T, NT, NRT>(acc);
    const T2 *A0 = panel + L.off + (int64_t)kk * m;
    const int4 *src4 = reinterpret_cast<const int4 *>(psrc + L.col0) + T_.g0;
    const int3 sh = wave_share<kRowWaves>(T_.g1 - T_.g0, w);
    ring_run<T, NT, NRT, false, row_depth<T>(), int4>(
        acc, sh.x, sh.y, sh.z, s, [&](int q) { return src4[q]; },  // the group's 4 source rows (scalar load)
        [&](int q, const int4 &sr, Grp<T, NT, NRT> &g) {
            const int row = pick4(sr, kk);
            grp_load<T, NT, NRT>(g, T(1), A0 + (int64_t)(T_.g0 + q) * 4 * m, m - 1, i16,
                                 reinterpret_cast<const T *>(xt) + (int64_t)row * 2 * RP + i16);
        });
    reduce_waves<kRowWaves, T, NT, NRT>(acc, red, w, lane);
    __syncthreads();  // every wave has read the reduction buffer before it is reused for staging
    if (w == 0) store_tile<T, NT, NRT>(acc, red, m, lane);
    __syncthreads();
    if (T_.slot >= 0) {
        T *dst = part + (int64_t)T_.slot * prow * (2 * RP);
        for (int e = threadIdx.x; e < m * 2 * RP; e += blockDim.x) dst[e] = red[e];
    } else {
        for (int e = threadIdx.x; e < 2 * m * R; e += blockDim.x) {
            const int j = e / (2 * m), q = e - j * 2 * m, row = q >> 1, p = q & 1;
            const int64_t idx = (col_of(cols, j0, j) * N + rs + row) * 2 + p;
            y[idx] = -0.5 * x[idx] + (double)red[row * 2 * RP + 2 * j + p];
        }
    }
}

template <class T, int NT, int MAXRT>
__global__ __launch_bounds__(64 * kRowWaves) void k_rows_mfma(
    const RowTask *__restrict__ tasks, const int *__restrict__ leaf_start, const LeafPanel *__restrict__ lpanel,
    const int *__restrict__ psrc, const typename Mfma<T>::T2 *__restrict__ panel,
    const typename Mfma<T>::T2 *__restrict__ xt, const double *__restrict__ x, int N, int R, double *__restrict__ y,
    const int *__restrict__ cols, int j0, const int *__restrict__ gate, T *__restrict__ part, int prow) {
    constexpr int PR = MAXRT > 4 ? kMaxLeaf : kPartRows;  // output staging rows
    __shared__ Lds<T, NT, kRowWaves, PR> red;
    if (gated(gate)) return;
    const int w = wave_id(), lane = threadIdx.x & 63;
    const RowTask T_ = tasks[blockIdx.x];
    if (T_.leaf < 0) return;  // padding task
    const int rs = leaf_start[T_.leaf], m = leaf_start[T_.leaf + 1] - rs;
    const LeafPanel L = lpanel[T_.leaf];
    const int nrt = (m + 15) >> 4;
    if (nrt <= 1)
        rows_run<T, NT, 1>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
    else if (nrt == 2 || MAXRT == 2)
        rows_run<T, NT, 2>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
    else if constexpr (MAXRT == 4) {
        if (nrt == 3)
            rows_run<T, NT, 3>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
        else
            rows_run<T, NT, 4>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
    } else if constexpr (MAXRT == 8) {
        if (nrt <= 4)
            rows_run<T, NT, 4>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
        else if (nrt <= 6)
            rows_run<T, NT, 6>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
        else
            rows_run<T, NT, 8>(T_, rs, m, L, psrc, panel, xt, x, N, R, y, cols, j0, part, prow, red.v, w, lane);
    }
}

// y = -x/2 + the task partials of each split leaf, summed in task order. Workgroup (leaf, chunk) takes
// kSumRows rows of one split leaf: every thread holds its few elements in registers and issues the loads
// of up to kSumBatch partials together (indices clamped, surplus values not added), so a workgroup waits
// about once per batch rather than once per partial; the summed rows go through LDS so that y rows are
// written contiguously.
constexpr int kSumRows = 8;
constexpr int kSumBatch = 8;
template <class T, int NT>
__global__ __launch_bounds__(256) void k_rows_sum(const int4 *__restrict__ split, const int *__restrict__ leaf_start,
                                                  const T *__restrict__ part, const double *__restrict__ x, int N,
                                                  int R, double *__restrict__ y, const int *__restrict__ cols, int j0,
                                                  const int *__restrict__ gate, int prow) {
    constexpr int W = 16 * NT, E = kSumRows * W, EPT = (E + 255) / 256;
    __shared__ T tile[E];
    if (gated(gate)) return;
    const int4 S = split[blockIdx.x];
    const int rs = leaf_start[S.x], m = leaf_start[S.x + 1] - rs;
    const int r0 = (int)blockIdx.y * kSumRows;
    if (r0 >= m) return;
    const int mr = min(kSumRows, m - r0), ne = mr * W;
    const T *src = part + (int64_t)S.y * prow * W + (int64_t)r0 * W;
    T acc[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++) acc[i] = T(0);
    for (int k0 = 0; k0 < S.z; k0 += kSumBatch) {
        T v[kSumBatch][EPT];
#pragma unroll
        for (int b = 0; b < kSumBatch; b++) {
            const T *sk = src + (int64_t)min(k0 + b, S.z - 1) * prow * W;
#pragma unroll
            for (int i = 0; i < EPT; i++) v[b][i] = sk[min((int)threadIdx.x + i * 256, ne - 1)];
        }
#pragma unroll
        for (int b = 0; b < kSumBatch; b++)
#pragma unroll
            for (int i = 0; i < EPT; i++) acc[i] = k0 + b < S.z ? acc[i] + v[b][i] : acc[i];
    }
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int e = (int)threadIdx.x + i * 256;
        if (e < ne) tile[e] = acc[i];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * mr * R; e += blockDim.x) {
        const int j = e / (2 * mr), q = e - j * 2 * mr, row = q >> 1, p = q & 1;
        const int64_t idx = (col_of(cols, j0, j) * N + rs + r0 + row) * 2 + p;
        y[idx] = -0.5 * x[idx] + (double)tile[row * W + 2 * j + p];
    }
}

// ---------------------------------------------------------------- layout (once per assembly)
// Stacked factors, column-major per stack: entry (c, p) of stack row o_b + p <- V_b[p][c] (forward,
// FP64 V row-major) or U_b[c][p] (transposed, FP64 U leaf-segmented: the row's leaf chunk).
template <class T, bool UT>
__global__ __launch_bounds__(256) void k_stack(const LRBlock *__restrict__ blocks, const StackMap *__restrict__ maps,
                                               const double2 *__restrict__ F, const int *__restrict__ leaf_start,
                                               const int *__restrict__ leaf_of,
                                               typename Mfma<T>::T2 *__restrict__ Vs) {
    const StackMap M = maps[blockIdx.x];
    const LRBlock B = blocks[M.blk];
    const int r = B.rank, n = UT ? B.re - B.rs : B.ce - B.cs;
    for (int64_t e = threadIdx.x; e < (int64_t)r * n; e += blockDim.x) {
        const int c = (int)(e / r), p = (int)(e - (int64_t)c * r);
        double2 v;
        if (UT) {
            const int l = leaf_of[B.rs + c];
            const int ls = leaf_start[l] - B.rs, ms = leaf_start[l + 1] - leaf_start[l];
            v = F[B.u_off + (int64_t)ls * r + (int64_t)p * ms + (c - ls)];
        } else {
            v = F[B.v_off + (int64_t)p * n + c];
        }
        Vs[M.dst + (int64_t)c * M.rtot + p] = Mfma<T>::make(v.x, v.y);
    }
}

// Fused panels: entry i of panel column c of leaf l <- the FP64 value at csrc[col0 + c].
template <class T>
__global__ __launch_bounds__(256) void k_fuse_panel(const int *__restrict__ leaf_start,
                                                    const LeafPanel *__restrict__ lpanel,
                                                    const ColSrc *__restrict__ csrc, const double2 *__restrict__ dense,
                                                    const double2 *__restrict__ U, const double2 *__restrict__ V,
                                                    typename Mfma<T>::T2 *__restrict__ panel) {
    // workgroup (leaf, y): its wavefronts take the leaf's panel columns y * 4 + w, + 4 gridDim.y, ...; the
    // lanes of a wavefront take the rows of a column
    const int l = blockIdx.x;
    const int m = leaf_start[l + 1] - leaf_start[l];
    const LeafPanel L = lpanel[l];
    const int ncol = L.ngroups * 4;
    const int w = threadIdx.x / ARH_WAVE, lane = threadIdx.x % ARH_WAVE;
    for (int c = blockIdx.y * 4 + w; c < ncol; c += gridDim.y * 4) {
        const ColSrc cs = csrc[L.col0 + c];
        const double2 *src = cs.kind == 1 ? dense : cs.kind == 2 ? U : V;
        typename Mfma<T>::T2 *dst = panel + L.off + (int64_t)c * m;
        for (int i = lane; i < m; i += ARH_WAVE) {
            double2 v = make_double2(0.0, 0.0);
            if (cs.kind) v = src[cs.off + (int64_t)i * cs.stride];
            dst[i] = Mfma<T>::make(v.x, v.y);
        }
    }
}

// Panel column lists of one leaf, written by one wavefront: the source row (psrc: x row, or N + the
// intermediates row) and the FP64 gather source (csrc) of every column of the leaf's fused panel.
//  forward:    the row leaf's dense columns (panel column c of the dense panel), then for each low-rank
//              block covering the leaf (lr list order) its rank U columns restricted to the leaf;
//  transposed: for each dense block in the column leaf's list, its mr rows as columns (a row of the
//              column-major block in its row leaf's panel, stride mr), then the V columns restricted to
//              the leaf of each low-rank block whose column range covers it;
// zero columns (source row 0, kind 0) up to the leaf's ncol. A list's column offsets are a wavefront
// prefix sum of the entries' widths. bad[0] is raised when a column count disagrees with the host's
// ncol, a source row falls outside [0, nsrc), or a low-rank block does not cover the leaf.
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < ARH_WAVE; d <<= 1) {
        const int t = __shfl_up(v, d, ARH_WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

__global__ __launch_bounds__(ARH_WAVE) void k_panel_cols(
    int tr, const LeafPanel *__restrict__ lpanel, const int *__restrict__ leaf_start,
    const int *__restrict__ pcol_ptr, const int64_t *__restrict__ panel_off, const int *__restrict__ pcol_idx,
    const int *__restrict__ dT_ptr, const int *__restrict__ dT_blk, const int *__restrict__ dblk_info,
    const int *__restrict__ lr_ptr, const int *__restrict__ lr_blk,
    const LRBlock *__restrict__ blocks, const int *__restrict__ t_row, int nsrc, int *__restrict__ psrc,
    ColSrc *__restrict__ csrc, int *__restrict__ bad) {
    const int l = blockIdx.x, lane = threadIdx.x;
    const int ls = leaf_start[l], m = leaf_start[l + 1] - ls;
    const LeafPanel L = lpanel[l];
    const int ncol = L.ngroups * 4;
    int *ps = psrc + L.col0;
    ColSrc *cs = csrc + L.col0;
    int base = 0;
    bool ok = true;
    if (!tr) {
        const int p0 = pcol_ptr[l], nd = pcol_ptr[l + 1] - p0;
        for (int c = lane; c < nd && c < ncol; c += ARH_WAVE) {
            ps[c] = pcol_idx[p0 + c];
            cs[c] = ColSrc{panel_off[l] + (int64_t)c * m, 1, 1};
        }
        base = nd;
    } else {
        const int e0 = dT_ptr[l], e1 = dT_ptr[l + 1];
        for (int c0 = e0; c0 < e1; c0 += ARH_WAVE) {
            const int e = c0 + lane;
            int b = 0, rl = 0, pc = 0, mr = 0;
            if (e < e1) {
                b = dT_blk[e];
                rl = dblk_info[2 * b], pc = dblk_info[2 * b + 1];
                mr = leaf_start[rl + 1] - leaf_start[rl];
            }
            const int inc = wave_incl_scan(mr, lane);
            const int start = base + inc - mr;
            const int rs = e < e1 ? leaf_start[rl] : 0;
            for (int j = 0; j < mr && start + j < ncol; j++) {
                ps[start + j] = rs + j;
                cs[start + j] = ColSrc{panel_off[rl] + (int64_t)pc * mr + j, mr, 1};
            }
            base += __shfl(inc, ARH_WAVE - 1, ARH_WAVE);
        }
    }
    const int e0 = lr_ptr[l], e1 = lr_ptr[l + 1];
    for (int c0 = e0; c0 < e1; c0 += ARH_WAVE) {
        const int e = c0 + lane;
        int r = 0, b = 0;
        LRBlock B{};
        if (e < e1) {
            b = lr_blk[e];
            B = blocks[b];
            r = B.rank;
            if (tr ? (ls < B.cs || ls + m > B.ce) : (ls < B.rs || ls + m > B.re)) ok = false;
        }
        const int inc = wave_incl_scan(r, lane);
        const int start = base + inc - r;
        for (int p = 0; p < r && start + p < ncol; p++) {
            const int src = t_row[b] + p;
            ps[start + p] = src;
            if (src < 0 || src >= nsrc) ok = false;
            cs[start + p] = tr ? ColSrc{B.v_off + (int64_t)p * (B.ce - B.cs) + (ls - B.cs), 1, 3}
                               : ColSrc{B.u_off + (int64_t)(ls - B.rs) * B.rank + (int64_t)p * m, 1, 2};
        }
        base += __shfl(inc, ARH_WAVE - 1, ARH_WAVE);
    }
    // the host rounded the leaf's column count up to a multiple of 4
    if (base > ncol || base + 4 <= ncol) ok = false;
    for (int c = base + lane; c < ncol; c += ARH_WAVE) {
        ps[c] = 0;
        cs[c] = ColSrc{0, 0, 0};
    }
    if (!ok) atomicOr(bad, 1);
}

// Work lists of one orientation, shared by both precisions (host + device copies).
//  forward:    stacks of V by column cluster; panels of row leaves: dense columns, then the U
//              columns (leaf segment) of the low-rank blocks covering the leaf.
//  transposed: stacks of U^T by row cluster; panels of column leaves: the transposed dense blocks
//              of the leaf, then the V columns (restricted to the leaf) of the low-rank blocks
//              whose column range covers it.
static void build_lists(arh_hmat *h, int tr) {
    arh_geom *g = h->g;
    MfmaLayout &M = h->mf[tr];
    hipStream_t s = g->ctx->stream;
    const int N = g->N;
    ARH_CHECK(g->max_leaf <= kMaxLeaf, ARH_EINVAL, "MFMA product supports leaves of at most 128 elements");
    // stacks: active low-rank blocks grouped by their column (forward) / row (transposed) cluster,
    // stacks in order of first appearance, members in active-block order
    auto key = [&](const LRBlock &B) {
        return tr ? ((int64_t)B.rs << 32) | (uint32_t)B.re : ((int64_t)B.cs << 32) | (uint32_t)B.ce;
    };
    const int nact = (int)h->active_lr.size();
    std::unordered_map<int64_t, int> sid;
    sid.reserve(2 * (size_t)nact + 1);
    std::vector<int> bsid(nact), first;
    for (int q = 0; q < nact; q++) {
        auto it = sid.emplace(key(h->hblocks[h->active_lr[q]]), (int)first.size());
        if (it.second) first.push_back(q);
        bsid[q] = it.first->second;
    }
    const int ns = (int)first.size();
    std::vector<int> mptr(ns + 1, 0), mem(nact);
    for (int q = 0; q < nact; q++) mptr[bsid[q] + 1]++;
    for (int k = 0; k < ns; k++) mptr[k + 1] += mptr[k];
    {
        std::vector<int> fill(mptr.begin(), mptr.end() - 1);
        for (int q = 0; q < nact; q++) mem[fill[bsid[q]]++] = h->active_lr[q];
    }
    M.stacks.clear();
    M.stacks.reserve(ns);
    std::vector<StackMap> maps;
    maps.reserve(nact);
    std::vector<int> t_row(h->hblocks.size(), -1);
    int64_t voff = 0, trow = N;
    for (int k = 0; k < ns; k++) {
        const LRBlock &B0 = h->hblocks[mem[mptr[k]]];
        VStack S{tr ? B0.rs : B0.cs, tr ? B0.re - B0.rs : B0.ce - B0.cs, 0, 0, voff, trow};
        for (int q = mptr[k]; q < mptr[k + 1]; q++) S.rtot += h->hblocks[mem[q]].rank;
        ARH_CHECK(trow + S.rtot < INT32_MAX, ARH_EINVAL, "stacked intermediates exceed the 32-bit row range");
        int o = 0;
        for (int q = mptr[k]; q < mptr[k + 1]; q++) {
            const int a = mem[q];
            maps.push_back(StackMap{a, S.rtot, voff + o});
            t_row[a] = (int)(trow + o);
            o += h->hblocks[a].rank;
        }
        M.stacks.push_back(S);
        voff += (int64_t)S.n * S.rtot;
        trow += S.rtot;
    }
    M.vs_total = voff;
    ARH_CHECK(trow - N == h->t_total, ARH_EINVAL, "stacked intermediates do not match the ranks");
    // phase-1 units: big stacks first, each list by decreasing column count
    std::vector<int2> big, small;
    for (int k = 0; k < (int)M.stacks.size(); k++)
        for (int rt = 0; rt * 32 < M.stacks[k].rtot; rt++)
            (M.stacks[k].n >= kBigCols ? big : small).push_back(int2{k, rt});
    auto by_cols = [&](const int2 &p, const int2 &q) {
        const int np = M.stacks[p.x].n, nq = M.stacks[q.x].n;
        return np != nq ? np > nq : (p.x != q.x ? p.x < q.x : p.y < q.y);
    };
    std::sort(big.begin(), big.end(), by_cols);
    std::sort(small.begin(), small.end(), by_cols);
    M.p1big = (int)big.size();
    M.p1units = big;
    M.p1units.insert(M.p1units.end(), small.begin(), small.end());
    // fused panels: column counts here, the column lists on the device (k_panel_cols)
    M.lpanel.assign(g->nleaf, LeafPanel{0, 0, 0});
    int64_t poff = 0, ncols = 0;
    for (int l = 0; l < g->nleaf; l++) {
        const int ls = g->leaf_start[l], m = g->leaf_start[l + 1] - ls;
        int64_t nc = 0;
        if (!tr) {
            nc = g->pcol_ptr[l + 1] - g->pcol_ptr[l];
            for (int e = g->lrR_ptr[l]; e < g->lrR_ptr[l + 1]; e++) nc += h->hblocks[g->lrR_blk[e]].rank;
        } else {
            for (int e = g->dT_ptr[l]; e < g->dT_ptr[l + 1]; e++) {
                const int b = g->dT_blk[e], rl = g->dblk_leafr[b];
                ARH_CHECK(g->dense[4 * b + 2] == ls && g->dense[4 * b + 3] == ls + m, ARH_EINVAL,
                          "dense block is not a leaf column");
                nc += g->leaf_start[rl + 1] - g->leaf_start[rl];
            }
            for (int e = g->lrC_ptr[l]; e < g->lrC_ptr[l + 1]; e++) nc += h->hblocks[g->lrC_blk[e]].rank;
        }
        nc = (nc + 3) / 4 * 4;
        ARH_CHECK(ncols + nc < INT32_MAX, ARH_EINVAL, "panel columns exceed the 32-bit range");
        M.lpanel[l] = LeafPanel{poff, (int)ncols, (int)(nc / 4)};
        ncols += nc;
        poff += nc * m;
    }
    M.panel_total = poff;
    upload(M.d_lpanel, M.lpanel, s);
    upload(M.d_trow, t_row, s);
    M.d_psrc.alloc(sizeof(int) * std::max<int64_t>(1, ncols));
    M.d_csrc.alloc(sizeof(ColSrc) * std::max<int64_t>(1, ncols));
    DBuf bad;
    bad.alloc(sizeof(int));
    ARH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_panel_cols, dim3(g->nleaf), dim3(ARH_WAVE), 0, s, tr, M.d_lpanel.as<LeafPanel>(),
                       g->d_leaf_start.as<int>(), g->d_pcol_ptr.as<int>(), g->d_panel_off.as<int64_t>(),
                       g->d_pcol_idx.as<int>(), g->d_dT_ptr.as<int>(), g->d_dT_blk.as<int>(),
                       g->d_dblk_info.as<int>(), tr ? g->d_lrC_ptr.as<int>() : g->d_lrR_ptr.as<int>(),
                       tr ? g->d_lrC_blk.as<int>() : g->d_lrR_blk.as<int>(), h->blocks.as<LRBlock>(),
                       M.d_trow.as<int>(), (int)(N + h->t_total), M.d_psrc.as<int>(), M.d_csrc.as<ColSrc>(),
                       bad.as<int>());
    ARH_HIP(hipGetLastError());
    // row tasks: leaves with more than tg groups are split evenly
    int64_t gt = 0;
    for (auto &L : M.lpanel) gt += L.ngroups;
    int tg = (int)std::max<int64_t>(8 * kRowWaves, (gt + kTaskSlots - 1) / kTaskSlots);
    tg = (tg + kRowWaves - 1) / kRowWaves * kRowWaves;
    M.rtasks.clear();
    M.rsplit.clear();
    M.nslots = 0;
    for (int l = 0; l < g->nleaf; l++) {
        const int n = M.lpanel[l].ngroups;
        const int nt = std::max(1, (n + tg - 1) / tg);
        if (nt > 1) M.rsplit.push_back(int4{l, M.nslots, nt, 0});
        for (int t = 0; t < nt; t++)
            M.rtasks.push_back(RowTask{l, (int)((int64_t)n * t / nt), (int)((int64_t)n * (t + 1) / nt),
                                       nt > 1 ? M.nslots + t : -1});
        if (nt > 1) M.nslots += nt;
    }
    // XCD-aware order: workgroups are dispatched round-robin over the kXcds dies, each with its own
    // L2, so die x gets a contiguous run of leaves (in tree order, which neighbouring leaves share
    // most of their low-rank blocks and hence their intermediates rows) with about 1/kXcds of the
    // groups; padding tasks (leaf -1) even out the runs. Otherwise: largest tasks first.
    if (kXcdOrder) {
        std::vector<std::vector<RowTask>> run(kXcds);
        int64_t acc = 0;
        for (const RowTask &T : M.rtasks) {
            const int x = (int)std::min<int64_t>(kXcds - 1, acc * kXcds / std::max<int64_t>(1, gt));
            run[x].push_back(T);
            acc += T.g1 - T.g0;
        }
        size_t len = 0;
        for (auto &r : run) len = std::max(len, r.size());
        M.rtasks.assign(len * kXcds, RowTask{-1, 0, 0, -1});
        for (int x = 0; x < kXcds; x++)
            for (size_t i = 0; i < run[x].size(); i++) M.rtasks[i * kXcds + x] = run[x][i];
    } else {
        std::stable_sort(M.rtasks.begin(), M.rtasks.end(),
                         [](const RowTask &p, const RowTask &q) { return p.g1 - p.g0 > q.g1 - q.g0; });
    }
    upload(M.d_stacks, M.stacks, s);
    upload(M.d_p1units, M.p1units, s);
    upload(M.d_rtasks, M.rtasks, s);
    upload(M.d_rsplit, M.rsplit, s);
    upload(M.d_maps, maps, s);
    M.nmaps = (int)maps.size();
    int nbad = 0;
    ARH_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ARH_HIP(hipStreamSynchronize(s));  // the host vectors above are freed on return
    ARH_CHECK(!nbad, ARH_EINVAL, "panel column lists disagree with the partition or the ranks");
    M.lists = true;
}

// The operator copy of one orientation and precision (0: FP32, 1: FP64) in the MFMA layout.
template <class T>
static void build_copy(arh_hmat *h, int tr, int pi) {
    typedef typename Mfma<T>::T2 T2;
    arh_geom *g = h->g;
    MfmaLayout &M = h->mf[tr];
    hipStream_t s = g->ctx->stream;
    M.vs[pi].alloc(sizeof(T2) * std::max<int64_t>(1, M.vs_total));
    M.panel[pi].alloc(sizeof(T2) * std::max<int64_t>(1, M.panel_total));
    if (M.nmaps) {
        if (tr)
            hipLaunchKernelGGL((k_stack<T, true>), dim3((unsigned)M.nmaps), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                               M.d_maps.as<StackMap>(), h->U.as<double2>(), g->d_leaf_start.as<int>(),
                               g->d_leaf_of.as<int>(), M.vs[pi].as<T2>());
        else
            hipLaunchKernelGGL((k_stack<T, false>), dim3((unsigned)M.nmaps), dim3(256), 0, s, h->blocks.as<LRBlock>(),
                               M.d_maps.as<StackMap>(), h->V.as<double2>(), g->d_leaf_start.as<int>(),
                               g->d_leaf_of.as<int>(), M.vs[pi].as<T2>());
    }
    hipLaunchKernelGGL(k_fuse_panel<T>, dim3(g->nleaf, 32), dim3(256), 0, s, g->d_leaf_start.as<int>(),
                       M.d_lpanel.as<LeafPanel>(), M.d_csrc.as<ColSrc>(), h->dense.as<double2>(), h->U.as<double2>(),
                       h->V.as<double2>(), M.panel[pi].as<T2>());
    ARH_HIP(hipGetLastError());
    ARH_HIP(hipStreamSynchronize(s));
    M.built[pi] = true;
}

void mfma_prepare(arh_hmat *h, int tr, int fp64) {
    MfmaLayout &M = h->mf[tr];
    if (!M.lists) build_lists(h, tr);
    if (M.built[fp64]) return;
    if (fp64)
        build_copy<double>(h, tr, 1);
    else
        build_copy<float>(h, tr, 0);
}

// Widest pass (NT output tiles) that fits the registers with 4 row tiles (leaves over 32 elements) and with
// 8 row tiles (leaves over 64 elements).
template <class T>
constexpr int max_nt_tall() {
    return sizeof(T) == 8 ? 3 : 5;
}
template <class T>
constexpr int max_nt_xtall() {
    return sizeof(T) == 8 ? 1 : 2;
}

// Rows of a partial tile: the largest leaf's row tiles (64 rows up to 4 tiles, else 128).
inline int part_rows(int maxrt) { return maxrt > 4 ? kMaxLeaf : kPartRows; }

template <class T, int NT>
static void launch_nt(arh_hmat *h, MfmaLayout &M, int R, const double2 *x, double2 *y, const int *cols, int j0,
                      const int *gate, hipStream_t s, int maxrt) {
    typedef typename Mfma<T>::T2 T2;
    constexpr int pi = sizeof(T) == 8 ? 1 : 0;
    arh_geom *g = h->g;
    Profiler &pf = g->ctx->prof;
    const int N = g->N;
    constexpr int RP = 8 * NT;
    M.xt[pi].alloc(sizeof(T2) * (size_t)(N + h->t_total) * RP);
    const int prow = part_rows(maxrt);
    if (M.nslots) M.part[pi].alloc(sizeof(T) * (size_t)M.nslots * prow * 2 * RP);
    T2 *xt = M.xt[pi].as<T2>();
    hipLaunchKernelGGL(k_pack_xf<T>, dim3((unsigned)std::min<int64_t>(2048, ((int64_t)N * RP + 255) / 256)), dim3(256),
                       0, s, x, N, R, RP, xt, cols, j0, gate);
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
    // Algorithmic bytes: every operator entry once (eb = one complex entry of T) and every packed rhs
    // / intermediates value once per rhs (also eb: xt holds T2 values), plus the FP64 x read for the
    // -x/2 term and the FP64 y write (16 B per rhs per element) in phase 2. Partial tiles of split
    // leaves and re-reads are not algorithmic.
    const double eb = 2.0 * sizeof(T);
    const int nu = (int)M.p1units.size(), nb = M.p1big;
    // phase-1 entries: V (forward) or U (transposed); both total sum_b r_b (m_b or n_b)
    const int64_t e1 = M.vs_total;
    if (nu)
        timed(pi ? PROF_LR64 : PROF_LR32, eb * e1 + eb * R * (N + h->t_total), 8.0 * R * e1, [&] {
            hipLaunchKernelGGL((k_lr_mfma<T, NT>), dim3(nb + (nu - nb + kLrWaves - 1) / kLrWaves),
                               dim3(64 * kLrWaves), 0, s, M.d_stacks.as<VStack>(), M.d_p1units.as<int2>(), nb, nu,
                               M.vs[pi].as<T2>(), xt, gate);
        });
    const int64_t ent = g->dense_entries + (h->u_total + h->v_total - e1);
    T *part = M.nslots ? M.part[pi].as<T>() : nullptr;
    const int ntask = (int)M.rtasks.size(), nsp = (int)M.rsplit.size();
    const double *xd = (const double *)x;
    double *yd = (double *)y;
    timed(pi ? PROF_ROWS64 : PROF_ROWS32, eb * ent + eb * R * (N + h->t_total) + 32.0 * R * N, 8.0 * R * ent, [&] {
        auto launch_rows = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(ntask), dim3(64 * kRowWaves), 0, s, M.d_rtasks.as<RowTask>(),
                               g->d_leaf_start.as<int>(), M.d_lpanel.as<LeafPanel>(), M.d_psrc.as<int>(),
                               M.panel[pi].as<T2>(), (const T2 *)xt, xd, N, R, yd, cols, j0, gate, part, prow);
        };
        if (maxrt <= 2) {
            launch_rows(k_rows_mfma<T, NT, 2>);
        } else if (maxrt <= 4) {
            if constexpr (NT <= max_nt_tall<T>())
                launch_rows(k_rows_mfma<T, NT, 4>);
            else
                ARH_CHECK(false, ARH_EINVAL, "internal: pass too wide for leaves of more than 32 elements");
        } else {
            if constexpr (NT <= max_nt_xtall<T>())
                launch_rows(k_rows_mfma<T, NT, 8>);
            else
                ARH_CHECK(false, ARH_EINVAL, "internal: pass too wide for leaves of more than 64 elements");
        }
    });
    if (nsp)
        hipLaunchKernelGGL((k_rows_sum<T, NT>), dim3(nsp, (g->max_leaf + kSumRows - 1) / kSumRows), dim3(256), 0, s,
                           M.d_rsplit.as<int4>(), g->d_leaf_start.as<int>(), part, xd, N, R, yd, cols, j0, gate, prow);
    ARH_HIP(hipGetLastError());
}

int mfma_max_leaf() { return kMaxLeaf; }

// y = (M - I/2) x, or its transpose, on the MFMA layout of one precision, right-hand sides in passes
// of up to kMaxRC. With a column list (device array of ncols indices into the nrhs columns of x
// and y), only those columns are computed; the others of y are left untouched. A non-null gate
// skips every pass when *gate is 0.
template <class T>
static void apply_t(arh_hmat *h, MfmaLayout &M, int R, const double2 *x, double2 *y, hipStream_t s, const int *cols,
                    int ncols, const int *gate) {
    arh_geom *g = h->g;
    const int maxrt = (g->max_leaf + 15) / 16;
    const int nc = cols ? ncols : R;
    ARH_CHECK(nc >= 0 && nc <= R, ARH_EINVAL, "column list longer than the number of right-hand sides");
    const int maxrc = maxrt > 4 ? 8 * max_nt_xtall<T>() : maxrt > 2 ? 8 * max_nt_tall<T>() : kMaxRC;
    for (int j0 = 0; j0 < nc; j0 += maxrc) {
        const int Rc = std::min(maxrc, nc - j0);
        switch ((Rc + 7) / 8) {
        case 1: launch_nt<T, 1>(h, M, Rc, x, y, cols, j0, gate, s, maxrt); break;
        case 2: launch_nt<T, 2>(h, M, Rc, x, y, cols, j0, gate, s, maxrt); break;
        case 3: launch_nt<T, 3>(h, M, Rc, x, y, cols, j0, gate, s, maxrt); break;
        case 4: launch_nt<T, 4>(h, M, Rc, x, y, cols, j0, gate, s, maxrt); break;
        default: launch_nt<T, 5>(h, M, Rc, x, y, cols, j0, gate, s, maxrt); break;
        }
    }
}

void hmat_apply_mfma(arh_hmat *h, int transpose, int fp64, int R, const double2 *x, double2 *y, hipStream_t s,
                     const int *cols, int ncols, const int *gate) {
    ARH_CHECK(h->g->max_leaf <= mfma_max_leaf(), ARH_EINVAL, "MFMA product supports leaves of at most 128 elements");
    const int tr = transpose ? 1 : 0;
    mfma_prepare(h, tr, fp64 ? 1 : 0);
    if (fp64)
        apply_t<double>(h, h->mf[tr], R, x, y, s, cols, ncols, gate);
    else
        apply_t<float>(h, h->mf[tr], R, x, y, s, cols, ncols, gate);
}

}  // namespace arh
