// Internal host-side structures of libarh (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/arh.h"
#include "trig.h"

namespace arh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ARH_HIP(call)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            throw ::arh::Error(ARH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + \
                                             ":" + std::to_string(__LINE__));                                \
    } while (0)

#define ARH_CHECK(cond, code, msg)                              \
    do {                                                        \
        if (!(cond)) throw ::arh::Error((code), (msg));          \
    } while (0)

// A wavenumber the kernels accept: finite, >= 0 and at most kTrigMaxK, so that every phase k r they
// form stays inside the domain of their sine and cosine (trig.h).
#define ARH_CHECK_K(k)                                                                               \
    ARH_CHECK(std::isfinite(k) && (k) >= 0 && (k) <= ::arh::kTrigMaxK, ARH_EINVAL,                    \
              "wavenumber must be finite, >= 0 and <= 1e6")

// Device memory cache (pool.cpp). dev_alloc rounds n up to the block it hands out; the block is
// ordered after its previous owner's last use when the streams differ. tl_stream is the stream of
// the context the calling thread is working for (set on entry to every C-ABI call with a context).
extern thread_local hipStream_t tl_stream;
void *dev_alloc(size_t &n, hipStream_t s);
void *dev_alloc(size_t &n, hipStream_t s, int *device);
void dev_free(void *p, size_t n, hipStream_t s, int device);
size_t dev_trim(int device);    // frees every cached block of the device; returns the bytes
size_t dev_cached(int device);  // bytes held in the cache

// Owning device buffer, from the device memory cache; it belongs to the stream it was allocated
// for (the calling context's).
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t s = nullptr;
    int dev = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), bytes(o.bytes), s(o.s), dev(o.dev) { o.p = nullptr, o.bytes = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        reset();
        p = o.p, bytes = o.bytes, s = o.s, dev = o.dev;
        o.p = nullptr, o.bytes = 0;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (p) dev_free(p, bytes, s, dev);
        p = nullptr, bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        reset();
        if (n == 0) n = 16;
        s = tl_stream;
        p = dev_alloc(n, s, &dev);
        bytes = n;
    }
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(p);
    }
};

template <class T>
void upload(DBuf &b, const std::vector<T> &v, hipStream_t s) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) ARH_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace arh

namespace arh {
// Live kernel timing with HIP events on the context's stream.
enum ProfKind { PROF_ROWS64 = 0, PROF_LR64 = 1, PROF_ROWS32 = 2, PROF_LR32 = 3 };
struct Profiler {
    struct Entry {
        hipEvent_t a, b;
        int kind;
        double bytes, flops;
        int slot = -1;  // gated launch: index into slot_ok (see below), else -1
    };
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<Entry> pending;
    // Gated launches (skipped on the device when the solver's active count is 0): a launcher
    // tags its entries with cur_slot; before flushing, the solver sets slot_ok[slot] to whether
    // that slot's launches ran, and launches that did not are left out of the totals.
    int cur_slot = -1;
    std::vector<char> slot_ok;
    arh_profile acc{};
    hipEvent_t get() {
        if (pool.empty()) {
            hipEvent_t e;
            ARH_HIP(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    arh_kernel_prof &slot(int kind) {
        switch (kind) {
        case PROF_ROWS64: return acc.rows_fp64;
        case PROF_LR64: return acc.lowrank_fp64;
        case PROF_ROWS32: return acc.rows_fp32;
        default: return acc.lowrank_fp32;
        }
    }
    // account finished launches (all of them when wait)
    void flush(bool wait) {
        size_t keep = 0;
        for (size_t i = 0; i < pending.size(); i++) {
            Entry &e = pending[i];
            bool done = wait ? (hipEventSynchronize(e.b) == hipSuccess) : (hipEventQuery(e.b) == hipSuccess);
            if (!done) {
                pending[keep++] = e;
                continue;
            }
            const bool ran = e.slot < 0 || e.slot >= (int)slot_ok.size() || slot_ok[e.slot];
            if (ran) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e.a, e.b);
                arh_kernel_prof &s = slot(e.kind);
                s.launches++, s.ms += ms, s.bytes += e.bytes, s.flops += e.flops;
            }
            pool.push_back(e.a), pool.push_back(e.b);
        }
        pending.resize(keep);
    }
    ~Profiler() {
        for (auto &e : pending) (void)hipEventDestroy(e.a), (void)hipEventDestroy(e.b);
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};
}  // namespace arh

struct arh_ctx_s {
    arh::Profiler prof;
    int refs = 1;          // the caller's handle + one per live geometry
    bool closed = false;   // arh_ctx_destroy called; freed when the last geometry goes
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    // reduction / scalar scratch reused by the solver
    arh::DBuf red;      // partial sums
    arh::DBuf scal;     // per-rhs scalars
    arh::DBuf hostpin;  // pinned host staging
    size_t hostpin_bytes = 0;
    void *pinned(size_t bytes);
};

// FP64 product work unit: up to 64 consecutive columns of one row leaf's column list, taken from
// its dense panel (kind 0) or from the leaf's segment of low-rank block blk's U (kind 1).
struct MvUnit {
    int leaf, kind, blk, c0, nc;
};

// Dense-leaf assembly task: row leaf, first panel column, column count (cols * m <= 256).
struct DenseTask {
    int leaf, c0, nc, pad;
};

// MFMA product: the low-rank factors of all blocks with one column cluster (forward: V) or one
// row cluster (transposed: U^T), stacked row-wise, column-major: Vs[v_off + c rtot + p], p < rtot =
// the sum of the blocks' ranks, c < n = the cluster's elements starting at cs.
struct VStack {
    int cs, n, rtot, pad;
    int64_t v_off;  // offset in the stacked factor
    int64_t t_row;  // row of the stack's first intermediate in the packed [x; t] buffer
};

// MFMA product layout builder: block blk's factor (V forward, U^T transposed) goes to stacked rows
// starting at dst (c = 0, p = 0) of a stack with rtot rows.
struct StackMap {
    int blk, rtot;
    int64_t dst;
};

// MFMA product layout builder: one panel column gathered from the FP64 operator, entries off +
// i * stride of the dense panels (kind 1), U (kind 2) or V (kind 3); kind 0 is a zero column.
struct ColSrc {
    int64_t off;
    int stride, kind;
};

// MFMA product: one leaf's fused panel (forward: the row leaf's dense columns, then the U columns
// of the low-rank blocks covering it; transposed: the column leaf's transposed dense blocks, then
// the V columns restricted to it), zero columns up to a multiple of 4; m x ncol column-major at
// `off`; the packed [x; t] row feeding panel column c is src[col0 + c].
struct LeafPanel {
    int64_t off;
    int col0, ngroups;  // ngroups = ncol / 4
};

// MFMA product row task: column groups [g0, g1) of row leaf `leaf`; slot >= 0 names the task's
// partial tile when the leaf is split over several tasks, -1 when the task writes y itself.
struct RowTask {
    int leaf, g0, g1, slot;
};

// One orientation (forward or transposed) of the MFMA product: operator copies in their own layout
// ([0] FP32, [1] FP64, each built on first use or at assembly), the packed rhs + intermediates
// buffer ([x; t] rows of RP complex), split leaves' partial tiles, and the shared work lists:
// stacks, phase-1 units (stack, 32-row tile; stacks with many columns first, one workgroup each),
// per-leaf panels and their source rows, row tasks, split leaves (leaf, first partial slot, task
// count), layout builder maps.
struct MfmaLayout {
    arh::DBuf panel[2], vs[2], xt[2], part[2];
    bool built[2] = {false, false};
    bool lists = false;
    arh::DBuf d_stacks, d_p1units, d_lpanel, d_psrc, d_rtasks, d_rsplit, d_maps, d_csrc, d_trow;
    std::vector<VStack> stacks;
    std::vector<int2> p1units;
    int p1big = 0, nmaps = 0;
    std::vector<LeafPanel> lpanel;
    std::vector<RowTask> rtasks;
    std::vector<int4> rsplit;
    int nslots = 0;
    int64_t panel_total = 0, vs_total = 0;
};

// Low-rank block descriptor (device). Offsets are in complex elements.
struct LRBlock {
    int rs, re, cs, ce;
    int lim;        // min(max_rank, m-1, n-1)
    int rank;       // stored rank (crosses with u, v != 0); filled by the ACA
    int rank_ref;   // rank as the reference counts it (aca.cpp:80)
    int l0;         // first row leaf
    int64_t us_off, vs_off;  // ACA scratch (U: m x lim col-major, V: lim x n row-major)
    int64_t u_off, v_off;    // packed storage (U leaf-segmented, V row-major)
    int64_t t_off;           // offset of the block's r intermediate values
    int64_t flag_off;        // used-row flags (m bytes)
};

struct arh_geom_s {
    int refs = 1;          // the caller's handle + one per live H-matrix
    bool closed = false;
    arh_ctx *ctx = nullptr;
    int V = 0, N = 0;
    arh::DBuf P, E, tri;  // Ps (V x 3), Es (N x 3), Tri records
    // partition (host copies)
    std::vector<int> dense;    // nd x 4
    std::vector<int> lowrank;  // na x 4
    int nd = 0, na = 0;
    // leaves
    int nleaf = 0, max_leaf = 0;
    std::vector<int> leaf_start;  // nleaf + 1
    std::vector<int> leaf_of;     // N
    arh::DBuf d_leaf_start, d_leaf_of;
    // dense panels: per row leaf, the row's dense blocks side by side (m x ncol column-major)
    std::vector<int> pcol_ptr;        // nleaf + 1: first panel column of each row leaf
    std::vector<int64_t> panel_off;   // nleaf + 1: complex offset of each panel
    std::vector<int> pcol_idx;        // global column (element) of each panel column
    std::vector<int> dblk_leafr, dblk_leafc, dblk_col;  // per dense block: row leaf, col leaf, panel col
    arh::DBuf d_pcol_ptr, d_panel_off, d_pcol_idx;
    int64_t dense_entries = 0;
    // transposed dense: per col leaf, list of dense blocks
    std::vector<int> dT_ptr, dT_blk;
    arh::DBuf d_dT_ptr, d_dT_blk, d_dblk_info;  // dblk_info: (row leaf, panel col) per block
    // dense assembly tasks
    std::vector<DenseTask> dtasks;
    arh::DBuf d_dtasks;
    // low-rank: per row leaf / per col leaf lists of blocks
    std::vector<int> lrR_ptr, lrR_blk, lrC_ptr, lrC_blk;
    arh::DBuf d_lrR_ptr, d_lrR_blk, d_lrC_ptr, d_lrC_blk;
    // ACA scratch sized for the largest max_rank seen
    int scratch_rank = -1;
    arh::DBuf aca_U, aca_V, aca_flags;
    int64_t aca_u_total = 0, aca_v_total = 0, aca_f_total = 0;
    std::vector<int64_t> us_off, vs_off, flag_off;
    std::vector<int> lim;
    // ACA launch order (largest work first, workgroup blocks leading): fixed by the partition and
    // max_rank, so it is computed once per max_rank
    int aca_order_rank = -1, aca_nbig = 0;
    arh::DBuf d_aca_order;
};

struct arh_hmat_s {
    arh_geom *g = nullptr;
    double k = 0;
    int storage = ARH_FP64;
    int max_rank = 0;
    double aca_tol = 0;
    arh::DBuf dense, dense32;    // panels (c128 / c64), column-major per leaf
    arh::DBuf denseT, denseT32;  // the same panels row-major per leaf (generic transposed product, built on use)
    arh::DBuf blocks;          // LRBlock[na]
    arh::DBuf U, V, U32, V32;  // packed factors
    std::vector<LRBlock> hblocks;
    int64_t u_total = 0, v_total = 0, t_total = 0;
    std::vector<int> active_lr;  // lr blocks with rank > 0
    arh::DBuf d_active_lr;
    arh::DBuf tbuf;  // intermediate t / s values (t_total x nrhs c128)
    MfmaLayout mf[2];  // MFMA product: [0] forward, [1] transposed
    bool generic32 = false;  // dense32 / U32 / V32 (the non-MFMA FP32 product) built
    // FP64 product: phase-1 units (block, 4-row group), phase-2 units per leaf, per-unit partials
    std::vector<MvUnit> mv_units;
    std::vector<int> mv_uptr;
    std::vector<int2> p1units64;
    arh::DBuf d_mv_units, d_mv_uptr, d_p1units64, mv_part;
    arh_hmat_info info{};
};

// Listener ACA of the LL x N receiver matrix: U (LL x lim, column-major) and V (lim x N,
// row-major) factors of which the first `rank` crosses are kept; device loop state, partial
// reductions, used-row flags and the listener coordinates.
struct arh_qaca_s {
    arh_geom *g = nullptr;
    double k = 0, tol = 0;
    int LL = 0, N = 0, lim = 0, rank = 0;
    int nbr = 0, nbn = 0, nbc = 0;  // residual row blocks (N), row normalisation blocks (N), column blocks (LL)
    arh::DBuf L, U, V, used, part, dots, st, t;
};

// Relief parameterisation: taps of the differentiable vertices in a tex_h x tex_w heightfield
// (four texel indices and weights per vertex) and, per texel, the CSR list of (vertex, tap) entries.
struct arh_relief_s {
    arh_ctx *ctx = nullptr;
    int n = 0, h = 0, w = 0;
    arh::DBuf idx, wts, ptr, ent;
};

namespace arh {
// geometry.hip
void geom_set_heights(arh_geom *g, const double *base, const int *vid_dev, int HH, const double *h);
void geom_build_tri(arh_geom *g);
// assemble.hip
void assemble_dense(arh_hmat *h);
void assemble_aca(arh_hmat *h);
void make_fp32_copy(arh_hmat *h);
void make_generic_fp32(arh_hmat *h);
// matvec.hip
// y = (M - I/2) x (or its transpose) for nrhs columns. cols (device, ncols entries) restricts the
// MFMA product to those columns (other products compute all); gate (device int) lets the MFMA
// product skip itself when 0. Both are optional.
void hmat_apply(arh_hmat *h, int transpose, int precision, int nrhs, const double2 *x, double2 *y, hipStream_t s,
                const int *cols = nullptr, int ncols = 0, const int *gate = nullptr);
// matvec_mfma.hip
void hmat_apply_mfma(arh_hmat *h, int transpose, int fp64, int nrhs, const double2 *x, double2 *y, hipStream_t s,
                     const int *cols = nullptr, int ncols = 0, const int *gate = nullptr);
void mfma_prepare(arh_hmat *h, int transpose, int fp64);
// matvec64.hip
void hmat_apply64(arh_hmat *h, int nrhs, const double2 *x, double2 *y, hipStream_t s);
int mfma_max_leaf();
// rhs / receivers / diffusion (receivers.hip)
void launch_rhs(arh_geom *g, double k, int kind, int nrhs, const double *vecs_dev, double2 *b);
void launch_receivers(arh_geom *g, double k, const double *L_dev, int LL, int transpose, int nrhs, const double2 *x,
                      double2 *y);
// Receivers with the coordinates in a static per-GPU array (no device allocation); at most
// kNoAllocMaxReceivers receivers, devices 0 .. kNoAllocMaxDevices - 1.
constexpr int kNoAllocMaxReceivers = 16384;
constexpr int kNoAllocMaxDevices = 64;
void launch_receivers_noalloc(arh_geom *g, double k, const double *L_host, int LL, int transpose, int nrhs,
                              const double2 *x, double2 *y);
void launch_scale(hipStream_t s, double2 *a, int64_t n, double f);
void launch_add(hipStream_t s, const double *x, double *y, int64_t n);
void launch_diffusion(hipStream_t s, int LL, int nrhs, const double2 *y, double *d, double *u, double2 *vf);
// relief.hip
void relief_build(arh_relief *r, const double *uvs);
void relief_eval(arh_relief *r, const double *tex, double *heights);
void relief_adjoint(arh_relief *r, const double *grad, double *tex_grad);
// qaca.hip
void qaca_build(arh_qaca *q, const double *L_host);
void qaca_apply(arh_qaca *q, int transpose, int nrhs, const double2 *x, double2 *y);
// solver.hip
void solve(arh_hmat *h, int transpose, int nrhs, const double2 *b, double2 *x, const arh_solve_opts &o,
           arh_solve_stats *st);
// grad.hip
void launch_shape_grad(arh_geom *g, double k, const double *L_dev, int LL, const int *diff, int HH, int nrhs,
                       const double2 *x, const double2 *lam, const double2 *vf, int src_kind, const double *src_dev,
                       double *A, double *B, double *C, double *G = nullptr);
}  // namespace arh
