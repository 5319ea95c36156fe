/*
 * arh.h -- C-ABI of the MI355X (gfx950) Helmholtz BEM library libarh.so.
 *
 * This is the boundary the acoustics3d drop-in module (pybind11, acoustics3d_impl) and any other
 * FFI caller (ctypes, a torch.autograd.Function) bind to. Plain pointers and sizes only.
 *
 * Conventions
 *  - Return value: ARH_OK (0) or a negative ARH_E* code; the message is in arh_last_error(ctx).
 *  - Complex vectors are interleaved (re, im) doubles ("c128"). A batch of nrhs vectors of
 *    length n is stored vector after vector (column-major n x nrhs).
 *  - Pointers documented "device" are HBM addresses (from arh_malloc or any HIP allocation);
 *    "host" pointers are read or written during the call only.
 *  - Mesh: Ps (V x 3 float64, row-major), Es (N x 3 int32, row-major, 0-based), y axis up,
 *    element normal n = (b - a) x (c - a), not normalised (constants.h:30-31, helpers.cpp:16-32).
 *  - Block partition rows/cols refer to the REORDERED element numbering returned by
 *    arh_partition (acoustics3d.cpp:47-48).
 *  - A context owns one HIP stream; a context and its objects are not re-entrant, different
 *    contexts are independent.
 *  - Wavenumbers k (rad/m) must be finite, >= 0 and at most 1e6: the kernels' FP64 sine and cosine
 *    reduce arguments k r accurately up to 1e9 (r in metres); larger k is rejected with ARH_EINVAL.
 *
 * Reference interfaces replaced (mickey1356/acoustic_reliefs, acoustics3d_cpp/):
 *  see the comment on each entry point.
 */
#ifndef ARH_H
#define ARH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARH_OK 0
#define ARH_EINVAL -1   /* invalid argument / shape */
#define ARH_EHIP -2     /* HIP runtime error */
#define ARH_ENAN -3     /* NaN in an assembled block (hmatrix.cpp:25-29 calls exit(1) instead) */
#define ARH_ECONV -4    /* solver did not reach tolerance (only when opted in) */
#define ARH_ENOMEM -5   /* device allocation failed or over the admission budget */
#define ARH_ESTATE -6   /* call order (e.g. not assembled) */
#define ARH_ECANCEL -7  /* serving: the request was cancelled before it was admitted */

/* Storage / arithmetic of an assembled H-matrix. */
#define ARH_FP64 0  /* complex FP64 leaves and factors */
#define ARH_MIXED 1 /* complex FP64 + complex FP32 copies: FP32 inner products on MFMA, FP64 residuals */

/* Right-hand-side kinds. */
#define ARH_RHS_POINT 0 /* b_i = e^{-ik|c_i - s|} / (4 pi |c_i - s|)   (helpers.cpp:46-50) */
#define ARH_RHS_PLANE 1 /* b_i = e^{-ik d . c_i}                         (sphere.h:22-26, d = +x there) */

typedef struct arh_ctx_s arh_ctx;
typedef struct arh_geom_s arh_geom;
typedef struct arh_hmat_s arh_hmat;
typedef struct arh_server_s arh_server;
typedef struct arh_qaca_s arh_qaca;
typedef struct arh_relief_s arh_relief;

typedef struct {
    double tol;        /* relative residual tolerance (BiCGSTAB::setTolerance, adjoint.cpp:93) */
    int maxit;         /* <= 0: 2 N (Eigen's default) */
    int true_residual; /* 1: stop only when the FP64 true residual ||b - A x|| / ||b|| <= tol
                          (restarts BiCGSTAB from the current x when the recursive residual drifted);
                          0: Eigen semantics (stop on the recursive residual) */
    int fail_on_noconv;/* 1: return ARH_ECONV when tol is not reached */
    int hist_cap;      /* capacity (entries per rhs) of stats->hist, 0 = none */
} arh_solve_opts;

typedef struct {
    int iters;          /* max over rhs of BiCGSTAB iterations (Eigen's iterations()) */
    int matvecs;        /* H-matrix products issued (hmatrix.h:54-60 mat_mults) */
    int refinements;    /* FP64 refinement sweeps (ARH_MIXED) */
    int restarts;       /* BiCGSTAB restarts */
    double max_err;     /* max over rhs of the recursive residual ||r|| / ||b|| (Eigen's error()) */
    double max_true_res;/* max over rhs of the FP64 true residual ||b - A x|| / ||b|| */
    int *iters_per_rhs; /* optional host array [nrhs] */
    double *true_res;   /* optional host array [nrhs] */
    double *hist;       /* optional host array [nrhs * hist_cap]: ||r_i|| / ||b|| per iteration */
    int *hist_len;      /* optional host array [nrhs] */
    double solve_ms;    /* wall time of the solve (out) */
    double post_ms;     /* arh_band_solve: wall time of receivers + diffusion coefficients (out) */
} arh_solve_stats;

typedef struct {
    int n;               /* elements */
    int n_dense, n_lowrank, n_leaves;
    int64_t dense_entries;     /* sum m n over dense blocks */
    int64_t lowrank_entries;   /* sum r (m + n) over low-rank blocks (stored) */
    int max_rank;              /* max ACA rank */
    int64_t sum_rank;
    int64_t bytes_fp64;        /* bytes of the complex FP64 operator */
    int64_t bytes_fp32;        /* bytes of the complex FP32 copy (0 unless ARH_MIXED) */
    int64_t bytes_total;       /* all device memory held by the H-matrix */
    double assemble_ms;        /* device time of the assembly (events) */
} arh_hmat_info;

/* Live kernel timing (HIP events on the context's stream) of the two H-matvec kernels:
 * lowrank = the per-block t = V x (or U^T x) pass, rows = the per-leaf dense-panel + U pass (or
 * the per-column-leaf pass when transposed). bytes = algorithmic bytes of the launches: matrix
 * entries read once (16 B complex FP64, 8 B complex FP32), the packed right-hand sides and
 * intermediates once per rhs in the operator's precision (16 B / 8 B), and in the rows pass the
 * FP64 x read and y write once per rhs (16 B each). flops = 8 per complex multiply-add. */
typedef struct {
    int64_t launches;
    double ms;
    double bytes;
    double flops;
} arh_kernel_prof;

typedef struct {
    arh_kernel_prof rows_fp64, lowrank_fp64, rows_fp32, lowrank_fp32;
} arh_profile;

/* ------------------------------------------------------------------ context */
/* New: HIP device + stream + error state. */
int arh_ctx_create(int device, arh_ctx **out);
void arh_ctx_destroy(arh_ctx *ctx);
/* Message of the last failed call on ctx (ctx == NULL: last failure in this thread). */
const char *arh_last_error(const arh_ctx *ctx);
/* Make ctx issue its work on an external HIP stream (e.g. torch's current stream); NULL = own stream. */
int arh_ctx_set_stream(arh_ctx *ctx, void *hip_stream);
int arh_ctx_synchronize(arh_ctx *ctx);
/* Enable (1, counters reset) or disable (0) live kernel timing; read the counters (waits for the
 * recorded launches to finish). */
int arh_ctx_profile(arh_ctx *ctx, int enable);
int arh_ctx_get_profile(arh_ctx *ctx, arh_profile *out);

/* Device memory helpers for FFI callers without a HIP binding. */
/* New (the reference has no device memory). The library keeps the device blocks its handles free
 * and hands them out again (a band or request loop then stops mapping fresh memory); a block is
 * reused on another stream only after the work of its previous stream. arh_device_trim returns
 * every cached block of a device to HIP; arh_device_cached reports the bytes held. Setting
 * ARH_POOL=0 in the environment disables the cache. */
int arh_device_trim(int device, size_t *freed_bytes);
int arh_device_cached(int device, size_t *bytes);
int arh_malloc(arh_ctx *ctx, size_t bytes, void **dptr);
int arh_free(arh_ctx *ctx, void *dptr);
int arh_upload(arh_ctx *ctx, void *dst_device, const void *src_host, size_t bytes);
int arh_download(arh_ctx *ctx, void *dst_host, const void *src_device, size_t bytes);
/* y[i] = y[i] + x[i], i < n (device float64), queued on ctx's stream: sums of device gradients kept in their
 * accumulation order (DiffBEM.gradient adds one band's sum over frequencies at a time, acoustics3d.cpp:224-227). */
int arh_add(arh_ctx *ctx, int64_t n, const double *x, double *y);

/* ------------------------------------------------------------------ set-up (host, once per mesh) */
/* Cluster tree + dual-tree block partition. Replaces ClusterTree(max_pts, Cs) +
 * get_direct_and_approx_blocks(factor, direct, approx) (clustertree.cpp:49-137, called from
 * acoustics3d.cpp:47-58). C: N x 3 centroids of the ORIGINAL element order (host).
 * perm[N] (host out): new position -> original element id (ClusterTree::reordering).
 * blocks (host out, capacity cap quadruples): n_dense then n_lowrank (rs, re, cs, ce) quadruples,
 * each in traversal order. blocks == NULL: only the counts are returned. */
int arh_partition(const double *C, int N, int leaf_size, double radius_factor, int *perm, int *blocks, int cap,
                  int *n_dense, int *n_lowrank);

/* Mesh + frozen partition on the device. Replaces DiffBEM::set_mesh's state
 * (acoustics3d.cpp:36-67). Ps, Es, blocks: host. Es already reordered. */
int arh_geom_create(arh_ctx *ctx, const double *Ps, int V, const int *Es, int N, const int *blocks, int n_dense,
                    int n_lowrank, arh_geom **out);
/* Move vertices (heights applied by the caller), keeping the partition (acoustics3d.cpp:139-146, 210-217). */
int arh_geom_set_vertices(arh_geom *g, const double *Ps_host);
int arh_geom_set_vertices_device(arh_geom *g, const double *Ps_device);
/* The vertices of a deformed mesh on the device (DiffBEM's differentiable points moved in y,
 * acoustics3d.cpp:139-142 and 210-213): Ps = base (device, V x 3) with y of vertex vids[t] (host; distinct ids,
 * or -1 to skip entry t) set to h[t] + base y (h device, HH). Queued on the context's stream; base and h must be
 * ready there. For callers whose heights live on the device (a torch optimiser): no host copy of the heights. */
int arh_geom_set_heights_device(arh_geom *g, const double *base, const int *vids, int HH, const double *h);
void arh_geom_destroy(arh_geom *g);

/* ------------------------------------------------------------------ H-matrix */
/* Near-field Green's-function quadrature of the dense leaves + partially pivoted ACA of the
 * admissible blocks at wavenumber k. Replaces HMatrix::compute_direct_and_approx_blocks_cpu
 * (hmatrix.cpp:10-85) with compute_block (integrate.cpp:87-100) and compute_ACA_block
 * (aca.cpp:38-82). max_rank: APPROX_ACA_MAX_K = 100 in the reference (constants.h:16). */
int arh_hmat_assemble(arh_geom *g, double k, double aca_tol, int max_rank, int storage, arh_hmat **out);
int arh_hmat_get_info(const arh_hmat *h, arh_hmat_info *info);
/* ACA ranks as the reference counts them, per low-rank block in partition order (host out). */
int arh_hmat_ranks(const arh_hmat *h, int *ranks);
/* Dense entries of dense block b (m x n column-major c128, host out) -- for checking. */
int arh_hmat_dense_block(const arh_hmat *h, int b, double *out);
/* Factors of low-rank block a as stored (crosses with u, v != 0): *rank = r; U (m x r column-major) and
 * V (r x n row-major), c128 host out, written when non-NULL. */
int arh_hmat_factors(const arh_hmat *h, int a, double *U, double *V, int *rank);
/* New (no reference counterpart; for callers that keep many H-matrices, as the gradient does
 * without recompute_matrices, adjoint.cpp:104-109): free the H-matrix's derived product copies
 * (FP32 copy, MFMA layouts, work buffers), keeping the assembled FP64 operator. The next product
 * rebuilds what it needs; results are unchanged. */
int arh_hmat_compact(arh_hmat *h);
void arh_hmat_destroy(arh_hmat *h);

/* y = (M - I/2) x or its (non-conjugate) transpose, nrhs vectors, device c128.
 * Replaces the HMatrix Eigen product (hmatrix.h:84-140, set_transposed at :51). precision:
 * ARH_FP64, or ARH_MIXED for the complex-FP32 operator (FP32 accumulate, c128 in/out). */
int arh_hmat_matvec(arh_hmat *h, int transpose, int precision, int nrhs, const double *x, double *y);

/* Solve (M - I/2) x = b (transpose: (M - I/2)^T x = b) for nrhs right-hand sides (device c128).
 * ARH_FP64 hmat: BiCGSTAB with Eigen's semantics per rhs (identity preconditioner, x0 = 0,
 * restarts, recursive-residual stop) -- replaces Eigen::BiCGSTAB<HMatrix> at adjoint.cpp:92-95,
 * 206-209 and acoustics3d.cpp:266-269, 306-309.
 * ARH_MIXED hmat: BiCGSTAB on the complex-FP32 operator (FP64 vectors and inner products) inside
 * FP64 true-residual iterative refinement. */
int arh_solve(arh_hmat *h, int transpose, int nrhs, const double *b, double *x, const arh_solve_opts *opts,
              arh_solve_stats *stats);

/* ------------------------------------------------------------------ right-hand sides, receivers */
/* b (device c128, nrhs x N): point sources src[3*j] (kind POINT) or plane-wave directions
 * (kind PLANE), host arrays of 3*nrhs doubles. compute_G_r (helpers.cpp:46-50) / compute_b
 * (sphere.h:22-26). */
int arh_rhs(arh_geom *g, double k, int kind, int nrhs, const double *vecs, double *b);

/* y (device c128, nrhs x LL) = Q x with Q_lj = M(L_l, element j) computed exactly on the fly
 * (the dense listener matrix, integrate.cpp:46-57; the reference's default approximates it by
 * compute_ACA_listener, aca.cpp:84-134). L: host LL x 3. transpose = 1: y (nrhs x N) = Q^T x
 * with x nrhs x LL (the adjoint right-hand side, adjoint.cpp:198). */
int arh_receivers(arh_geom *g, double k, const double *L, int LL, int transpose, int nrhs, const double *x,
                  double *y);

/* No counterpart in the reference: arh_receivers for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results as arh_receivers; the entries Q_lj are evaluated on the fly and the receiver
 * coordinates are staged in a static per-GPU array of 16384 receivers (384 KiB, held until the
 * process ends), so at most LL = 16384 (ARH_EINVAL above). No hipMalloc / hipFree; work is queued on
 * ctx's stream and the call returns without waiting for it. Calls that overlap on one GPU (other
 * streams or threads) share the array and race. */
int arh_receivers_noalloc(arh_geom *g, double k, const double *L, int LL, int transpose, int nrhs, const double *x,
                          double *y);

/* Listener ACA (compute_ACA_listener, aca.cpp:84-134): partially pivoted cross approximation
 * Q ~= U V of the whole LL x N listener matrix Q_lj = M(L_l, element j), started at row 0, rows
 * divided by the pivot without a zero guard, stopped when sqrt(|u|^2 |v|^2 / err) <= tol (the
 * current cross discarded) or after min(max_rank, LL - 1, N - 1) crosses (the reference passes
 * Q_ACA_MAX_K = 5000, constants.h:17). Built on the device at create; keeps a reference to g.
 * L: host LL x 3. The reference's uses: acoustics3d.cpp:275 and adjoint.cpp:101 (y = U (V x)),
 * adjoint.cpp:177-178 (V^T (U^T vf)), sphere.h:80 (tol = the solver tolerance). */
int arh_qaca_create(arh_geom *g, double k, const double *L, int LL, double tol, int max_rank, arh_qaca **out);
int arh_qaca_rank(const arh_qaca *q, int *rank);
/* y (device c128) = U (V x) for x nrhs x N, y nrhs x LL; transpose = 1: y (nrhs x N) =
 * V^T (U^T x) for x nrhs x LL. */
int arh_qaca_apply(arh_qaca *q, int transpose, int nrhs, const double *x, double *y);
/* Host copies of the kept factors: U (LL x rank, column-major c128), V (rank x N, row-major c128);
 * either may be NULL. */
int arh_qaca_factors(const arh_qaca *q, double *U, double *V);
void arh_qaca_destroy(arh_qaca *q);

/* ------------------------------------------------------------------ relief parameterisation */
/* Heights of the differentiable vertices sampled from a heightfield texture, and the adjoint: eval_uv
 * (acoustics_opt.py:24-34) and the texture-gradient scatter of acoustic_gradient (acoustics_opt.py:45-56),
 * bit for bit, for fixed uvs (host n x 2, in [0, 1]; 0,0 = the texture's first row and column).
 * arh_relief_eval: heights (device, n) from tex (device, tex_h x tex_w row-major).
 * arh_relief_adjoint: tex_grad (device, tex_h x tex_w, overwritten) from grad (device, n). Both return when
 * the result is ready. Keeps a reference to ctx. */
int arh_relief_create(arh_ctx *ctx, const double *uvs, int n, int tex_h, int tex_w, arh_relief **out);
int arh_relief_eval(arh_relief *r, const double *tex, double *heights);
int arh_relief_adjoint(arh_relief *r, const double *grad, double *tex_grad);
void arh_relief_destroy(arh_relief *r);

/* Diffusion coefficient per rhs from y (device c128, nrhs x LL) with n_freqs = 1:
 * yt = |y|, d = ((sum L)^2 - sum L^2) / ((LL-1) sum L^2), L = yt^2 (adjoint.cpp:112-127, 388-397).
 * d (host out, nrhs). u (device out, nrhs x LL, may be NULL) = dd/dyt; vf (device out c128,
 * nrhs x LL, may be NULL) = u conj(y) / |y| (adjoint.cpp:145). */
int arh_diffusion(arh_ctx *ctx, int LL, int nrhs, const double *y, double *d, double *u, double *vf);

/* ------------------------------------------------------------------ shape gradient */
/* d(coefficient)/d(height) contributions of one frequency for nrhs (x, lambda, vf) triples,
 * summed over the triples (adjoint.cpp:219-231, 244-386 restructured per element; heights are
 * the y coordinate of vertices diff[0..HH-1]). x, lam: device c128 nrhs x N; vf: device c128
 * nrhs x LL; L: host LL x 3; src_kind/src: the right-hand side kind and host 3*nrhs vectors.
 * A, B, C: device float64 [HH] outputs (overwritten); the gradient is A - B + C. */
int arh_shape_grad(arh_geom *g, double k, const double *L, int LL, const int *diff, int HH, int nrhs, const double *x,
                   const double *lam, const double *vf, int src_kind, const double *src, double *A, double *B,
                   double *C);

/* arh_shape_grad that adds the gradient instead of returning its terms: grad[h] (device, HH) += (A - B) + C,
 * the order of the reference sum A - B + C (adjoint.cpp:232). */
int arh_shape_grad_accumulate(arh_geom *g, double k, const double *L, int LL, const int *diff, int HH, int nrhs,
                              const double *x, const double *lam, const double *vf, int src_kind, const double *src,
                              double *grad);

/* ------------------------------------------------------------------ one band, batched */
/* The unit of the relief workload for one frequency: assemble, solve nrhs plane waves or point
 * sources to opts->tol, receivers, diffusion coefficient per rhs. d (host out, nrhs); y (device
 * out c128 nrhs x LL, may be NULL); x (device out c128 nrhs x N, may be NULL). */
int arh_band_solve(arh_geom *g, double k, double aca_tol, int max_rank, int storage, int kind, int nrhs,
                   const double *vecs, const double *L, int LL, const arh_solve_opts *opts, double *d, double *y,
                   double *x, arh_solve_stats *stats, arh_hmat_info *info);

/* ------------------------------------------------------------------ serving (continuous batching) */
/* New (the reference has no serving path; its callers loop over DiffBEM.band_value one call at a
 * time, acoustics_pipeline.py:28-33). A server owns `workers` device streams (one arh_ctx each).
 * Requests -- one band solve each (arh_band_solve's unit) -- are admitted in arrival order as
 * soon as a worker is free and their H-matrix footprint fits the HBM budget; admitted requests run
 * concurrently on the device. Every request is computed by the same deterministic kernels whatever
 * else is in flight, so its result is the same alone and in the mix. */
typedef struct {
    int workers;        /* concurrent requests (device streams) */
    int64_t hbm_budget; /* bytes the in-flight requests may hold (0 = 90% of free HBM at creation) */
    int device;
    int maxit;          /* BiCGSTAB iteration limit per request (0 = 2N, Eigen's default) */
} arh_server_opts;

typedef struct {
    int status;            /* 0 queued, 1 running, 2 done, < 0 the request's ARH_E* code */
    int nrhs;
    double t_submit, t_start, t_done; /* seconds on the server's steady clock */
    int64_t footprint;     /* admitted HBM bytes (estimate) */
    int iters;             /* BiCGSTAB iterations (max over rhs) */
    double max_true_res;   /* FP64 true residual (max over rhs) */
    double assemble_ms, solve_ms, post_ms; /* split of the service time (wall, on the worker) */
    double geom_ms;        /* device geometry set-up when the worker's cached mesh changes (wall) */
} arh_request_info;

int arh_server_create(const arh_server_opts *opts, arh_server **out);
/* Register a mesh (original element order; the server partitions and reorders it). Host inputs. */
int arh_server_add_mesh(arh_server *s, const double *Ps, int V, const int *Es, int N, int leaf_size,
                        double radius_factor, int *mesh_id);
/* Queue one band solve; vecs (3*nrhs) and L (3*LL) are copied. Returns immediately. */
int arh_server_submit(arh_server *s, int mesh_id, double k, double aca_tol, int storage, int kind, int nrhs,
                      const double *vecs, const double *L, int LL, double tol, int64_t *request_id);
/* Block until the request is done; d (host, nrhs) receives the diffusion coefficients. */
int arh_server_wait(arh_server *s, int64_t request_id, double *d, arh_request_info *info);
int arh_server_poll(arh_server *s, int64_t request_id, arh_request_info *info);
/* Kernel profile of the server's product kernels (as arh_ctx_profile): enable clears the totals;
 * each request admitted while enabled adds its launches when it completes. */
int arh_server_profile(arh_server *s, int enable);
int arh_server_get_profile(arh_server *s, arh_profile *out);
/* Drop every queued (not yet admitted) request: each ends with status ARH_ECANCEL, and
 * arh_server_wait on it returns ARH_ECANCEL. Requests already running finish normally.
 * n_cancelled (may be NULL) receives how many were dropped. */
int arh_server_cancel(arh_server *s, int *n_cancelled);
/* Cancel what is queued (as arh_server_cancel), let the running requests finish, stop and join the
 * worker threads, and wait until no caller is blocked in arh_server_wait; then free the server.
 * Call it before the process tears down the HIP runtime (e.g. from an exit hook): worker threads
 * still issuing HIP work during that teardown crash the process. */
void arh_server_destroy(arh_server *s);

#ifdef __cplusplus
}
#endif
#endif /* ARH_H */
