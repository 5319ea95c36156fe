// Listener ACA on gfx950: one partially pivoted ACA over the whole LL x N receiver matrix
// Q_lj = entry(listener l, element j) (compute_ACA_listener, aca.cpp:84-134), then y = U (V x) and
// its transpose V^T (U^T x) (adjoint.cpp:101-102, 177-178).
//
// Unlike the block ACA (one workgroup per block), this is a single large block, so every stage of a
// cross is a grid-wide kernel: residual row (N entries) with block-partial first-max; pivot, row
// normalisation and residual column (LL entries, masked first-max) in one launch; the non-conjugated dot
// products with the stored crosses (chunk partials over a 2-D grid); and a one-workgroup finish that
// applies the stopping rule. A residual row or column block takes 64 entries (one per lane) and splits the sum
// over the stored crosses between its kQS wavefronts, so the grid grows with the rank as well as with
// N or LL. All reductions are two-stage with a fixed order. The loop control lives on the device (a `done` flag every kernel
// checks), so the host launches a few crosses at a time and synchronises once per batch.
//
// Semantics kept from aca.cpp:84-134: start at row 0 (used = {0}); V row = v / v_j with no zero-pivot
// guard (aca.cpp:104); error estimate with |u|^2 |v|^2 + 2 sum |u_p^T u| |v v_p^T|; stop when
// sqrt(|u|^2 |v|^2 / err) <= eps discarding the current cross; crosses are kept even when zero;
// iteration limit min(max_rank, LL - 1, N - 1).
#include <algorithm>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

constexpr int kQT = 256;  // threads per block
constexpr int kQS = 8;    // wavefronts of a residual row / column block (slices of the stored crosses)

struct QAcaState {
    double err;
    int i, it, done, lim;
};

// block first-max of (value, index); ties to the lower index; value -1 = none
__device__ __forceinline__ void block_argmax_q(double &v, int &idx, double *sv, int *si) {
    wave_argmax(v, idx);
    const int w = threadIdx.x / ARH_WAVE, lane = threadIdx.x % ARH_WAVE;
    __syncthreads();
    if (lane == 0) sv[w] = v, si[w] = idx;
    __syncthreads();
    double bv = sv[0];
    int bi = si[0];
    for (int q = 1; q < kQT / ARH_WAVE; q++)
        if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) bv = sv[q], bi = si[q];
    v = bv;
    idx = bi;
}

__device__ __forceinline__ double block_sum_q(double v, double *sv) {
    v = wave_sum(v);
    const int w = threadIdx.x / ARH_WAVE, lane = threadIdx.x % ARH_WAVE;
    __syncthreads();
    if (lane == 0) sv[w] = v;
    __syncthreads();
    double s = sv[0];
    for (int q = 1; q < kQT / ARH_WAVE; q++) s += sv[q];
    return s;
}

// Slice w of the stored-cross sum of a residual entry: sum over p in wave w's share of [0, it) of
// a(p) b(p), where the lane's own operand is own[p * ostride] and the shared one is shared[p * sstride].
__device__ __forceinline__ double2 cross_slice(int it, int w, const double2 *__restrict__ own, int64_t ostride,
                                               const double2 *__restrict__ shared, int64_t sstride) {
    const int chunk = (it + kQS - 1) / kQS, p0 = w * chunk, p1 = min(it, p0 + chunk);
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int p = p0; p < p1; p++) {
        const double2 a = shared[p * sstride], b = own[p * ostride];
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
    }
    return acc;
}

// entry - (slice 0 + ... ): wave 0 subtracts the kQS slice sums in wave order (the other waves leave
// theirs in red and return false)
__device__ __forceinline__ bool combine_slices(double2 &v, const double2 &mine, double2 (*red)[ARH_WAVE], int w,
                                               int lane) {
    if (w > 0) red[w - 1][lane] = mine;
    __syncthreads();
    if (w != 0) return false;
    v.x -= mine.x;
    v.y -= mine.y;
#pragma unroll
    for (int q = 0; q < kQS - 1; q++) v.x -= red[q][lane].x, v.y -= red[q][lane].y;
    return true;
}

// residual row i into V row `it` (unnormalised) + block partial first-max of |v|^2 (64 columns a block)
__global__ __launch_bounds__(kQS * ARH_WAVE) void k_qa_row(const QAcaState *__restrict__ S,
                                                          const double *__restrict__ L, const Tri *__restrict__ tri,
                                                          double k, int LL, int N, const double2 *__restrict__ U,
                                                          double2 *__restrict__ V, double *__restrict__ pmax) {
    __shared__ double2 red[kQS - 1][ARH_WAVE];
    const QAcaState st = *S;
    if (st.done) return;
    const int i = st.i, it = st.it;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / ARH_WAVE)), lane = threadIdx.x % ARH_WAVE;
    const int c = blockIdx.x * ARH_WAVE + lane, cc = min(c, N - 1);  // loads clamped, stores guarded
    double2 v = make_double2(0.0, 0.0);
    if (w == 0) v = bem_entry(k, L[3 * i], L[3 * i + 1], L[3 * i + 2], tri[cc]);
    const double2 mine = cross_slice(it, w, V + cc, N, U + i, LL);
    if (!combine_slices(v, mine, red, w, lane)) return;
    double best = -1.0;
    int bi = 0x7fffffff;
    if (c < N) {
        V[(int64_t)it * N + c] = v;
        const double a = cnorm2(v);
        best = a == a ? a : -1.0, bi = c;  // NaN never wins
    }
    wave_argmax(best, bi);
    // the partial carries v at the block's first maximum, so the pivot value is known without reading the
    // row while it is being normalised
    const int src = bi - (int)blockIdx.x * ARH_WAVE;
    const int sl = src >= 0 && src < ARH_WAVE ? src : 0;
    const double vr = __shfl(v.x, sl, ARH_WAVE), vi = __shfl(v.y, sl, ARH_WAVE);
    if (lane == 0) {
        double *o = pmax + 4 * blockIdx.x;
        o[0] = best, o[1] = (double)bi, o[2] = vr, o[3] = vi;
    }
}

// One launch after the residual row, workgroups [0, nbn) normalising V row `it` (512 columns each, block
// partials of |v|^2) and [nbn, nbn + nbc) forming the residual column j (64 rows each). Every workgroup
// first finds the pivot column j = the first maximum over the row blocks' partials (ties to the lower
// index, so it is the sequential first maximum) and takes v_j from the winning partial.
__global__ __launch_bounds__(kQS * ARH_WAVE) void k_qa_normcol(const QAcaState *__restrict__ S,
                                                              const double *__restrict__ L,
                                                              const Tri *__restrict__ tri, double k, int LL, int N,
                                                              double2 *__restrict__ U, double2 *__restrict__ V,
                                                              const unsigned char *__restrict__ used,
                                                              const double *__restrict__ pmax, int nbr, int nbn,
                                                              double *__restrict__ pvn, double *__restrict__ pcol) {
    __shared__ double2 red[kQS - 1][ARH_WAVE];
    __shared__ double sv[kQS];
    __shared__ int sj;
    __shared__ double2 spiv;
    const QAcaState st = *S;
    if (st.done) return;
    const int it = st.it;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / ARH_WAVE)), lane = threadIdx.x % ARH_WAVE;
    if (w == 0) {
        double bv = -1.0;
        int bj = 0x7fffffff;
        for (int b = lane; b < nbr; b += ARH_WAVE) {
            const double v = pmax[4 * b];
            const int j = (int)pmax[4 * b + 1];
            if (v > bv || (v == bv && j < bj)) bv = v, bj = j;
        }
        wave_argmax(bv, bj);
        if (lane == 0) {
            const double *o = pmax + 4 * (bj / ARH_WAVE);  // the row block holding column bj
            sj = bj, spiv = make_double2(o[2], o[3]);
        }
    }
    __syncthreads();
    const int j = sj;
    if ((int)blockIdx.x < nbn) {
        // V row `it` /= v_j (no zero-pivot guard, aca.cpp:104) + block partial of |v|^2
        const int c = blockIdx.x * (kQS * ARH_WAVE) + threadIdx.x;
        double a = 0.0;
        if (c < N) {
            const double2 v = cdiv(V[(int64_t)it * N + c], spiv);
            V[(int64_t)it * N + c] = v;
            a = cnorm2(v);
        }
        a = wave_sum(a);
        if (lane == 0) sv[w] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = sv[0];
            for (int q = 1; q < kQS; q++) t += sv[q];
            pvn[blockIdx.x] = t;
        }
        return;
    }
    // residual column j into U column `it` + block partials of |u|^2 and of the masked first-max
    const int bc = (int)blockIdx.x - nbn;
    const int l = bc * ARH_WAVE + lane, ll = min(l, LL - 1);
    double2 u = make_double2(0.0, 0.0);
    if (w == 0) u = bem_entry(k, L[3 * ll], L[3 * ll + 1], L[3 * ll + 2], tri[j]);
    const double2 mine = cross_slice(it, w, U + ll, LL, V + j, N);
    if (!combine_slices(u, mine, red, w, lane)) return;
    double a = 0.0, best = -1.0;
    int bi = 0x7fffffff;
    if (l < LL) {
        U[(int64_t)it * LL + l] = u;
        a = cnorm2(u);
        if (!used[l]) best = a == a ? a : -1.0, bi = l;
    }
    a = wave_sum(a);
    wave_argmax(best, bi);
    if (lane == 0) pcol[3 * bc] = a, pcol[3 * bc + 1] = best, pcol[3 * bc + 2] = (double)bi;
}

// Chunk partials of the non-conjugated products u^T u_p and v v_p^T with the stored crosses p < it
// (aca.cpp:64-69 / :114-116): block (p, c) sums chunk c of kQC entries, chunks [0, ncu) over U's
// LL rows and [ncu, ncu + ncv) over V's N columns; pu[p * ncu + c], pv[p * ncv + c'].
constexpr int kQC = 8 * kQT;
__global__ __launch_bounds__(kQT) void k_qa_dots(const QAcaState *__restrict__ S, int LL, int N, int ncu, int ncv,
                                                const double2 *__restrict__ U, const double2 *__restrict__ V,
                                                double2 *__restrict__ pu, double2 *__restrict__ pv) {
    __shared__ double sv[kQT / ARH_WAVE];
    const QAcaState st = *S;
    const int p = blockIdx.x, c = blockIdx.y;
    if (st.done || p >= st.it) return;
    const bool isu = c < ncu;
    const int n = isu ? LL : N, c0 = (isu ? c : c - ncu) * kQC;
    const double2 *a = isu ? U + (int64_t)st.it * LL : V + (int64_t)st.it * N;
    const double2 *b = isu ? U + (int64_t)p * LL : V + (int64_t)p * N;
    double2 acc = make_double2(0.0, 0.0);
    const int e1 = min(n, c0 + kQC);
    for (int e = c0 + threadIdx.x; e < e1; e += kQT) cfma(acc, a[e], b[e]);
    acc.x = block_sum_q(acc.x, sv);
    acc.y = block_sum_q(acc.y, sv);
    if (threadIdx.x == 0) {
        if (isu)
            pu[(int64_t)p * ncu + c] = acc;
        else
            pv[(int64_t)p * ncv + (c - ncu)] = acc;
    }
}

// error estimate, stopping rule, next row (one workgroup; fixed-order reductions)
__global__ __launch_bounds__(kQT) void k_qa_finish(QAcaState *__restrict__ S, double eps,
                                                  const double *__restrict__ pvn, int nbn,
                                                  const double *__restrict__ pcol, int nbc,
                                                  const double2 *__restrict__ pu, int ncu,
                                                  const double2 *__restrict__ pv, int ncv,
                                                  unsigned char *__restrict__ used) {
    __shared__ double sv[kQT / ARH_WAVE];
    __shared__ int si[kQT / ARH_WAVE];
    const QAcaState st0 = *S;
    if (st0.done) return;
    double vn = 0.0, un = 0.0, d = 0.0, bv = -1.0;
    int bi = 0x7fffffff;
    for (int b = threadIdx.x; b < nbn; b += kQT) vn += pvn[b];
    for (int b = threadIdx.x; b < nbc; b += kQT) {
        un += pcol[3 * b];
        const double v = pcol[3 * b + 1];
        const int l = (int)pcol[3 * b + 2];
        if (v > bv || (v == bv && l < bi)) bv = v, bi = l;
    }
    for (int p = threadIdx.x; p < st0.it; p += kQT) {
        double2 su = make_double2(0.0, 0.0), sq = make_double2(0.0, 0.0);
        for (int c = 0; c < ncu; c++) su.x += pu[(int64_t)p * ncu + c].x, su.y += pu[(int64_t)p * ncu + c].y;
        for (int c = 0; c < ncv; c++) sq.x += pv[(int64_t)p * ncv + c].x, sq.y += pv[(int64_t)p * ncv + c].y;
        d += cabs_(su) * cabs_(sq);
    }
    vn = block_sum_q(vn, sv);
    un = block_sum_q(un, sv);
    d = block_sum_q(d, sv);
    block_argmax_q(bv, bi, sv, si);
    if (threadIdx.x != 0) return;
    QAcaState st = st0;
    st.err += un * vn + 2.0 * d;
    if (sqrt(un * vn / st.err) <= eps) {
        st.done = 1;  // current cross discarded: rank = it
    } else {
        st.i = bv < 0.0 ? 0 : bi;
        used[st.i] = 1;
        st.it++;
        if (st.it >= st.lim) st.done = 1;
    }
    *S = st;
}

// products: t = V x (rank x nrhs, one block per (p, j)); y = U t
__global__ __launch_bounds__(kQT) void k_qa_vx(int N, int r, const double2 *__restrict__ V, const double2 *__restrict__ x,
                                              double2 *__restrict__ t) {
    __shared__ double sv[kQT / ARH_WAVE];
    const int p = blockIdx.x, j = blockIdx.y;
    double2 s = make_double2(0.0, 0.0);
    for (int c = threadIdx.x; c < N; c += kQT) {
        const double2 a = V[(int64_t)p * N + c], b = x[(int64_t)j * N + c];
        s.x += a.x * b.x - a.y * b.y;
        s.y += a.x * b.y + a.y * b.x;
    }
    s.x = block_sum_q(s.x, sv);
    s.y = block_sum_q(s.y, sv);
    if (threadIdx.x == 0) t[(int64_t)j * r + p] = s;
}

__global__ void k_qa_ut(int LL, int r, int nrhs, const double2 *__restrict__ U, const double2 *__restrict__ t,
                        double2 *__restrict__ y) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)LL * nrhs) return;
    const int j = (int)(e / LL), l = (int)(e % LL);
    double2 s = make_double2(0.0, 0.0);
    for (int p = 0; p < r; p++) {
        const double2 a = U[(int64_t)p * LL + l], b = t[(int64_t)j * r + p];
        s.x += a.x * b.x - a.y * b.y;
        s.y += a.x * b.y + a.y * b.x;
    }
    y[e] = s;
}

// transposed: s = U^T x (rank x nrhs, one block per (p, j)); y = V^T s
__global__ __launch_bounds__(kQT) void k_qa_utx(int LL, int r, const double2 *__restrict__ U,
                                               const double2 *__restrict__ x, double2 *__restrict__ t) {
    __shared__ double sv[kQT / ARH_WAVE];
    const int p = blockIdx.x, j = blockIdx.y;
    double2 s = make_double2(0.0, 0.0);
    for (int l = threadIdx.x; l < LL; l += kQT) {
        const double2 a = U[(int64_t)p * LL + l], b = x[(int64_t)j * LL + l];
        s.x += a.x * b.x - a.y * b.y;
        s.y += a.x * b.y + a.y * b.x;
    }
    s.x = block_sum_q(s.x, sv);
    s.y = block_sum_q(s.y, sv);
    if (threadIdx.x == 0) t[(int64_t)j * r + p] = s;
}

__global__ void k_qa_vts(int N, int r, int nrhs, const double2 *__restrict__ V, const double2 *__restrict__ t,
                         double2 *__restrict__ y) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)N * nrhs) return;
    const int j = (int)(e / N), c = (int)(e % N);
    double2 s = make_double2(0.0, 0.0);
    for (int p = 0; p < r; p++) {
        const double2 a = V[(int64_t)p * N + c], b = t[(int64_t)j * r + p];
        s.x += a.x * b.x - a.y * b.y;
        s.y += a.x * b.y + a.y * b.x;
    }
    y[e] = s;
}

// build the factors (host loop over batches of crosses)
void qaca_build(arh_qaca *q, const double *L_host) {
    arh_geom *g = q->g;
    hipStream_t s = g->ctx->stream;
    const int LL = q->LL, N = q->N, lim = q->lim;
    const std::vector<double> hL(L_host, L_host + 3 * (size_t)LL);  // alive until the stream is synchronised
    upload(q->L, hL, s);
    q->U.alloc(sizeof(double2) * std::max<int64_t>(1, (int64_t)LL * lim));
    q->V.alloc(sizeof(double2) * std::max<int64_t>(1, (int64_t)lim * N));
    q->nbr = (N + ARH_WAVE - 1) / ARH_WAVE;               // residual row blocks
    q->nbn = (N + kQS * ARH_WAVE - 1) / (kQS * ARH_WAVE);  // row normalisation blocks
    q->nbc = (LL + ARH_WAVE - 1) / ARH_WAVE;               // residual column blocks
    q->part.alloc(sizeof(double) * (4 * (size_t)q->nbr + q->nbn + 3 * (size_t)q->nbc + 16));
    const int ncu = (LL + kQC - 1) / kQC, ncv = (N + kQC - 1) / kQC;
    q->dots.alloc(sizeof(double2) * (size_t)std::max(1, lim) * (ncu + ncv));
    std::vector<unsigned char> used(LL, 0);
    used[0] = 1;
    upload(q->used, used, s);
    QAcaState st0{};
    st0.i = 0, st0.it = 0, st0.done = lim <= 0 ? 1 : 0, st0.lim = lim;
    q->st.alloc(sizeof(QAcaState));
    ARH_HIP(hipMemcpyAsync(q->st.p, &st0, sizeof(QAcaState), hipMemcpyHostToDevice, s));
    double *pmax = q->part.as<double>(), *pvn = pmax + 4 * q->nbr, *pcol = pvn + q->nbn;
    QAcaState *S = q->st.as<QAcaState>();
    const Tri *tri = g->tri.as<Tri>();
    const double2 *Uc = q->U.as<double2>();
    const double2 *Vc = q->V.as<double2>();
    double2 *pu = q->dots.as<double2>(), *pv = pu + (size_t)std::max(1, lim) * ncu;
    int launched = 0;  // crosses launched so far (an upper bound of it on the device)
    QAcaState hs = st0;
    while (!hs.done) {
        // a few crosses per synchronisation at first, more once the rank grows (launches after the
        // stop return at once)
        const int batch = std::min(std::max(8, std::min(32, launched / 4)), lim - launched);
        for (int b = 0; b < batch; b++) {
            const int itb = launched + b;  // the cross index if none of the previous ones stopped
            hipLaunchKernelGGL(k_qa_row, dim3(q->nbr), dim3(kQS * ARH_WAVE), 0, s, S, q->L.as<double>(), tri, q->k,
                               LL, N, Uc, q->V.as<double2>(), pmax);
            hipLaunchKernelGGL(k_qa_normcol, dim3(q->nbn + q->nbc), dim3(kQS * ARH_WAVE), 0, s, S, q->L.as<double>(),
                               tri, q->k, LL, N, q->U.as<double2>(), q->V.as<double2>(), q->used.as<unsigned char>(),
                               pmax, q->nbr, q->nbn, pvn, pcol);
            if (itb > 0)
                hipLaunchKernelGGL(k_qa_dots, dim3(itb, ncu + ncv), dim3(kQT), 0, s, S, LL, N, ncu, ncv, Uc, Vc, pu,
                                   pv);
            hipLaunchKernelGGL(k_qa_finish, dim3(1), dim3(kQT), 0, s, S, q->tol, pvn, q->nbn, pcol, q->nbc, pu, ncu,
                               pv, ncv, q->used.as<unsigned char>());
        }
        ARH_HIP(hipGetLastError());
        launched += batch;
        ARH_HIP(hipMemcpyAsync(&hs, q->st.p, sizeof(QAcaState), hipMemcpyDeviceToHost, s));
        ARH_HIP(hipStreamSynchronize(s));
        ARH_CHECK(hs.done || launched < lim, ARH_EINVAL, "internal: listener ACA ran past its limit");
    }
    q->rank = hs.it;
}

void qaca_apply(arh_qaca *q, int transpose, int nrhs, const double2 *x, double2 *y) {
    hipStream_t s = q->g->ctx->stream;
    const int r = q->rank;
    const int64_t ny = (int64_t)(transpose ? q->N : q->LL) * nrhs;
    if (r == 0) {
        ARH_HIP(hipMemsetAsync(y, 0, sizeof(double2) * ny, s));
        return;
    }
    q->t.alloc(sizeof(double2) * (size_t)r * nrhs);
    if (!transpose) {
        hipLaunchKernelGGL(k_qa_vx, dim3(r, nrhs), dim3(kQT), 0, s, q->N, r, q->V.as<double2>(), x, q->t.as<double2>());
        hipLaunchKernelGGL(k_qa_ut, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, s, q->LL, r, nrhs,
                           q->U.as<double2>(), q->t.as<double2>(), y);
    } else {
        hipLaunchKernelGGL(k_qa_utx, dim3(r, nrhs), dim3(kQT), 0, s, q->LL, r, q->U.as<double2>(), x,
                           q->t.as<double2>());
        hipLaunchKernelGGL(k_qa_vts, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, s, q->N, r, nrhs,
                           q->V.as<double2>(), q->t.as<double2>(), y);
    }
    ARH_HIP(hipGetLastError());
}

}  // namespace arh
