// Per-element geometry records and the device-side index structures of a frozen partition.
#include <algorithm>
#include <map>

#include "arh_internal.h"
#include "device_common.h"

namespace arh {

// helpers.cpp:16-32: centroid (a + b + c) / 3 and n = ab x ac (components as the reference writes them).
__global__ void k_tri(const double *__restrict__ P, const int *__restrict__ E, int N, Tri *__restrict__ out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    int ia = E[3 * e], ib = E[3 * e + 1], ic = E[3 * e + 2];
    double ax = P[3 * ia], ay = P[3 * ia + 1], az = P[3 * ia + 2];
    double bx = P[3 * ib], by = P[3 * ib + 1], bz = P[3 * ib + 2];
    double cx = P[3 * ic], cy = P[3 * ic + 1], cz = P[3 * ic + 2];
    Tri t;
    t.ax = ax, t.ay = ay, t.az = az;
    t.bx = bx - ax, t.by = by - ay, t.bz = bz - az;
    t.cx = cx - ax, t.cy = cy - ay, t.cz = cz - az;
    t.nx = t.cz * t.by - t.cy * t.bz;
    t.ny = t.cx * t.bz - t.cz * t.bx;
    t.nz = t.cy * t.bx - t.cx * t.by;
    t.gx = (ax + bx + cx) / 3.0;
    t.gy = (ay + by + cy) / 3.0;
    t.gz = (az + bz + cz) / 3.0;
    t.pad = 0.0;
    out[e] = t;
}

// Vertex vid[t] (>= 0) gets y = h[t] + base y (the y offset of DiffBEM's differentiable points,
// acoustics3d.cpp:139-142); the other coordinates are the base ones (copied before the launch).
__global__ void k_offset_y(const double *__restrict__ base, const int *__restrict__ vid, int HH,
                           const double *__restrict__ h, double *__restrict__ P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= HH) return;
    const int v = vid[t];
    if (v >= 0) P[3 * v + 1] = h[t] + base[3 * v + 1];
}

void geom_set_heights(arh_geom *g, const double *base, const int *vid_dev, int HH, const double *h) {
    hipStream_t s = g->ctx->stream;
    ARH_HIP(hipMemcpyAsync(g->P.p, base, sizeof(double) * 3 * g->V, hipMemcpyDeviceToDevice, s));
    if (HH > 0)
        hipLaunchKernelGGL(k_offset_y, dim3((HH + 255) / 256), dim3(256), 0, s, base, vid_dev, HH, h, g->P.as<double>());
    ARH_HIP(hipGetLastError());
    geom_build_tri(g);
}

void geom_build_tri(arh_geom *g) {
    g->tri.alloc(sizeof(Tri) * (size_t)g->N);
    hipLaunchKernelGGL(k_tri, dim3((g->N + 255) / 256), dim3(256), 0, g->ctx->stream, g->P.as<double>(),
                       g->E.as<int>(), g->N, g->tri.as<Tri>());
    ARH_HIP(hipGetLastError());
}

// Builds leaves, dense panels, transposed lists, low-rank lists and the assembly task list.
void geom_build_index(arh_geom *g) {
    const int N = g->N, nd = g->nd, na = g->na;
    hipStream_t s = g->ctx->stream;
    // leaves = row ranges of dense blocks (every leaf meets itself in a dense block)
    std::map<int, int> starts;
    for (int b = 0; b < nd; b++) {
        int rs = g->dense[4 * b], re = g->dense[4 * b + 1];
        auto it = starts.find(rs);
        ARH_CHECK(it == starts.end() || it->second == re, ARH_EINVAL, "partition: inconsistent dense row ranges");
        starts[rs] = re;
    }
    g->leaf_start.clear();
    int pos = 0;
    for (auto &kv : starts) {
        ARH_CHECK(kv.first == pos && kv.second > pos, ARH_EINVAL, "partition: dense row ranges do not tile [0, N)");
        g->leaf_start.push_back(pos);
        pos = kv.second;
    }
    ARH_CHECK(pos == N, ARH_EINVAL, "partition: dense row ranges do not cover [0, N)");
    g->leaf_start.push_back(N);
    g->nleaf = (int)g->leaf_start.size() - 1;
    const int nl = g->nleaf;
    g->leaf_of.assign(N, -1);
    g->max_leaf = 0;
    for (int l = 0; l < nl; l++) g->max_leaf = std::max(g->max_leaf, g->leaf_start[l + 1] - g->leaf_start[l]);
    for (int l = 0; l < nl; l++)
        for (int i = g->leaf_start[l]; i < g->leaf_start[l + 1]; i++) g->leaf_of[i] = l;
    auto leaf_exact = [&](int a, int b) {  // [a, b) is exactly one leaf
        return a >= 0 && a < N && g->leaf_start[g->leaf_of[a]] == a && g->leaf_start[g->leaf_of[a] + 1] == b;
    };
    auto leaf_union = [&](int a, int b) {  // [a, b) is a union of whole leaves
        return a >= 0 && b <= N && a < b && g->leaf_start[g->leaf_of[a]] == a &&
               (b == N || g->leaf_start[g->leaf_of[b]] == b);
    };
    // dense panels
    g->dblk_leafr.assign(nd, 0);
    g->dblk_leafc.assign(nd, 0);
    g->dblk_col.assign(nd, 0);
    std::vector<std::vector<int>> rowblocks(nl), colblocks(nl);
    for (int b = 0; b < nd; b++) {
        int rs = g->dense[4 * b], re = g->dense[4 * b + 1], cs = g->dense[4 * b + 2], ce = g->dense[4 * b + 3];
        ARH_CHECK(leaf_exact(rs, re) && leaf_exact(cs, ce), ARH_EINVAL, "partition: dense block is not leaf x leaf");
        g->dblk_leafr[b] = g->leaf_of[rs];
        g->dblk_leafc[b] = g->leaf_of[cs];
        rowblocks[g->leaf_of[rs]].push_back(b);
        colblocks[g->leaf_of[cs]].push_back(b);
    }
    g->pcol_ptr.assign(nl + 1, 0);
    g->panel_off.assign(nl + 1, 0);
    g->pcol_idx.clear();
    g->dtasks.clear();
    for (int l = 0; l < nl; l++) {
        int m = g->leaf_start[l + 1] - g->leaf_start[l];
        int c = 0;
        for (int b : rowblocks[l]) {
            g->dblk_col[b] = c;
            for (int j = g->dense[4 * b + 2]; j < g->dense[4 * b + 3]; j++) g->pcol_idx.push_back(j), c++;
        }
        g->pcol_ptr[l + 1] = g->pcol_ptr[l] + c;
        g->panel_off[l + 1] = g->panel_off[l] + (int64_t)m * c;
        int cpt = std::max(1, 256 / m);
        for (int c0 = 0; c0 < c; c0 += cpt) g->dtasks.push_back(DenseTask{l, c0, std::min(cpt, c - c0), 0});
    }
    g->dense_entries = g->panel_off[nl];
    // transposed dense lists
    g->dT_ptr.assign(nl + 1, 0);
    g->dT_blk.clear();
    for (int l = 0; l < nl; l++) {
        for (int b : colblocks[l]) g->dT_blk.push_back(b);
        g->dT_ptr[l + 1] = (int)g->dT_blk.size();
    }
    std::vector<int> dinfo(2 * (size_t)nd);
    for (int b = 0; b < nd; b++) dinfo[2 * b] = g->dblk_leafr[b], dinfo[2 * b + 1] = g->dblk_col[b];
    // low-rank lists
    std::vector<std::vector<int>> lrR(nl), lrC(nl);
    for (int a = 0; a < na; a++) {
        int rs = g->lowrank[4 * a], re = g->lowrank[4 * a + 1], cs = g->lowrank[4 * a + 2], ce = g->lowrank[4 * a + 3];
        ARH_CHECK(leaf_union(rs, re) && leaf_union(cs, ce), ARH_EINVAL,
                  "partition: low-rank block ranges are not unions of leaves");
        for (int l = g->leaf_of[rs]; l < nl && g->leaf_start[l] < re; l++) lrR[l].push_back(a);
        for (int l = g->leaf_of[cs]; l < nl && g->leaf_start[l] < ce; l++) lrC[l].push_back(a);
    }
    g->lrR_ptr.assign(nl + 1, 0);
    g->lrC_ptr.assign(nl + 1, 0);
    g->lrR_blk.clear();
    g->lrC_blk.clear();
    for (int l = 0; l < nl; l++) {
        g->lrR_blk.insert(g->lrR_blk.end(), lrR[l].begin(), lrR[l].end());
        g->lrC_blk.insert(g->lrC_blk.end(), lrC[l].begin(), lrC[l].end());
        g->lrR_ptr[l + 1] = (int)g->lrR_blk.size();
        g->lrC_ptr[l + 1] = (int)g->lrC_blk.size();
    }
    // exact cover check: sum of block areas == N^2 (blocks are disjoint by construction of a tree partition)
    int64_t area = 0;
    for (int b = 0; b < nd; b++)
        area += (int64_t)(g->dense[4 * b + 1] - g->dense[4 * b]) * (g->dense[4 * b + 3] - g->dense[4 * b + 2]);
    for (int a = 0; a < na; a++)
        area += (int64_t)(g->lowrank[4 * a + 1] - g->lowrank[4 * a]) * (g->lowrank[4 * a + 3] - g->lowrank[4 * a + 2]);
    ARH_CHECK(area == (int64_t)N * N, ARH_EINVAL, "partition: blocks do not cover N x N exactly once");

    upload(g->d_leaf_start, g->leaf_start, s);
    upload(g->d_leaf_of, g->leaf_of, s);
    upload(g->d_pcol_ptr, g->pcol_ptr, s);
    upload(g->d_panel_off, g->panel_off, s);
    upload(g->d_pcol_idx, g->pcol_idx, s);
    upload(g->d_dT_ptr, g->dT_ptr, s);
    upload(g->d_dT_blk, g->dT_blk, s);
    upload(g->d_dblk_info, dinfo, s);
    upload(g->d_dtasks, g->dtasks, s);
    upload(g->d_lrR_ptr, g->lrR_ptr, s);
    upload(g->d_lrR_blk, g->lrR_blk, s);
    upload(g->d_lrC_ptr, g->lrC_ptr, s);
    upload(g->d_lrC_blk, g->lrC_blk, s);
    ARH_HIP(hipStreamSynchronize(s));
}

}  // namespace arh
