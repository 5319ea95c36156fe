// Cluster tree over element centroids and the dual-tree block partition (host, once per mesh).
//
// Restates ClusterTree (clustertree.cpp:49-101) and dual_tree_traversal (clustertree.cpp:103-137):
//  * node centre = mean of its centroids (sequential sum in member order, divided by the count),
//    radius = max distance of a member centroid to the centre;
//  * a node with <= leaf_size members is a leaf, otherwise it splits the longest bounding-box
//    axis (first axis on ties) at the centre coordinate, members with coordinate < centre go
//    left, the order of members is preserved on both sides;
//  * leaves in left-to-right order define the element reordering;
//  * from (root, root): admissible (dist > factor (r_s + r_t)) -> low-rank block, two leaves ->
//    dense block, otherwise split the source when the target is a leaf or the source is larger,
//    else split the target; left child before right child.
// The reference recurses without end when a split leaves one side empty (coincident centroids);
// here that raises ARH_EINVAL.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "arh_internal.h"

namespace {

struct PNode {
    int s, e;
    double cx, cy, cz, rad;
    int left, right;
};

struct Builder {
    const double *C;
    int leaf;
    std::vector<int> order;  // working permutation: position -> original element id
    std::vector<PNode> nodes;

    int make(int s, int e) {
        const int n = e - s;
        double sx = 0, sy = 0, sz = 0;
        for (int t = s; t < e; t++) {
            const double *c = C + 3 * (size_t)order[t];
            sx += c[0], sy += c[1], sz += c[2];
        }
        PNode nd{s, e, sx / n, sy / n, sz / n, 0.0, -1, -1};
        double r = 0;
        for (int t = s; t < e; t++) {
            const double *c = C + 3 * (size_t)order[t];
            double dx = c[0] - nd.cx, dy = c[1] - nd.cy, dz = c[2] - nd.cz;
            r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        nd.rad = r;
        nodes.push_back(nd);
        return (int)nodes.size() - 1;
    }

    void build(int N) {
        order.resize(N);
        for (int i = 0; i < N; i++) order[i] = i;
        std::vector<int> stack{make(0, N)};
        std::vector<int> tmp;
        while (!stack.empty()) {
            int id = stack.back();
            stack.pop_back();
            int s = nodes[id].s, e = nodes[id].e;
            if (e - s <= leaf) continue;
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int t = s; t < e; t++) {
                const double *c = C + 3 * (size_t)order[t];
                for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], c[d]), hi[d] = std::max(hi[d], c[d]);
            }
            int dim = 0;
            for (int d = 1; d < 3; d++)
                if (hi[d] - lo[d] > hi[dim] - lo[dim]) dim = d;
            const double gamma = dim == 0 ? nodes[id].cx : (dim == 1 ? nodes[id].cy : nodes[id].cz);
            // stable partition of order[s:e)
            tmp.clear();
            int w = s;
            for (int t = s; t < e; t++) {
                int el = order[t];
                if (C[3 * (size_t)el + dim] < gamma)
                    order[w++] = el;
                else
                    tmp.push_back(el);
            }
            std::copy(tmp.begin(), tmp.end(), order.begin() + w);
            if (w == s || w == e)
                throw arh::Error(ARH_EINVAL, "cluster tree: a split left one side empty (coincident element centroids)");
            int l = make(s, w);
            int r = make(w, e);
            nodes[id].left = l;
            nodes[id].right = r;
            stack.push_back(r);
            stack.push_back(l);
        }
    }

    bool is_leaf(int id) const { return nodes[id].left < 0; }

    void traverse(double factor, std::vector<int> &dir, std::vector<int> &app) const {
        std::vector<std::pair<int, int>> stack{{0, 0}};
        while (!stack.empty()) {
            auto [s, t] = stack.back();
            stack.pop_back();
            const PNode &S = nodes[s], &T = nodes[t];
            double dx = S.cx - T.cx, dy = S.cy - T.cy, dz = S.cz - T.cz;
            double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dist > factor * (S.rad + T.rad)) {
                app.insert(app.end(), {S.s, S.e, T.s, T.e});
            } else if (is_leaf(s) && is_leaf(t)) {
                dir.insert(dir.end(), {S.s, S.e, T.s, T.e});
            } else {
                bool split_src = is_leaf(t) || ((S.e - S.s) > (T.e - T.s));
                if (split_src) {
                    stack.push_back({S.right, t});
                    stack.push_back({S.left, t});
                } else {
                    stack.push_back({s, T.right});
                    stack.push_back({s, T.left});
                }
            }
        }
    }
};

}  // namespace

namespace arh {
void partition(const double *C, int N, int leaf, double factor, std::vector<int> &perm, std::vector<int> &dir,
               std::vector<int> &app) {
    ARH_CHECK(N > 0 && C, ARH_EINVAL, "partition: empty mesh");
    ARH_CHECK(leaf >= 1, ARH_EINVAL, "partition: leaf size must be >= 1");
    for (size_t i = 0; i < 3 * (size_t)N; i++)
        ARH_CHECK(std::isfinite(C[i]), ARH_EINVAL, "partition: non-finite centroid");
    Builder b;
    b.C = C;
    b.leaf = leaf;
    b.build(N);
    perm = b.order;
    dir.clear();
    app.clear();
    b.traverse(factor, dir, app);
}
}  // namespace arh
