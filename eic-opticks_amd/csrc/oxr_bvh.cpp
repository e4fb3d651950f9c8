// oxr_bvh.cpp -- binned SAH build (16 bins per axis, one item per leaf) with BFS output.
#include "oxr_bvh.h"
#include <algorithm>
#include <cmath>
#include <limits>
#include <deque>

namespace oxh {

Box Box::empty() {
    Box b;
    for (int k = 0; k < 3; k++) { b.lo[k] = std::numeric_limits<double>::infinity(); b.hi[k] = -std::numeric_limits<double>::infinity(); }
    return b;
}
void Box::grow(const Box& o) {
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], o.lo[k]); hi[k] = std::max(hi[k], o.hi[k]); }
}
void Box::grow_point(const double* p) {
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
}
double Box::area() const {
    double dx = std::max(0.0, hi[0] - lo[0]), dy = std::max(0.0, hi[1] - lo[1]), dz = std::max(0.0, hi[2] - lo[2]);
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

namespace {

struct TmpNode { Box box; int left = -1, right = -1, item = -1; };

struct Builder {
    const std::vector<Box>& boxes;
    std::vector<int> idx;
    std::vector<TmpNode> tmp;
    explicit Builder(const std::vector<Box>& b) : boxes(b) {}

    int build(int begin, int end, int depth) {
        int me = int(tmp.size());
        tmp.emplace_back();
        Box bb = Box::empty(), cb = Box::empty();
        for (int i = begin; i < end; i++) {
            bb.grow(boxes[idx[i]]);
            double c[3] = {boxes[idx[i]].centroid(0), boxes[idx[i]].centroid(1), boxes[idx[i]].centroid(2)};
            cb.grow_point(c);
        }
        tmp[me].box = bb;
        int n = end - begin;
        if (n == 1) { tmp[me].item = idx[begin]; return me; }
        // binned SAH over the centroid box
        const int NB = 16;
        int best_axis = -1, best_split = -1;
        double best_cost = std::numeric_limits<double>::infinity();
        for (int axis = 0; axis < 3; axis++) {
            double ext = cb.extent(axis);
            if (!(ext > 0.0)) continue;
            Box bin[NB]; int cnt[NB];
            for (int k = 0; k < NB; k++) { bin[k] = Box::empty(); cnt[k] = 0; }
            for (int i = begin; i < end; i++) {
                int k = int(NB * (boxes[idx[i]].centroid(axis) - cb.lo[axis]) / ext);
                k = std::min(NB - 1, std::max(0, k));
                bin[k].grow(boxes[idx[i]]); cnt[k]++;
            }
            double la[NB], ra[NB]; int lc[NB], rc[NB];
            Box acc = Box::empty(); int c = 0;
            for (int k = 0; k < NB; k++) { acc.grow(bin[k]); c += cnt[k]; la[k] = acc.area(); lc[k] = c; }
            acc = Box::empty(); c = 0;
            for (int k = NB - 1; k >= 0; k--) { acc.grow(bin[k]); c += cnt[k]; ra[k] = acc.area(); rc[k] = c; }
            for (int s = 0; s < NB - 1; s++) {
                if (lc[s] == 0 || rc[s + 1] == 0) continue;
                double cost = la[s] * lc[s] + ra[s + 1] * rc[s + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = s; }
            }
        }
        int mid;
        if (best_axis < 0) {
            // coincident centroids (concentric items): median by box size, so nested
            // surfaces group by radius
            mid = begin + n / 2;
            std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end,
                             [&](int p, int q) { return boxes[p].area() < boxes[q].area(); });
        } else if (depth >= 24) {
            // a deep branch: spatial median on the widest axis
            int ax = 0;
            for (int k = 1; k < 3; k++) if (cb.extent(k) > cb.extent(ax)) ax = k;
            mid = begin + n / 2;
            std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end,
                             [&](int p, int q) { return boxes[p].centroid(ax) < boxes[q].centroid(ax); });
        } else {
            double ext = cb.extent(best_axis);
            auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](int q) {
                int k = int(NB * (boxes[q].centroid(best_axis) - cb.lo[best_axis]) / ext);
                k = std::min(NB - 1, std::max(0, k));
                return k <= best_split;
            });
            mid = int(it - idx.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        int l = build(begin, mid, depth + 1);
        int r = build(mid, end, depth + 1);
        tmp[me].left = l; tmp[me].right = r;
        return me;
    }
};

// expand outward so float slab tests stay conservative for the CSG hits inside
void put_box(const Box& b, float* out6, double slack) {
    double m = 0.0;
    for (int k = 0; k < 3; k++) m = std::max(m, std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k])));
    double eps = 1e-5 * m + 1e-3 + slack;
    for (int k = 0; k < 3; k++) {
        out6[k] = std::nextafter(float(b.lo[k] - eps), -std::numeric_limits<float>::infinity());
        out6[3 + k] = std::nextafter(float(b.hi[k] + eps), std::numeric_limits<float>::infinity());
    }
}

// Radial bound of a subtree about its own centre, merged bottom-up: a sphere item is
// (its centre, r, r), any other item (box centre, 0, half diagonal); a parent takes its
// box centre and the children's ranges shifted by their centres' distance from it (the
// triangle inequality keeps the ranges valid, and exact for shared centres).
struct Rad { double c[3]; double rin, rout; };

double dist3(const double* a, const double* b) {
    double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return std::sqrt(x * x + y * y + z * z);
}

Rad item_rad(const Box& b, const RadItem& it) {
    Rad r;
    if (it.sphere) {
        for (int k = 0; k < 3; k++) r.c[k] = it.c[k];
        r.rin = r.rout = it.r;
    } else {
        for (int k = 0; k < 3; k++) r.c[k] = b.centroid(k);
        r.rin = 0.0;
        r.rout = 0.5 * std::sqrt(b.extent(0) * b.extent(0) + b.extent(1) * b.extent(1) + b.extent(2) * b.extent(2));
    }
    return r;
}

// the child's range about the parent's centre c
void shifted(const Rad& ch, const double* c, double& rin, double& rout) {
    double d = dist3(ch.c, c);
    rin = std::max(0.0, ch.rin - d);
    rout = ch.rout + d;
}

} // namespace

Bvh build_bvh(const std::vector<Box>& boxes, int node_base, double slack, const std::vector<RadItem>* rad) {
    Bvh out;
    int n = int(boxes.size());
    if (n == 0) { out.root_ref = int32_t(0x80000000u); return out; }
    Builder B(boxes);
    B.idx.resize(n);
    for (int i = 0; i < n; i++) B.idx[i] = i;
    int root = B.build(0, n, 0);
    if (B.tmp[root].item >= 0) { out.root_ref = ~B.tmp[root].item; return out; }
    // BFS numbering of internal nodes
    std::vector<int> order;
    std::vector<int> bfs_index(B.tmp.size(), -1);
    std::vector<int> level(B.tmp.size(), 0);
    std::deque<int> q;
    q.push_back(root);
    while (!q.empty()) {
        int t = q.front(); q.pop_front();
        bfs_index[t] = int(order.size());
        order.push_back(t);
        out.depth = std::max(out.depth, level[t] + 1);
        for (int c : {B.tmp[t].left, B.tmp[t].right}) {
            if (B.tmp[c].item < 0) { level[c] = level[t] + 1; q.push_back(c); }
        }
    }
    out.nodes.resize(order.size());
    for (size_t k = 0; k < order.size(); k++) {
        const TmpNode& t = B.tmp[order[k]];
        float b0[6], b1[6];
        put_box(B.tmp[t.left].box, b0, slack);
        put_box(B.tmp[t.right].box, b1, slack);
        NodeOut& o = out.nodes[k];
        o.a[0] = b0[0]; o.a[1] = b0[1]; o.a[2] = b0[2]; o.a[3] = b0[3];
        o.b[0] = b0[4]; o.b[1] = b0[5]; o.b[2] = b1[0]; o.b[3] = b1[1];
        o.c[0] = b1[2]; o.c[1] = b1[3]; o.c[2] = b1[4]; o.c[3] = b1[5];
        auto ref = [&](int c) { return B.tmp[c].item >= 0 ? int32_t(~B.tmp[c].item) : int32_t(node_base + bfs_index[c]); };
        o.ch[0] = ref(t.left); o.ch[1] = ref(t.right); o.ch[2] = 0; o.ch[3] = 0;
    }
    if (rad) {
        // bottom-up over the temporary tree (children are created after their parent)
        std::vector<Rad> R(B.tmp.size());
        for (int t = int(B.tmp.size()) - 1; t >= 0; t--) {
            const TmpNode& tn = B.tmp[size_t(t)];
            if (tn.item >= 0) { R[size_t(t)] = item_rad(boxes[size_t(tn.item)], (*rad)[size_t(tn.item)]); continue; }
            Rad& r = R[size_t(t)];
            for (int k = 0; k < 3; k++) r.c[k] = tn.box.centroid(k);
            double i0, o0, i1, o1;
            shifted(R[size_t(tn.left)], r.c, i0, o0);
            shifted(R[size_t(tn.right)], r.c, i1, o1);
            r.rin = std::min(i0, i1);
            r.rout = std::max(o0, o1);
        }
        out.rad.resize(order.size());
        for (size_t k = 0; k < order.size(); k++) {
            const TmpNode& t = B.tmp[size_t(order[k])];
            const Rad& r = R[size_t(order[k])];
            RadOut& o = out.rad[k];
            double scale = 0.0;
            for (int j = 0; j < 3; j++) { o.c[j] = float(r.c[j]); scale = std::max(scale, std::fabs(r.c[j])); }
            o.c[3] = 0.f;
            double ci[3] = {double(o.c[0]), double(o.c[1]), double(o.c[2])};   // the centre the device uses
            int j = 0;
            for (int ch : {t.left, t.right}) {
                double rin, rout;
                shifted(R[size_t(ch)], ci, rin, rout);
                // widened like the boxes: the float distance |o - c| and the leaves' float roots
                // stay inside these bounds for any ray origin in the scene
                double m = 1e-5 * (scale + rout) + 0.05 + slack;
                o.r[2 * j] = std::nextafter(float(std::max(0.0, rin - m)), -std::numeric_limits<float>::infinity());
                if (o.r[2 * j] < 0.f) o.r[2 * j] = 0.f;
                o.r[2 * j + 1] = std::nextafter(float(rout + m), std::numeric_limits<float>::infinity());
                j++;
            }
        }
    }
    out.root_ref = node_base;
    return out;
}

} // namespace oxh
