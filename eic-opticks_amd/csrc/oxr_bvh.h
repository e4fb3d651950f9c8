// oxr_bvh.h -- host-side binned-SAH BVH builder producing the 64-byte two-child node
// layout traversed by oxr_trace.h. Replaces the OptiX GAS/IAS builds
// (CSGOptiX/GAS_Builder.cc:112-159, CSGOptiX/SBT.cc:393-553).
#pragma once
#include <vector>
#include <cstdint>

namespace oxh {

struct Box {
    double lo[3], hi[3];
    static Box empty();
    void grow(const Box& b);
    void grow_point(const double* p);
    double area() const;
    double extent(int axis) const { return hi[axis] - lo[axis]; }
    double centroid(int axis) const { return 0.5 * (lo[axis] + hi[axis]); }
};

// Mirrors oxd::BNode: child0 box, child1 box, refs (>=0 node, <0 ~item).
struct NodeOut {
    float a[4], b[4], c[4];
    int32_t ch[4];
};

// Radial bound of one node: a centre c and, for each child, the range [r_in, r_out] of
// distances from c at which the surfaces of the child's items can lie. A ray from o can
// reach those surfaces no nearer than max(r_in - |o - c|, |o - c| - r_out) along it.
struct RadOut {
    float c[4];   // centre, w unused
    float r[4];   // r_in, r_out of child 0, r_in, r_out of child 1
};

// What an item's surface is known to be: a sphere (centre, radius) or anything inside box.
struct RadItem {
    double c[3];
    double r;
    bool sphere;
};

struct Bvh {
    std::vector<NodeOut> nodes;   // BFS order, root at 0 (empty if the root is a leaf)
    std::vector<RadOut> rad;      // one per node when radial items were given
    int32_t root_ref = 0;         // 0 for an internal root; ~item if the tree is one leaf
    int depth = 0;                // internal nodes on the longest root-to-leaf path
};

// Build over items with the given boxes. node_base is added to every internal ref so
// several trees can share one concatenated array. Stored child boxes are widened by a
// relative and absolute margin plus slack (mm), the bound on the rounding of the
// device's slab test for ray origins anywhere in the scene.
// With rad (one RadItem per box) each node also gets its radial bound, widened by the
// same margins. Items whose box centroids coincide (concentric shells) are split by size,
// so nested surfaces group by radius.
Bvh build_bvh(const std::vector<Box>& boxes, int node_base, double slack, const std::vector<RadItem>* rad = nullptr);

} // namespace oxh
