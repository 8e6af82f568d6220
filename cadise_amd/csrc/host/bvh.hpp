// Host BVH build. Two stages:
//   1. build_binary: the reference's binary BVH algorithm (BvhBuilder,
//      reference core/intersector/accelerator/bvh/bvhBuilder.cpp:18-358):
//      12-bucket SAH (or equal split), leaf size 1, DFS pre-order layout, and
//      the ordered primitive list. With reference_quirk the empty box keeps
//      the reference's max = FLT_MIN, which reproduces its tree node for node;
//      without it the same algorithm runs on true empty boxes (the tree the
//      GPU traverses: the quirk stretches every box of negative-coordinate
//      geometry to the origin and leaves tens of primitives per leaf).
//   2. collapse_wide: the binary tree folded into a W-wide tree (W = 4 or 8)
//      with tight child boxes and small subtrees folded into multi-primitive
//      leaves, laid out for the gfx950 traversal kernel.
#pragma once

#include <cstdint>
#include <vector>

#include "scene.hpp"

namespace mpt {

struct BinaryNode {
    float   bmin[3], bmax[3];
    int64_t a;      // leaf: first ordered primitive; interior: second child index
    int64_t b;      // leaf: primitive count; interior: split axis
    int32_t leaf;
};

struct BinaryBvh {
    std::vector<BinaryNode> nodes;
    std::vector<int32_t>    ordered;   // creation indices in leaf order
    int32_t depth = 0;
};

void build_binary(const HostScene& s, bool reference_quirk, BinaryBvh* out);

// Root bound of the reference's tree: the union of every primitive's padded
// bound accumulated on the reference's empty box (max starts at FLT_MIN).
void reference_root_bound(const HostScene& s, float lo[3], float hi[3]);

// Wide node, structure-of-arrays over its W children. child[i] >= 0 with
// count[i] == 0: interior child node index. count[i] > 0: leaf of count[i]
// primitives starting at ordered position child[i]. Empty slot: count 0,
// child -1, inverted box.
template <int W>
struct alignas(16) WideNode {
    float lo_x[W], lo_y[W], lo_z[W];
    float hi_x[W], hi_y[W], hi_z[W];
    int32_t child[W];
    int32_t count[W];
};
static_assert(sizeof(WideNode<4>) == 128, "BVH4 node must be one 128-byte line");
static_assert(sizeof(WideNode<8>) == 256, "BVH8 node must be two 128-byte lines");

struct WideBvh {
    int32_t width = 4;
    std::vector<uint8_t> nodes;   // WideNode<width> array, root at 0
    int64_t nnodes = 0;
    int32_t depth = 0;
    int32_t max_stack = 0;        // upper bound on traversal stack entries
};

void collapse_wide(const HostScene& s, const BinaryBvh& bin, int width, int leaf_max, WideBvh* out);

}  // namespace mpt
