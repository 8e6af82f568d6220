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

// The device's node: child boxes as 8-bit offsets on a per-node grid with
// origin o and power-of-two steps 2^(e - 127) per axis (exps: the biased
// exponents in bytes 0..2). A child's box is [o + q_lo 2^e, o + q_hi 2^e], each
// corner evaluated as fma(q, 2^e, o) in binary32; the host picks q so that this
// box contains the child's tight box. child: leaf = first primitive (leaf
// order), interior = node index, -1 = empty slot; count: primitives of a leaf,
// 0 for an interior child. BVH4: 64 bytes (two nodes per 128-byte line), BVH8:
// 128 bytes (one line).
template <int W>
struct alignas(16) QNode {
    float    origin[3];
    uint32_t exps;
    int32_t  child[W];
    uint8_t  q[6][W];            // lo_x, lo_y, lo_z, hi_x, hi_y, hi_z
    uint16_t count[W];
    uint8_t  pad[W == 8 ? 16 : 0];
};
static_assert(sizeof(QNode<4>) == 64, "BVH4 node must be half a 128-byte line");
static_assert(sizeof(QNode<8>) == 128, "BVH8 node must be one 128-byte line");

struct WideBvh {
    int32_t width = 4;
    std::vector<uint8_t> nodes;   // QNode<width> array, root at 0
    int64_t nnodes = 0;
    int32_t depth = 0;
    int32_t max_stack = 0;        // upper bound on traversal stack entries
    bool    ok = true;            // false: a node could not be encoded (a leaf of more than
                                  // 65535 primitives, or boxes beyond the binary32 range)
};

void collapse_wide(const HostScene& s, const BinaryBvh& bin, int width, int leaf_max, WideBvh* out);

}  // namespace mpt
