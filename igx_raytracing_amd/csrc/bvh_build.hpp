// Host BVH builders: binned SAH (binary) and its collapse into quantised
// 4-wide / 8-wide nodes.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "rt/rt_c.h"
#include "rt_format.h"

namespace rtx {

struct HostBvh {
    uint32_t width = 4;
    uint32_t num_tris = 0;
    uint32_t root_ref = kRefNone;
    uint32_t num_nodes = 0;
    uint32_t node_bytes = 64;
    uint32_t max_depth = 0;
    uint32_t max_stack = 0;
    uint32_t num_leaves = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double build_ms = 0.0;
    std::vector<uint8_t> nodes;  // num_nodes * node_bytes
    std::vector<TriRec> tris;    // leaf order
};

// A binary tree in the builder's own representation (used by both the SAH
// builder and the LBVH host path).  Leaves hold [first, first+count) of the
// permuted primitive order `order`.
struct BinTree {
    struct Node {
        float lo[3], hi[3];
        uint32_t left = 0, right = 0;  // valid for interior nodes
        uint32_t first = 0, count = 0; // count > 0 marks a leaf
    };
    std::vector<Node> nodes;  // nodes[0] is the root
    std::vector<uint32_t> order;
    // With reference splitting: the box of the part of triangle order[i]
    // that leaf position i covers (lo xyz, hi xyz), else empty.  A triangle
    // may then appear in several leaves.
    std::vector<float> piece;
};

// Effective options (defaults filled in, values validated).
bool resolve_opts(const rt_build_opts *in, rt_build_opts &out, std::string &err);

// Binned SAH over n input-order triangles tri_xyz[n][9].
bool build_sah_tree(const float *tri_xyz, uint32_t n, const rt_build_opts &o, BinTree &tree,
                    std::string &err);

// Pads leaf boxes outward by pad_rel * (scene scale) and refits interior
// boxes as unions of their children.
void pad_and_refit(BinTree &tree, const float *tri_xyz, float pad_rel);

// Emits the device layout (nodes of width o.node_width + leaf-ordered
// triangle records) from a binary tree.
bool emit_layout(const BinTree &tree, const float *tri_xyz, uint32_t n, const rt_build_opts &o,
                 HostBvh &out, std::string &err);

// Full host build: SAH tree, padding, layout.
bool build_host(const float *tri_xyz, uint32_t n, const rt_build_opts &o, HostBvh &out,
                std::string &err);

}  // namespace rtx
