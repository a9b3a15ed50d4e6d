// Device data layout of a scene: child references, BVH2/BVH4/BVH8 nodes and
// the leaf-ordered triangle records.  Shared by the host builders and the
// HIP kernels; plain structs with fixed sizes.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RT_HD __host__ __device__
#else
#define RT_HD
#endif

namespace rtx {

// Child reference (32 bits):
//   0xFFFFFFFF            empty slot
//   bit 31 = 0            interior node, bits 0..30 = node index
//   bit 31 = 1            leaf: bits 28..30 = count-1 (1..8 triangles),
//                         bits 0..27 = first triangle record
constexpr uint32_t kRefNone = 0xFFFFFFFFu;
constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kLeafOffMask = 0x0FFFFFFFu;
constexpr uint32_t kMaxLeafOffset = 0x0FFFFFF0u;
constexpr uint32_t kMaxLeafTris = 8;

RT_HD inline uint32_t make_leaf(uint32_t first, uint32_t count) {
    return kRefLeaf | ((count - 1u) << 28) | first;
}
RT_HD inline bool ref_is_leaf(uint32_t r) { return (r & kRefLeaf) != 0u; }
RT_HD inline uint32_t leaf_first(uint32_t r) { return r & kLeafOffMask; }
RT_HD inline uint32_t leaf_count(uint32_t r) { return ((r >> 28) & 7u) + 1u; }

// Binary node, 64 bytes: both children's float boxes and references.
struct alignas(64) NodeB2 {
    float lo0[3], hi0[3];
    float lo1[3], hi1[3];
    uint32_t child[2];
    uint32_t pad[2];
};
static_assert(sizeof(NodeB2) == 64, "NodeB2 must be 64 bytes");

// 4-wide node, 64 bytes.  Child c's box on axis k is
//   [origin[k] + qlo[k][c] * 2^(ex[k]-127), origin[k] + qhi[k][c] * 2^(ex[k]-127)]
// with 8-bit q rounded outward; an empty slot has qlo = 255, qhi = 0 and
// child = kRefNone, so its box is empty for every ray.
struct alignas(64) NodeB4 {
    float origin[3];
    uint8_t ex[3];      // biased exponents of the per-axis scale
    uint8_t nchild;
    uint8_t qlo[3][4];  // [axis][child]
    uint8_t qhi[3][4];
    uint32_t child[4];
    uint32_t pad[2];
};
static_assert(sizeof(NodeB4) == 64, "NodeB4 must be 64 bytes");

// 8-wide node, 128 bytes (two 64-byte halves), same encoding as NodeB4.
struct alignas(128) NodeB8 {
    float origin[3];
    uint8_t ex[3];
    uint8_t nchild;
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
    uint32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(NodeB8) == 128, "NodeB8 must be 128 bytes");

// Triangle record, 48 bytes, in leaf order:
//   [p0.x p0.y p0.z id] [e1.x e1.y e1.z 0] [e2.x e2.y e2.z 0]
// with e1 = p1 - p0 and e2 = p2 - p0 rounded to FP32 exactly as
// primitive.glsl:245-246 computes them, so storing them changes no bit.
struct alignas(16) TriRec {
    float p0[3];
    uint32_t id;
    float e1[3];
    uint32_t pad1;
    float e2[3];
    uint32_t pad2;
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

}  // namespace rtx
