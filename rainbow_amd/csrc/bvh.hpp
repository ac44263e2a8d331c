// Host-side SAH BVH builder for the tracer (replaces the reference's median
// split BVH, shared/accelerators/detail/bounding_volume_hierarchy.hpp:162-233,
// and the two-level split of scenes/scene.cpp:30-80: every primitive of every
// entity goes into one tree).
#pragma once

#include <cstdint>
#include <vector>

namespace rbw {

struct aabb {
    float lo[3], hi[3];
};

// Binary BVH in the "two child boxes per node" layout (64 bytes per node):
//   n[0] = (left.lo.x, left.hi.x, left.lo.y, left.hi.y)
//   n[1] = (right.lo.x, right.hi.x, right.lo.y, right.hi.y)
//   n[2] = (left.lo.z, left.hi.z, right.lo.z, right.hi.z)
//   n[3] = (left child, right child, 0, 0) as int bits
// A child >= 0 is an inner node index; a child with bit 31 set is a leaf:
// bits 27..30 = primitive count - 1, bits 0..26 = first primitive slot.
struct bvh2_node {
    float lxy[4];
    float rxy[4];
    float lrz[4];
    int32_t child[4];
};
static_assert(sizeof(bvh2_node) == 64, "bvh2 node is 64 bytes");

constexpr int32_t k_leaf_bit = int32_t(0x80000000u);
constexpr uint32_t k_leaf_max = 16;
constexpr int32_t k_empty_child = int32_t(0xFFFFFFFFu);  // leaf flag with count 16 at slot 2^27-1: never hit

// Wide BVH with 8-bit quantised child boxes (80 bytes per node), the layout of
// Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs
// Through Compressed Wide BVHs", HPG 2017:
//   p (float3 origin), e (3 x uint8 exponent), imask (inner-child mask),
//   child_base, prim_base, meta[8], qlo_x[8], qlo_y[8], qlo_z[8], qhi_x[8], qhi_y[8], qhi_z[8]
struct bvh8_node {
    float p[3];
    uint8_t e[3];
    uint8_t imask;
    uint32_t child_base;
    uint32_t prim_base;
    uint8_t meta[8];
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
};
static_assert(sizeof(bvh8_node) == 80, "bvh8 node is 80 bytes");

// Binary SAH tree in builder form (children indices, node boxes)
struct build_node {
    aabb box;
    int32_t left = -1, right = -1;  // inner: child build nodes; leaf: left = -1
    uint32_t first = 0, count = 0;  // leaf: slots [first, first+count) of `order`
};

struct bvh_build_result {
    std::vector<build_node> tree;    // root at index 0
    std::vector<uint32_t> order;     // leaf slot -> primitive index
    std::vector<bvh2_node> nodes2;   // root at index 0
    uint32_t depth = 0;
    uint32_t leaves = 0;
};

// bounds: conservative primitive boxes.  Builds the binned-SAH tree and the bvh2 layout.
void build_bvh(const std::vector<aabb>& bounds, bvh_build_result& out, uint32_t max_leaf = 4);

// Collapses the binary tree into the compressed 8-wide layout.  Leaves hold at
// most 3 primitives in consecutive slots of `order8`.
void build_bvh8(const std::vector<aabb>& bounds, const bvh_build_result& bin, std::vector<bvh8_node>& nodes8,
                std::vector<uint32_t>& order8);

}  // namespace rbw
