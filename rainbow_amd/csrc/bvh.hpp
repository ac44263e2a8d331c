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

// Four-wide BVH, one 128-byte cache line per node (full-precision child boxes, SoA by
// plane, so a visit is seven aligned 16-byte loads of one line):
//   lox, hix, loy, hiy, loz, hiz   the four child boxes
//   child                          child refs in the bvh2 encoding (inner node index,
//                                  leaf bit | count - 1 | first slot, or k_empty_child)
// Leaves are the binary tree's leaves and use its leaf slots, so the two trees share
// one copy of the primitives.
struct bvh4_node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(bvh4_node) == 128, "bvh4 node is one 128-byte line");

// Eight-wide BVH with 8-bit quantised child boxes, 80 bytes per node (five
// 16-byte loads), after Ylitie, Karras and Laine, "Efficient Incoherent Ray
// Traversal on GPUs Through Compressed Wide BVHs", HPG 2017; the field coding
// below is this build's own:
//   p          node box minimum (float3)
//   e          per axis the biased exponent of the quantisation step s = 2^(e-127)
//   imask      bit i: child slot i is an inner node
//   child_base inner children are nodes child_base + (number of inner slots below i)
//   prim_base  leaf children hold primitive slots prim_base + offset ...
//   meta[i]    leaf child: 0x80 | (count - 1) << 5 | offset  (count 1..4, offset 0..31)
//              inner child or empty slot: 0
//   qlo/qhi    child box = p + q * s per axis (rounded once), conservative
// Children sit in slots by the octant of their centre relative to the node
// centre, so visiting slots in the order j ^ octant(ray direction) is roughly
// front to back.
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

constexpr uint8_t k_meta_leaf = 0x80;

// child box of slot i as the traversal sees it
inline void bvh8_child_box(const bvh8_node& n, int i, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; a++) {
        const uint32_t bits = uint32_t(n.e[a]) << 23;
        float s;
        __builtin_memcpy(&s, &bits, 4);
        lo[a] = n.p[a] + float(n.qlo[a][i]) * s;
        hi[a] = n.p[a] + float(n.qhi[a][i]) * s;
    }
}

struct bvh8_result {
    std::vector<bvh8_node> nodes;   // root at index 0
    std::vector<uint32_t> order;    // leaf slot -> primitive index
    uint32_t depth = 0;             // levels of nodes
};

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

struct bvh4_result {
    std::vector<bvh4_node> nodes;   // root at index 0, depth-first order
    uint32_t depth = 0;             // levels of nodes
    uint32_t stack = 0;             // most entries a walk's stack can hold (3 per level at most)
};

// Collapses the binary SAH tree into the four-wide layout: each node takes a binary
// node's children and opens its largest inner child until it has four.  Binary leaves
// stay leaves (same slots as the binary tree's).
void build_bvh4(const bvh_build_result& bin, bvh4_result& out);

// Collapses the binary SAH tree into the eight-wide layout: each wide node
// takes a binary node's children and keeps opening its largest inner child
// until it has eight.  Binary leaves (1..4 primitives) stay leaves.
void build_bvh8(const bvh_build_result& bin, bvh8_result& out);

}  // namespace rbw
