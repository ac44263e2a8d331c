// Binned SAH BVH builder (host).  32 bins per axis, leaves of up to
// `max_leaf` primitives, node traversal cost 1 against primitive cost 1.
#include "bvh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>

namespace rbw {
namespace {

constexpr int k_bins = 32;

inline void grow(aabb& b, const aabb& o) {
    for (int a = 0; a < 3; a++) {
        b.lo[a] = std::min(b.lo[a], o.lo[a]);
        b.hi[a] = std::max(b.hi[a], o.hi[a]);
    }
}

inline aabb empty_box() {
    aabb b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = FLT_MAX;
        b.hi[a] = -FLT_MAX;
    }
    return b;
}

inline float area(const aabb& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

struct builder {
    const std::vector<aabb>& bounds;
    std::vector<float> cx, cy, cz;
    bvh_build_result& out;
    uint32_t max_leaf;

    builder(const std::vector<aabb>& b, bvh_build_result& o, uint32_t ml) : bounds(b), out(o), max_leaf(ml) {
        const size_t n = b.size();
        cx.resize(n); cy.resize(n); cz.resize(n);
        for (size_t i = 0; i < n; i++) {
            cx[i] = 0.5f * (b[i].lo[0] + b[i].hi[0]);
            cy[i] = 0.5f * (b[i].lo[1] + b[i].hi[1]);
            cz[i] = 0.5f * (b[i].lo[2] + b[i].hi[2]);
        }
    }
    float centroid(uint32_t p, int axis) const { return axis == 0 ? cx[p] : (axis == 1 ? cy[p] : cz[p]); }

    int32_t make_leaf(uint32_t first, uint32_t count, const aabb& box) {
        build_node n;
        n.box = box;
        n.first = first;
        n.count = count;
        out.tree.push_back(n);
        out.leaves++;
        return int32_t(out.tree.size() - 1);
    }

    int32_t build(uint32_t first, uint32_t count, uint32_t depth) {
        out.depth = std::max(out.depth, depth);
        uint32_t* ord = out.order.data() + first;
        aabb box = empty_box(), cbox = empty_box();
        for (uint32_t i = 0; i < count; i++) {
            grow(box, bounds[ord[i]]);
            aabb c;
            for (int a = 0; a < 3; a++) c.lo[a] = c.hi[a] = centroid(ord[i], a);
            grow(cbox, c);
        }
        if (count == 1) return make_leaf(first, count, box);
        // best split over the three axes
        float best_cost = FLT_MAX;
        int best_axis = -1, best_bin = -1;
        for (int axis = 0; axis < 3; axis++) {
            const float lo = cbox.lo[axis], hi = cbox.hi[axis];
            if (!(hi > lo)) continue;
            const float scale = k_bins / (hi - lo);
            aabb bb[k_bins];
            uint32_t bc[k_bins] = {};
            for (int b = 0; b < k_bins; b++) bb[b] = empty_box();
            for (uint32_t i = 0; i < count; i++) {
                int b = int((centroid(ord[i], axis) - lo) * scale);
                b = std::min(std::max(b, 0), k_bins - 1);
                bc[b]++;
                grow(bb[b], bounds[ord[i]]);
            }
            float ra[k_bins];
            uint32_t rc[k_bins];
            aabb acc = empty_box();
            uint32_t cnt = 0;
            for (int b = k_bins - 1; b > 0; b--) {
                grow(acc, bb[b]);
                cnt += bc[b];
                ra[b] = area(acc);
                rc[b] = cnt;
            }
            acc = empty_box();
            cnt = 0;
            for (int b = 0; b < k_bins - 1; b++) {
                grow(acc, bb[b]);
                cnt += bc[b];
                if (cnt == 0 || rc[b + 1] == 0) continue;
                const float cost = cnt * area(acc) + rc[b + 1] * ra[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_bin = b;
                }
            }
        }
        const float parent_area = area(box);
        const float leaf_cost = float(count);
        const float split_cost = best_axis >= 0 && parent_area > 0 ? 1.0f + best_cost / parent_area : FLT_MAX;
        uint32_t mid;
        if (best_axis < 0 || (split_cost >= leaf_cost && count <= max_leaf)) {
            if (count <= max_leaf) return make_leaf(first, count, box);
            if (best_axis < 0) {
                mid = count / 2;  // coincident centroids: split by index
            } else {
                goto partition;
            }
        } else {
        partition:
            const float lo = cbox.lo[best_axis], hi = cbox.hi[best_axis];
            const float scale = k_bins / (hi - lo);
            uint32_t* m = std::partition(ord, ord + count, [&](uint32_t p) {
                int b = int((centroid(p, best_axis) - lo) * scale);
                b = std::min(std::max(b, 0), k_bins - 1);
                return b <= best_bin;
            });
            mid = uint32_t(m - ord);
            if (mid == 0 || mid == count) mid = count / 2;
        }
        const int32_t id = int32_t(out.tree.size());
        out.tree.emplace_back();
        out.tree[id].box = box;
        const int32_t l = build(first, mid, depth + 1);
        const int32_t r = build(first + mid, count - mid, depth + 1);
        out.tree[id].left = l;
        out.tree[id].right = r;
        return id;
    }
};

inline void store_box(const aabb& b, float* xy, float* z, int which) {
    xy[0] = b.lo[0]; xy[1] = b.hi[0]; xy[2] = b.lo[1]; xy[3] = b.hi[1];
    z[2 * which + 0] = b.lo[2];
    z[2 * which + 1] = b.hi[2];
}

}  // namespace

void build_bvh(const std::vector<aabb>& bounds, bvh_build_result& out, uint32_t max_leaf) {
    if (bounds.empty()) throw std::runtime_error("empty scene");
    if (max_leaf > k_leaf_max) max_leaf = k_leaf_max;
    out = bvh_build_result();
    out.order.resize(bounds.size());
    for (uint32_t i = 0; i < bounds.size(); i++) out.order[i] = i;
    out.tree.reserve(2 * bounds.size() / std::max(1u, max_leaf / 2) + 8);
    builder b(bounds, out, max_leaf);
    b.build(0, uint32_t(bounds.size()), 1);
    // bvh2 layout: one node per inner build node; a leaf root gets a wrapper node
    const std::vector<build_node>& t = out.tree;
    std::vector<int32_t> inner_id(t.size(), -1);
    int32_t next = 0;
    if (t[0].left < 0) {
        next = 1;
    } else {
        for (size_t i = 0; i < t.size(); i++)
            if (t[i].left >= 0) inner_id[i] = next++;
    }
    out.nodes2.assign(size_t(next), bvh2_node{});
    auto child_ref = [&](int32_t c) -> int32_t {
        if (t[c].left >= 0) return inner_id[c];
        return int32_t(uint32_t(k_leaf_bit) | ((t[c].count - 1) << 27) | t[c].first);
    };
    if (t[0].left < 0) {
        bvh2_node& n = out.nodes2[0];
        store_box(t[0].box, n.lxy, n.lrz, 0);
        const aabb e = empty_box();
        store_box(e, n.rxy, n.lrz, 1);
        n.child[0] = child_ref(0);
        n.child[1] = k_empty_child;
    } else {
        for (size_t i = 0; i < t.size(); i++) {
            if (t[i].left < 0) continue;
            bvh2_node& n = out.nodes2[size_t(inner_id[i])];
            store_box(t[t[i].left].box, n.lxy, n.lrz, 0);
            store_box(t[t[i].right].box, n.rxy, n.lrz, 1);
            n.child[0] = child_ref(t[i].left);
            n.child[1] = child_ref(t[i].right);
        }
    }
    if (bounds.size() >= (1u << 27)) throw std::runtime_error("too many primitives for the bvh2 leaf encoding");
}

}  // namespace rbw
