// Binned SAH BVH builder (host).  32 bins per axis, leaves of up to
// `max_leaf` primitives, node traversal cost 1 against primitive cost 1.
#include "bvh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
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

namespace {

// exponent e with 2^e >= x (x > 0)
inline int ceil_log2(double x) {
    int k;
    const double m = std::frexp(x, &k);  // x = m 2^k, m in [0.5, 1)
    return m == 0.5 ? k - 1 : k;
}

inline float step_of(uint8_t e) {
    const uint32_t bits = uint32_t(e) << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    return s;
}

}  // namespace

void build_bvh4(const bvh_build_result& bin, bvh4_result& out) {
    const std::vector<build_node>& t = bin.tree;
    out = bvh4_result();
    if (t.empty()) throw std::runtime_error("empty tree");
    out.nodes.reserve(t.size() / 3 + 1);
    auto leaf_ref = [&](int32_t b) {
        return int32_t(uint32_t(k_leaf_bit) | ((t[b].count - 1) << 27) | t[b].first);
    };
    // depth-first: a node's subtree follows it; pushed = stack entries its ancestors may leave
    struct maker {
        const std::vector<build_node>& t;
        bvh4_result& out;
        const decltype(leaf_ref)& leaf;
        uint32_t make(int32_t b, uint32_t level, uint32_t pushed) {
            out.depth = std::max(out.depth, level);
            const uint32_t id = uint32_t(out.nodes.size());
            out.nodes.emplace_back();
            int32_t ch[4];
            int nch = 0;
            if (t[b].left < 0) {
                ch[nch++] = b;
            } else {
                ch[nch++] = t[b].left;
                ch[nch++] = t[b].right;
                while (nch < 4) {
                    int best = -1;
                    float best_area = -1.0f;
                    for (int k = 0; k < nch; k++)
                        if (t[ch[k]].left >= 0 && area(t[ch[k]].box) > best_area) {
                            best_area = area(t[ch[k]].box);
                            best = k;
                        }
                    if (best < 0) break;
                    const int32_t o = ch[best];
                    ch[best] = t[o].left;
                    ch[nch++] = t[o].right;
                }
            }
            const uint32_t below = pushed + uint32_t(nch - 1);
            out.stack = std::max(out.stack, below);
            bvh4_node n;
            std::memset(&n, 0, sizeof(n));
            for (int k = 0; k < 4; k++) {
                if (k >= nch) {
                    n.lox[k] = n.loy[k] = n.loz[k] = FLT_MAX;
                    n.hix[k] = n.hiy[k] = n.hiz[k] = -FLT_MAX;
                    n.child[k] = k_empty_child;
                    continue;
                }
                const aabb& bx = t[ch[k]].box;
                n.lox[k] = bx.lo[0]; n.hix[k] = bx.hi[0];
                n.loy[k] = bx.lo[1]; n.hiy[k] = bx.hi[1];
                n.loz[k] = bx.lo[2]; n.hiz[k] = bx.hi[2];
                n.child[k] = t[ch[k]].left >= 0 ? int32_t(make(ch[k], level + 1, below)) : leaf(ch[k]);
            }
            out.nodes[id] = n;
            return id;
        }
    };
    maker m{t, out, leaf_ref};
    m.make(0, 1, 0);
}

void build_bvh8(const bvh_build_result& bin, bvh8_result& out) {
    const std::vector<build_node>& t = bin.tree;
    out = bvh8_result();
    if (t.empty()) throw std::runtime_error("empty tree");
    // primitive range of every binary subtree (the builder partitions in place, so a
    // subtree's primitives are consecutive slots); subtrees of at most k_leaf8 primitives
    // become one wide-tree leaf
    constexpr uint32_t k_leaf8 = 2;
    std::vector<uint32_t> first(t.size()), count(t.size());
    for (size_t k = t.size(); k-- > 0;) {
        if (t[k].left < 0) {
            first[k] = t[k].first;
            count[k] = t[k].count;
        } else {
            first[k] = std::min(first[size_t(t[k].left)], first[size_t(t[k].right)]);
            count[k] = count[size_t(t[k].left)] + count[size_t(t[k].right)];
        }
    }
    auto is_leaf = [&](int32_t b) { return t[b].left < 0 || count[size_t(b)] <= k_leaf8; };
    out.order.reserve(bin.order.size());
    struct item {
        int32_t bnode;
        uint32_t node8, level;
    };
    std::vector<item> queue;
    queue.push_back({0, 0, 1});
    out.nodes.push_back(bvh8_node{});
    for (size_t qi = 0; qi < queue.size(); qi++) {
        const item it = queue[qi];
        out.depth = std::max(out.depth, it.level);
        // children: open the largest inner child until there are eight
        std::vector<int32_t> ch;
        if (is_leaf(it.bnode)) {
            ch.push_back(it.bnode);
        } else {
            ch.push_back(t[it.bnode].left);
            ch.push_back(t[it.bnode].right);
            while (ch.size() < 8) {
                // the largest subtree too big for a leaf; when none is left, split the
                // largest small subtree into finer leaves while slots are free
                int best = -1;
                float best_area = -1.0f;
                for (int pass = 0; pass < 2 && best < 0; pass++)
                    for (size_t k = 0; k < ch.size(); k++) {
                        const bool open = pass == 0 ? !is_leaf(ch[k]) : t[ch[k]].left >= 0;
                        if (open && area(t[ch[k]].box) > best_area) {
                            best_area = area(t[ch[k]].box);
                            best = int(k);
                        }
                    }
                if (best < 0) break;
                const int32_t b = ch[size_t(best)];
                ch[size_t(best)] = t[b].left;
                ch.push_back(t[b].right);
            }
        }
        aabb nb = empty_box();
        for (int32_t c : ch) grow(nb, t[c].box);
        // slots by octant of the child centre; larger children choose first
        std::vector<int32_t> order_by_area(ch.begin(), ch.end());
        std::stable_sort(order_by_area.begin(), order_by_area.end(),
                         [&](int32_t a, int32_t b) { return area(t[a].box) > area(t[b].box); });
        int32_t child_in_slot[8];
        for (int i = 0; i < 8; i++) child_in_slot[i] = -1;
        for (int32_t c : order_by_area) {
            int oct = 0;
            for (int a = 0; a < 3; a++) {
                const float cc = 0.5f * (t[c].box.lo[a] + t[c].box.hi[a]);
                const float nc = 0.5f * (nb.lo[a] + nb.hi[a]);
                if (cc > nc) oct |= 4 >> a;
            }
            int pick = -1, pick_d = 99;
            for (int s = 0; s < 8; s++) {
                if (child_in_slot[s] >= 0) continue;
                const int d = __builtin_popcount(unsigned(s ^ oct));
                if (d < pick_d) {
                    pick_d = d;
                    pick = s;
                }
            }
            child_in_slot[pick] = c;
        }
        bvh8_node n;
        std::memset(&n, 0, sizeof(n));
        // quantisation grid: p = node minimum, step 2^e with 255 steps covering the box
        for (int a = 0; a < 3; a++) {
            n.p[a] = nb.lo[a];
            const double range = double(nb.hi[a]) - double(nb.lo[a]);
            int e = range > 0.0 ? ceil_log2(range * (1.0 + 1.0 / (1 << 18)) / 255.0) : -126;
            e = std::max(-126, std::min(127, e));
            n.e[a] = uint8_t(e + 127);
        }
        // inner children first (consecutive node indices in slot order), then leaves
        n.child_base = uint32_t(out.nodes.size());
        n.prim_base = uint32_t(out.order.size());
        for (int s = 0; s < 8; s++) {
            const int32_t c = child_in_slot[s];
            if (c < 0 || is_leaf(c)) continue;
            n.imask |= uint8_t(1u << s);
            queue.push_back({c, uint32_t(out.nodes.size()), it.level + 1});
            out.nodes.push_back(bvh8_node{});
        }
        for (int s = 0; s < 8; s++) {
            const int32_t c = child_in_slot[s];
            if (c < 0) continue;
            if (is_leaf(c)) {
                const uint32_t off = uint32_t(out.order.size()) - n.prim_base;
                const uint32_t cnt = count[size_t(c)];
                if (cnt < 1 || cnt > 4 || off > 31) throw std::runtime_error("bvh8 leaf does not fit its meta byte");
                n.meta[s] = uint8_t(k_meta_leaf | ((cnt - 1) << 5) | off);
                for (uint32_t k = 0; k < cnt; k++) out.order.push_back(bin.order[first[size_t(c)] + k]);
            }
            for (int a = 0; a < 3; a++) {
                const float st = step_of(n.e[a]);
                double ql = std::floor((double(t[c].box.lo[a]) - double(n.p[a])) / double(st));
                double qh = std::ceil((double(t[c].box.hi[a]) - double(n.p[a])) / double(st));
                int lo = int(std::max(0.0, std::min(255.0, ql)));
                int hi = int(std::max(0.0, std::min(255.0, qh)));
                // conservative in the traversal's float arithmetic: p + q s rounded once
                while (lo > 0 && n.p[a] + float(lo) * st > t[c].box.lo[a]) lo--;
                while (hi < 255 && n.p[a] + float(hi) * st < t[c].box.hi[a]) hi++;
                if (n.p[a] + float(lo) * st > t[c].box.lo[a] || n.p[a] + float(hi) * st < t[c].box.hi[a])
                    throw std::runtime_error("bvh8 quantisation is not conservative");
                n.qlo[a][s] = uint8_t(lo);
                n.qhi[a][s] = uint8_t(hi);
            }
        }
        out.nodes[it.node8] = n;
    }
}

}  // namespace rbw
