// Structural check of the eight-wide BVH (rainbow_amd/csrc/bvh.cpp) on random
// triangle soups: every primitive lands in exactly one leaf, every quantised
// child box contains its subtree's primitive boxes in the traversal's float
// arithmetic, and a closest-hit walk over the wide tree finds the same nearest
// triangle as a brute-force scan.  Prints "ok" or the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../rainbow_amd/csrc/bvh.hpp"

using namespace rbw;

struct tri {
    float v[3][3];
};

static bool mt(const tri& T, const float o[3], const float d[3], float tmax, float& t) {
    float e1[3], e2[3], p[3], tv[3], q[3];
    for (int a = 0; a < 3; a++) {
        e1[a] = T.v[1][a] - T.v[0][a];
        e2[a] = T.v[2][a] - T.v[0][a];
        tv[a] = o[a] - T.v[0][a];
    }
    p[0] = d[1] * e2[2] - d[2] * e2[1];
    p[1] = d[2] * e2[0] - d[0] * e2[2];
    p[2] = d[0] * e2[1] - d[1] * e2[0];
    q[0] = tv[1] * e1[2] - tv[2] * e1[1];
    q[1] = tv[2] * e1[0] - tv[0] * e1[2];
    q[2] = tv[0] * e1[1] - tv[1] * e1[0];
    const float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0f) return false;
    const float inv = 1.0f / det;
    const float b1 = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) * inv;
    const float b2 = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    return b1 >= 0.0f && b2 >= 0.0f && b1 + b2 <= 1.0f && t > 0.0f && t < tmax;
}

static bool slab(const float lo[3], const float hi[3], const float o[3], const float inv[3], float tmax) {
    float t0 = 0.0f, t1 = tmax;
    for (int a = 0; a < 3; a++) {
        float n = (lo[a] - o[a]) * inv[a], f = (hi[a] - o[a]) * inv[a];
        if (n > f) std::swap(n, f);
        t0 = std::max(t0, n);
        t1 = std::min(t1, f * 1.0000004f);
    }
    return t0 <= t1;
}

static double frand(unsigned& s) {
    s = s * 1664525u + 1013904223u;
    return (s >> 8) * (1.0 / 16777216.0);
}

int main(int argc, char** argv) {
    const int ntris = argc > 1 ? atoi(argv[1]) : 20000;
    unsigned seed = 12345;
    std::vector<tri> T(ntris);
    std::vector<aabb> B(ntris);
    for (int i = 0; i < ntris; i++) {
        float c[3];
        for (int a = 0; a < 3; a++) c[a] = float(frand(seed) * 2.0 - 1.0);
        const float r = float(0.002 + 0.05 * frand(seed) * frand(seed));
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) T[i].v[k][a] = c[a] + float((frand(seed) * 2.0 - 1.0) * r);
        for (int a = 0; a < 3; a++) {
            B[i].lo[a] = std::min(T[i].v[0][a], std::min(T[i].v[1][a], T[i].v[2][a]));
            B[i].hi[a] = std::max(T[i].v[0][a], std::max(T[i].v[1][a], T[i].v[2][a]));
        }
    }
    bvh_build_result bin;
    build_bvh(B, bin, 4);
    bvh8_result w;
    build_bvh8(bin, w);
    // 1. every primitive exactly once
    std::vector<int> seen(ntris, 0);
    for (uint32_t p : w.order) seen[p]++;
    for (int i = 0; i < ntris; i++)
        if (seen[i] != 1) {
            printf("primitive %d appears %d times\n", i, seen[i]);
            return 1;
        }
    // 2. conservative child boxes (post-order over the wide tree)
    std::vector<aabb> sub(w.nodes.size());
    for (size_t k = w.nodes.size(); k-- > 0;) {
        const bvh8_node& n = w.nodes[k];
        aabb acc;
        for (int a = 0; a < 3; a++) {
            acc.lo[a] = 3.4e38f;
            acc.hi[a] = -3.4e38f;
        }
        int rank = 0;
        for (int s = 0; s < 8; s++) {
            aabb cb;
            bool present = false;
            if (n.imask & (1u << s)) {
                cb = sub[n.child_base + rank++];
                present = true;
            } else if (n.meta[s] & k_meta_leaf) {
                const uint32_t cnt = ((n.meta[s] >> 5) & 3u) + 1u, off = n.meta[s] & 31u;
                for (int a = 0; a < 3; a++) {
                    cb.lo[a] = 3.4e38f;
                    cb.hi[a] = -3.4e38f;
                }
                for (uint32_t q = 0; q < cnt; q++)
                    for (int a = 0; a < 3; a++) {
                        cb.lo[a] = std::min(cb.lo[a], B[w.order[n.prim_base + off + q]].lo[a]);
                        cb.hi[a] = std::max(cb.hi[a], B[w.order[n.prim_base + off + q]].hi[a]);
                    }
                present = true;
            }
            if (!present) continue;
            float lo[3], hi[3];
            bvh8_child_box(n, s, lo, hi);
            for (int a = 0; a < 3; a++)
                if (lo[a] > cb.lo[a] || hi[a] < cb.hi[a]) {
                    printf("node %zu slot %d axis %d: box [%g,%g] misses [%g,%g]\n", k, s, a, lo[a], hi[a], cb.lo[a],
                           cb.hi[a]);
                    return 1;
                }
            for (int a = 0; a < 3; a++) {
                acc.lo[a] = std::min(acc.lo[a], cb.lo[a]);
                acc.hi[a] = std::max(acc.hi[a], cb.hi[a]);
            }
        }
        sub[k] = acc;
    }
    // 3. closest hit through the wide tree equals brute force
    int mism = 0, hits = 0;
    for (int r = 0; r < 2000; r++) {
        float o[3], d[3], inv[3];
        for (int a = 0; a < 3; a++) o[a] = float(frand(seed) * 2.4 - 1.2);
        double l = 0;
        for (int a = 0; a < 3; a++) {
            d[a] = float(frand(seed) * 2.0 - 1.0);
            l += d[a] * d[a];
        }
        for (int a = 0; a < 3; a++) {
            d[a] = float(d[a] / std::sqrt(l));
            inv[a] = 1.0f / d[a];
        }
        int bf = -1;
        float bt = 3.4e38f, t;
        for (int i = 0; i < ntris; i++)
            if (mt(T[i], o, d, bt, t)) {
                bt = t;
                bf = i;
            }
        int wf = -1;
        float wt = 3.4e38f;
        std::vector<uint32_t> stack{0};
        while (!stack.empty()) {
            const bvh8_node& n = w.nodes[stack.back()];
            stack.pop_back();
            int rank = 0;
            for (int s = 0; s < 8; s++) {
                const bool inner = n.imask & (1u << s);
                const uint32_t child = inner ? n.child_base + rank++ : 0;
                if (!inner && !(n.meta[s] & k_meta_leaf)) continue;
                float lo[3], hi[3];
                bvh8_child_box(n, s, lo, hi);
                if (!slab(lo, hi, o, inv, wt)) continue;
                if (inner) {
                    stack.push_back(child);
                } else {
                    const uint32_t cnt = ((n.meta[s] >> 5) & 3u) + 1u, off = n.meta[s] & 31u;
                    for (uint32_t q = 0; q < cnt; q++) {
                        const uint32_t p = w.order[n.prim_base + off + q];
                        if (mt(T[p], o, d, wt, t)) {
                            wt = t;
                            wf = int(p);
                        }
                    }
                }
            }
        }
        hits += bf >= 0;
        if (bf != wf && !(bf >= 0 && wf >= 0 && bt == wt)) mism++;
    }
    printf("nodes %zu depth %u hits %d mismatches %d\n", w.nodes.size(), w.depth, hits, mism);
    if (mism) return 1;
    printf("ok\n");
    return 0;
}
