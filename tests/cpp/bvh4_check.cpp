// Structural check of the four-wide BVH (rainbow_amd/csrc/bvh.cpp) on random
// triangle soups: its leaves are exactly the binary tree's leaves (every primitive
// reached once), every child box contains its subtree's primitive boxes, the
// recorded stack bound covers every root-to-node path, and a closest-hit walk over
// the four-wide tree finds the same nearest triangle as a brute-force scan.
// Prints "ok" or the first failure.
#include <algorithm>
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

static void child_box(const bvh4_node& n, int c, float lo[3], float hi[3]) {
    lo[0] = n.lox[c]; lo[1] = n.loy[c]; lo[2] = n.loz[c];
    hi[0] = n.hix[c]; hi[1] = n.hiy[c]; hi[2] = n.hiz[c];
}

int main(int argc, char** argv) {
    const int ntris = argc > 1 ? atoi(argv[1]) : 20000;
    unsigned seed = 777;
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
    bvh4_result w;
    build_bvh4(bin, w);
    // 1. every primitive reached exactly once; 2. conservative child boxes; 3. stack bound
    std::vector<int> seen(ntris, 0);
    std::vector<int> visited(w.nodes.size(), 0);
    uint32_t max_stack = 0;
    struct item {
        uint32_t node, pushed;
    };
    std::vector<item> work{{0, 0}};
    while (!work.empty()) {
        const item it = work.back();
        work.pop_back();
        if (visited[it.node]++) {
            printf("node %u reached twice\n", it.node);
            return 1;
        }
        const bvh4_node& n = w.nodes[it.node];
        int nch = 0;
        for (int c = 0; c < 4; c++) nch += n.child[c] != k_empty_child;
        if (nch == 0 || (nch < 2 && w.nodes.size() > 1)) {
            printf("node %u has %d children\n", it.node, nch);
            return 1;
        }
        const uint32_t below = it.pushed + uint32_t(nch - 1);
        max_stack = std::max(max_stack, below);
        for (int c = 0; c < 4; c++) {
            const int32_t ch = n.child[c];
            if (ch == k_empty_child) continue;
            float lo[3], hi[3];
            child_box(n, c, lo, hi);
            // the primitives below this child
            std::vector<uint32_t> prims;
            std::vector<int32_t> sub{ch};
            while (!sub.empty()) {
                const int32_t r = sub.back();
                sub.pop_back();
                if (r >= 0) {
                    for (int k = 0; k < 4; k++)
                        if (w.nodes[r].child[k] != k_empty_child) sub.push_back(w.nodes[r].child[k]);
                } else {
                    const uint32_t first = uint32_t(r) & 0x07FFFFFFu, cnt = ((uint32_t(r) >> 27) & 15u) + 1u;
                    for (uint32_t q = 0; q < cnt; q++) prims.push_back(bin.order[first + q]);
                }
            }
            for (uint32_t p : prims)
                for (int a = 0; a < 3; a++)
                    if (lo[a] > B[p].lo[a] || hi[a] < B[p].hi[a]) {
                        printf("node %u child %d axis %d misses primitive %u\n", it.node, c, a, p);
                        return 1;
                    }
            if (ch >= 0) {
                work.push_back({uint32_t(ch), below});
            } else {
                const uint32_t first = uint32_t(ch) & 0x07FFFFFFu, cnt = ((uint32_t(ch) >> 27) & 15u) + 1u;
                for (uint32_t q = 0; q < cnt; q++) seen[bin.order[first + q]]++;
            }
        }
    }
    for (int i = 0; i < ntris; i++)
        if (seen[i] != 1) {
            printf("primitive %d reached %d times\n", i, seen[i]);
            return 1;
        }
    for (size_t k = 0; k < w.nodes.size(); k++)
        if (visited[k] != 1) {
            printf("node %zu unreachable\n", k);
            return 1;
        }
    if (max_stack != w.stack) {
        printf("stack bound %u recorded, %u found\n", w.stack, max_stack);
        return 1;
    }
    // 4. closest hit through the four-wide tree equals brute force
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
        std::vector<int32_t> stack{0};
        while (!stack.empty()) {
            const int32_t ref = stack.back();
            stack.pop_back();
            if (ref < 0) {
                const uint32_t first = uint32_t(ref) & 0x07FFFFFFu, cnt = ((uint32_t(ref) >> 27) & 15u) + 1u;
                for (uint32_t q = 0; q < cnt; q++) {
                    const uint32_t p = bin.order[first + q];
                    if (mt(T[p], o, d, wt, t)) {
                        wt = t;
                        wf = int(p);
                    }
                }
                continue;
            }
            const bvh4_node& n = w.nodes[ref];
            for (int c = 0; c < 4; c++) {
                if (n.child[c] == k_empty_child) continue;
                float lo[3], hi[3];
                child_box(n, c, lo, hi);
                if (slab(lo, hi, o, inv, wt)) stack.push_back(n.child[c]);
            }
        }
        hits += bf >= 0;
        if (bf != wf && !(bf >= 0 && wf >= 0 && bt == wt)) mism++;
    }
    printf("nodes %zu depth %u stack %u hits %d mismatches %d\n", w.nodes.size(), w.depth, w.stack, hits, mism);
    if (mism) return 1;
    printf("ok\n");
    return 0;
}
