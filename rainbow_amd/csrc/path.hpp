// Device-side path state and the per-path stages shared by the single-render
// wavefront (tracer.hip) and the batched multi-job wavefront (batch.hip):
// BVH traversal, the SoA path state, samplers, and the surface lookup of an
// extension hit.  The arithmetic follows rainbow-cpu's path integrator
// (integrators/path_integrator.cpp:15-47, integrators/integrator.cpp:101-404).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rbw.h"
#include "bvh.hpp"
#include "device.hpp"
#include "guiding.hpp"
#include "shading.hpp"

namespace rbw {

#define RBW_STACK 64

// Both walks keep the top of each lane's stack in LDS, older entries in private
// memory: RBW_LDS_STACK 32-bit entries of the binary walk, or RBW_LDS_STACK / 2
// two-word groups of the wide walk.  Word e of lane t sits at
// rbw_lds_stack[e * blockDim.x + t] (conflict-free).  A kernel instantiated with
// LDS = true must be launched with lds_stack_bytes(blockDim.x) of dynamic LDS.
constexpr int RBW_LDS_STACK = 16;
extern __shared__ uint32_t rbw_lds_stack[];
inline size_t lds_stack_bytes(uint32_t block) { return size_t(RBW_LDS_STACK) * block * sizeof(uint32_t); }

// The shadow/MIS kernels are latency bound on dependent node fetches: with the binary
// walk six waves per SIMD (at most 80 VGPRs, a few spilled) beat the five its natural
// 95 VGPRs allow; the wide walk's larger state is left unconstrained (W: the width).
#define RBW_CONNECT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(W == 2 ? 6 : (W == 4 ? RBW_CONNECT4_WAVES : 1))))
// The shading kernel keeps four waves per SIMD (at most 128 VGPRs).
#define RBW_SHADE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4)))
// The binary walk of extension rays: seven waves per SIMD (at most 72 VGPRs, 4 spilled)
// beat the six of its natural 79 VGPRs; eight (64 VGPRs) spill 11 and run slower.
#define RBW_EXTEND_OCCUPANCY __attribute__((amdgpu_waves_per_eu(W == 2 ? 7 : (W == 4 ? RBW_EXTEND4_WAVES : 1))))
#ifndef RBW_EXTEND4_WAVES
#define RBW_EXTEND4_WAVES 6
#endif
#ifndef RBW_CONNECT4_WAVES
#define RBW_CONNECT4_WAVES 6
#endif

// ------------------------------------------------------------------ traversal

// acceptance bound of the next candidate: inside the same instanced mesh the
// reference compares against the mesh-local length directly (mesh.cpp:431);
// any other candidate sees the world length mapped into its local space.
__device__ __forceinline__ float thr_from(uint32_t best_inst, uint32_t cand_inst, float best_t, float len_w, float L0) {
    if (cand_inst != RBW_NONE && cand_inst == best_inst) return best_t;
    return len_w * L0;
}


// per-ray state of a closest-hit walk
struct trav_state {
    hitrec best;
    uint32_t best_inst_entity;  // entity of the best hit when it belongs to an instanced mesh
    float len_w;                // world length of the ray (shrinks with each accepted hit)
    float thr_local;            // current acceptance bound in local units
};

// the primitives of one leaf, slots [first, first + count) of (pa, pb, pc);
// returns true when an any-hit query may stop
template <bool VISIBLE_ONLY, bool ANY_HIT, bool COUNT>
__device__ __forceinline__ bool test_leaf(const dscene& S, const float4* __restrict__ pa_, const float4* __restrict__ pb_,
                                          const float4* __restrict__ pc_, uint32_t first, uint32_t count,
                                          const local_ray& lr, f3 o_w, const wt_setup& wt, uint32_t tri_test,
                                          uint32_t target_entity, trace_counters& cnt, trav_state& st) {
    for (uint32_t k = 0; k < count; k++) {
        const uint32_t slot = first + k;
        const float4 pa = pa_[slot];
        const uint32_t word = __float_as_uint(pa.w);
        if (VISIBLE_ONLY && (word & k_prim_invisible)) continue;
        if (COUNT) cnt.prims++;
        const uint32_t id = word & k_prim_id_mask;
        float t = 0.0f, b1 = 0.0f, b2 = 0.0f;
        bool hit;
        uint32_t inst_entity = RBW_NONE;
        if (word & k_prim_sphere) {
            const float4 pb = pb_[slot];
            const float thr = thr_from(st.best_inst_entity, RBW_NONE, st.best.t, st.len_w, lr.L0);
            const f3 ol = translate(o_w, f3{-pa.x, -pa.y, -pa.z});
            hit = sphere_test(pb.x, ol, lr.d, thr, t);
        } else {
            const float4 pb = pb_[slot], pc = pc_[slot];
            if (word & k_prim_instanced) inst_entity = S.tri_entity[id];
            const float thr = thr_from(st.best_inst_entity, inst_entity, st.best.t, st.len_w, lr.L0);
            const f3 p0 = f3{pa.x, pa.y, pa.z}, p1 = f3{pb.x, pb.y, pb.z}, p2 = f3{pc.x, pc.y, pc.z};
            if (tri_test == RBW_TRI_WATERTIGHT)
                hit = wt_test(p0, p1, p2, lr.o, lr.d, thr, wt.kz, wt.Sx, wt.Sy, wt.Sz, t, b1, b2);
            else
                hit = mt_test(p0, p1, p2, lr.o, lr.d, thr, t, b1, b2);
        }
        if (hit) {
            if (ANY_HIT) {
                const uint32_t ent = (word & k_prim_sphere) ? S.sphere_entity[id - S.num_tris] : S.tri_entity[id];
                if (ent != target_entity) {
                    st.best.prim = id;
                    st.best.t = t;
                    return true;
                }
                continue;
            }
            st.best.prim = id;
            st.best.t = t;
            st.best.b1 = b1;
            st.best.b2 = b2;
            st.best_inst_entity = inst_entity;
            st.len_w = t * lr.L1;  // entity.cpp:384
            st.thr_local = (inst_entity != RBW_NONE) ? t : st.len_w * lr.L0;
        }
    }
    return false;
}

// the nearest hit among the visible emitters' triangles (mesh lights, world space), with the
// leaf test's acceptance rule; used only while few emitter triangles exist (k_lights_few)
__device__ __forceinline__ hitrec nearest_emitter_hit(const dscene& S, const uint32_t* emitters, uint32_t num_emitters,
                                                      f3 o_w, f3 d_w, uint32_t tri_test) {
    const local_ray lr = make_local_ray(o_w, d_w);
    wt_setup wt;
    if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
    hitrec best;
    best.prim = RBW_MISS;
    best.t = 0.0f;
    best.b1 = best.b2 = 0.0f;
    float len_w = __builtin_inff();
    for (uint32_t k = 0; k < num_emitters; k++) {
        const dentity E = S.ents[emitters[k]];
        if (E.material == RBW_NONE) continue;
        for (uint32_t i = 0; i < E.count; i++) {
            const uint32_t tri = E.first + i;
            const f3 p0 = load3(S.pos, S.idx[3 * tri]), p1 = load3(S.pos, S.idx[3 * tri + 1]),
                     p2 = load3(S.pos, S.idx[3 * tri + 2]);
            float t, b1, b2;
            const float thr = len_w * lr.L0;
            const bool hit = tri_test == RBW_TRI_WATERTIGHT
                                 ? wt_test(p0, p1, p2, lr.o, lr.d, thr, wt.kz, wt.Sx, wt.Sy, wt.Sz, t, b1, b2)
                                 : mt_test(p0, p1, p2, lr.o, lr.d, thr, t, b1, b2);
            if (hit) {
                best.prim = tri;
                best.t = t;
                best.b1 = b1;
                best.b2 = b2;
                len_w = t * lr.L1;
            }
        }
    }
    return best;
}

// closest hit (or closest visible hit) through the binary BVH: near child first,
// the far child on a private stack
template <bool VISIBLE_ONLY, bool ANY_HIT, bool COUNT, bool LDS>
__device__ __forceinline__ hitrec traverse2(const dscene& S, f3 o_w, f3 d_w, float len_w, uint32_t tri_test,
                                            uint32_t target_entity, trace_counters& cnt) {
    const local_ray lr = make_local_ray(o_w, d_w);
    wt_setup wt;
    if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
    trav_state st;
    st.best.prim = RBW_MISS;
    st.best.t = 0.0f;
    st.best.b1 = st.best.b2 = 0.0f;
    st.best_inst_entity = RBW_NONE;
    st.len_w = len_w;
    st.thr_local = len_w * lr.L0;
    uint32_t stack[RBW_STACK];
    int sp = 0;
    uint32_t* ring = rbw_lds_stack + threadIdx.x;
    const uint32_t ring_stride = blockDim.x;
    auto push = [&](uint32_t v) {
        if (LDS) {
            uint32_t* e = ring + (uint32_t)(sp & (RBW_LDS_STACK - 1)) * ring_stride;
            if (sp >= RBW_LDS_STACK) stack[sp - RBW_LDS_STACK] = *e;  // the oldest ring entry moves out
            *e = v;
        } else {
            stack[sp] = v;
        }
        sp++;
    };
    auto pop = [&]() -> uint32_t {
        sp--;
        if (LDS) {
            uint32_t* e = ring + (uint32_t)(sp & (RBW_LDS_STACK - 1)) * ring_stride;
            const uint32_t v = *e;
            if (sp >= RBW_LDS_STACK) *e = stack[sp - RBW_LDS_STACK];  // and comes back
            return v;
        }
        return stack[sp];
    };
    int32_t cur = (int32_t)S.root;
    const f3 o = lr.o;
    const f3 inv = lr.inv;
    const f3 oi = f3{o.x * inv.x, o.y * inv.y, o.z * inv.z};
    // while-while walk with postponed leaves (Aila and Laine, "Understanding the efficiency of
    // ray traversal on GPUs", 2009): a wave walks inner nodes until every lane holds a leaf
    // to test (a lane that holds one keeps walking ahead), then tests leaves together, so
    // node and leaf work are not interleaved lane by lane
    bool done = false, has_leaf = false;
    int32_t leaf = 0;
    auto advance = [&]() {
        if (sp == 0) done = true;
        else cur = (int32_t)pop();
    };
    for (;;) {
        for (;;) {
            if (!done && !has_leaf && cur < 0) {
                if (cur != k_empty_child) {
                    leaf = cur;
                    has_leaf = true;
                }
                advance();
            }
            const bool walk = !done && cur >= 0;
            if (__ballot(walk) == 0) break;
            if (walk) {
                const float4* n = S.nodes + 4 * (size_t)cur;
                const float4 a = n[0], b = n[1], c = n[2];
                const int4 ch = *reinterpret_cast<const int4*>(n + 3);
                if (COUNT) cnt.nodes++;
                const float tmax = st.thr_local * 1.0000002f;
                const float lx0 = __builtin_fmaf(a.x, inv.x, -oi.x), lx1 = __builtin_fmaf(a.y, inv.x, -oi.x);
                const float ly0 = __builtin_fmaf(a.z, inv.y, -oi.y), ly1 = __builtin_fmaf(a.w, inv.y, -oi.y);
                const float lz0 = __builtin_fmaf(c.x, inv.z, -oi.z), lz1 = __builtin_fmaf(c.y, inv.z, -oi.z);
                const float rx0 = __builtin_fmaf(b.x, inv.x, -oi.x), rx1 = __builtin_fmaf(b.y, inv.x, -oi.x);
                const float ry0 = __builtin_fmaf(b.z, inv.y, -oi.y), ry1 = __builtin_fmaf(b.w, inv.y, -oi.y);
                const float rz0 = __builtin_fmaf(c.z, inv.z, -oi.z), rz1 = __builtin_fmaf(c.w, inv.z, -oi.z);
                const float lmin = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fmaxf(fminf(lz0, lz1), 0.0f));
                const float lmax = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fminf(fmaxf(lz0, lz1), tmax));
                const float rmin = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fmaxf(fminf(rz0, rz1), 0.0f));
                const float rmax = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fminf(fmaxf(rz0, rz1), tmax));
                const bool hl = lmin <= lmax * 1.0000004f;
                const bool hr = rmin <= rmax * 1.0000004f;
                if (hl && hr) {
                    const bool lfirst = lmin <= rmin;
                    cur = lfirst ? ch.x : ch.y;
                    push((uint32_t)(lfirst ? ch.y : ch.x));
                } else if (hl) {
                    cur = ch.x;
                } else if (hr) {
                    cur = ch.y;
                } else {
                    advance();
                }
            }
            // every lane holds a leaf or is done: test them
            if (__ballot(!(has_leaf || done)) == 0) break;
        }
        while (has_leaf) {
            has_leaf = false;
            const uint32_t first = (uint32_t)leaf & 0x07FFFFFFu;
            const uint32_t count = (((uint32_t)leaf >> 27) & 15u) + 1u;
            if (test_leaf<VISIBLE_ONLY, ANY_HIT, COUNT>(S, S.prim_a, S.prim_b, S.prim_c, first, count, lr, o_w, wt,
                                                        tri_test, target_entity, cnt, st))
                return st.best;
            if (!done && cur < 0) {
                if (cur != k_empty_child) {
                    leaf = cur;
                    has_leaf = true;
                }
                advance();
            }
        }
        if (done) break;
    }
    return st.best;
}


// closest hit (or closest visible hit) through the eight-wide BVH.  The walk keeps
// a group (a node's inner children still to visit, as a bit mask in visiting
// order j, child slot j ^ octant); a visited node tests its eight child boxes,
// tests hit leaves at once in visiting order and makes its hit inner children
// the next group, the rest of the current group going onto the stack.
template <bool VISIBLE_ONLY, bool ANY_HIT, bool COUNT, bool LDS>
__device__ __forceinline__ hitrec traverse8(const dscene& S, f3 o_w, f3 d_w, float len_w, uint32_t tri_test,
                                            uint32_t target_entity, trace_counters& cnt) {
    const local_ray lr = make_local_ray(o_w, d_w);
    wt_setup wt;
    if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
    trav_state st;
    st.best.prim = RBW_MISS;
    st.best.t = 0.0f;
    st.best.b1 = st.best.b2 = 0.0f;
    st.best_inst_entity = RBW_NONE;
    st.len_w = len_w;
    st.thr_local = len_w * lr.L0;
    const f3 o = lr.o;
    const f3 inv = lr.inv;
    const f3 oi = f3{o.x * inv.x, o.y * inv.y, o.z * inv.z};
    const uint32_t oct = (inv.x < 0.0f ? 4u : 0u) | (inv.y < 0.0f ? 2u : 0u) | (inv.z < 0.0f ? 1u : 0u);
    uint64_t stack[RBW_STACK];
    int sp = 0;
    constexpr int K8 = RBW_LDS_STACK / 2;
    uint32_t* ring = rbw_lds_stack + threadIdx.x;
    const uint32_t ring_stride = blockDim.x;
    auto push = [&](uint64_t v) {
        if (LDS) {
            uint32_t* e = ring + (uint32_t)(sp & (K8 - 1)) * 2u * ring_stride;
            if (sp >= K8) stack[sp - K8] = (uint64_t)e[0] | ((uint64_t)e[ring_stride] << 32);
            e[0] = (uint32_t)v;
            e[ring_stride] = (uint32_t)(v >> 32);
        } else {
            stack[sp] = v;
        }
        sp++;
    };
    auto pop = [&]() -> uint64_t {
        sp--;
        if (LDS) {
            uint32_t* e = ring + (uint32_t)(sp & (K8 - 1)) * 2u * ring_stride;
            const uint64_t v = (uint64_t)e[0] | ((uint64_t)e[ring_stride] << 32);
            if (sp >= K8) {
                const uint64_t o = stack[sp - K8];
                e[0] = (uint32_t)o;
                e[ring_stride] = (uint32_t)(o >> 32);
            }
            return v;
        }
        return stack[sp];
    };
    // the root as a group of one inner child
    uint32_t g_base = S.root8, g_imask = 1u << oct, g_pending = 1u;
    for (;;) {
        if (g_pending == 0u) {
            if (sp == 0) break;
            const uint64_t e = pop();
            g_base = (uint32_t)e;
            g_imask = (uint32_t)(e >> 32) & 0xFFu;
            g_pending = (uint32_t)(e >> 40) & 0xFFu;
            continue;
        }
        const uint32_t j = (uint32_t)__builtin_ctz(g_pending);
        g_pending &= g_pending - 1u;
        const uint32_t slot = j ^ oct;
        const uint32_t node = g_base + (uint32_t)__builtin_popcount(g_imask & ((1u << slot) - 1u));
        if (g_pending) push((uint64_t)g_base | ((uint64_t)g_imask << 32) | ((uint64_t)g_pending << 40));
        const float4* n = S.nodes8 + 5 * (size_t)node;
        const float4 n0 = n[0], n1 = n[1], n2 = n[2], n3 = n[3], n4 = n[4];
        if (COUNT) cnt.nodes++;
        const uint32_t ew = __float_as_uint(n0.w);
        const float sx = __uint_as_float((ew & 0xFFu) << 23), sy = __uint_as_float(((ew >> 8) & 0xFFu) << 23),
                    sz = __uint_as_float(((ew >> 16) & 0xFFu) << 23);
        const uint32_t imask = ew >> 24;
        const uint32_t child_base = __float_as_uint(n1.x), prim_base = __float_as_uint(n1.y);
        const uint32_t meta0 = __float_as_uint(n1.z), meta1 = __float_as_uint(n1.w);
        const uint32_t qlx[2] = {__float_as_uint(n2.x), __float_as_uint(n2.y)};
        const uint32_t qly[2] = {__float_as_uint(n2.z), __float_as_uint(n2.w)};
        const uint32_t qlz[2] = {__float_as_uint(n3.x), __float_as_uint(n3.y)};
        const uint32_t qhx[2] = {__float_as_uint(n3.z), __float_as_uint(n3.w)};
        const uint32_t qhy[2] = {__float_as_uint(n4.x), __float_as_uint(n4.y)};
        const uint32_t qhz[2] = {__float_as_uint(n4.z), __float_as_uint(n4.w)};
        const float tmax = st.thr_local * 1.0000002f;
        uint32_t inner_hits = 0u, leaf_hits = 0u;
        float tnear[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int w = c >> 2, sh = (c & 3) * 8;
            // child box p + q s (q s is exact), then the slab distances as in the binary walk
            const float bx0 = __builtin_fmaf((float)((qlx[w] >> sh) & 0xFFu), sx, n0.x);
            const float bx1 = __builtin_fmaf((float)((qhx[w] >> sh) & 0xFFu), sx, n0.x);
            const float by0 = __builtin_fmaf((float)((qly[w] >> sh) & 0xFFu), sy, n0.y);
            const float by1 = __builtin_fmaf((float)((qhy[w] >> sh) & 0xFFu), sy, n0.y);
            const float bz0 = __builtin_fmaf((float)((qlz[w] >> sh) & 0xFFu), sz, n0.z);
            const float bz1 = __builtin_fmaf((float)((qhz[w] >> sh) & 0xFFu), sz, n0.z);
            const float tx0 = __builtin_fmaf(bx0, inv.x, -oi.x), tx1 = __builtin_fmaf(bx1, inv.x, -oi.x);
            const float ty0 = __builtin_fmaf(by0, inv.y, -oi.y), ty1 = __builtin_fmaf(by1, inv.y, -oi.y);
            const float tz0 = __builtin_fmaf(bz0, inv.z, -oi.z), tz1 = __builtin_fmaf(bz1, inv.z, -oi.z);
            const float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
            const float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
            const bool hit = tmin <= tfar * 1.0000004f;
            tnear[c] = tmin;
            const bool inner = (imask >> c) & 1u;
            const bool leaf = (((w ? meta1 : meta0) >> sh) & (uint32_t)k_meta_leaf) != 0u;
            const uint32_t jbit = 1u << ((uint32_t)c ^ oct);
            inner_hits |= (hit && inner) ? jbit : 0u;
            leaf_hits |= (hit && leaf) ? jbit : 0u;
        }
        // hit leaves nearest box first; a hit in one may already be nearer than the next boxes
        while (leaf_hits) {
            uint32_t ls = 0u;
            float tl = 3.4e38f;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const bool cand = (leaf_hits >> ((uint32_t)c ^ oct)) & 1u;
                if (cand && tnear[c] < tl) {
                    tl = tnear[c];
                    ls = (uint32_t)c;
                }
            }
            leaf_hits &= ~(1u << (ls ^ oct));
            if (tl > st.thr_local * 1.0000004f) break;
            const uint32_t meta = ((ls & 4u ? meta1 : meta0) >> ((ls & 3u) * 8u)) & 0xFFu;
            const uint32_t first = prim_base + (meta & 31u), count = ((meta >> 5) & 3u) + 1u;
            if (test_leaf<VISIBLE_ONLY, ANY_HIT, COUNT>(S, S.prim8_a, S.prim8_b, S.prim8_c, first, count, lr, o_w, wt,
                                                        tri_test, target_entity, cnt, st))
                return st.best;
        }
        g_base = child_base;
        g_imask = imask;
        g_pending = inner_hits;
    }
    return st.best;
}

// One four-wide node visit of a walk at node `cur`: the child boxes' slab tests against
// [0, tmax], the hit children sorted by entry distance (a sorting network; misses last),
// the nearest made `cur` and the others pushed far to near.  Returns false when no child
// is hit (the caller pops).
template <class PUSH>
__device__ __forceinline__ bool visit4(const dscene& S, int32_t& cur, f3 inv, f3 oi, float tmax, PUSH&& push) {
    const float4* n = S.nodes4 + 8 * (size_t)cur;
    const float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5];
    const int4 ch = *reinterpret_cast<const int4*>(n + 6);
    float d[4];
    int32_t r[4] = {ch.x, ch.y, ch.z, ch.w};
    const float blx[4] = {lx.x, lx.y, lx.z, lx.w}, bhx[4] = {hx.x, hx.y, hx.z, hx.w};
    const float bly[4] = {ly.x, ly.y, ly.z, ly.w}, bhy[4] = {hy.x, hy.y, hy.z, hy.w};
    const float blz[4] = {lz.x, lz.y, lz.z, lz.w}, bhz[4] = {hz.x, hz.y, hz.z, hz.w};
    uint32_t nhit = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float x0 = __builtin_fmaf(blx[c], inv.x, -oi.x), x1 = __builtin_fmaf(bhx[c], inv.x, -oi.x);
        const float y0 = __builtin_fmaf(bly[c], inv.y, -oi.y), y1 = __builtin_fmaf(bhy[c], inv.y, -oi.y);
        const float z0 = __builtin_fmaf(blz[c], inv.z, -oi.z), z1 = __builtin_fmaf(bhz[c], inv.z, -oi.z);
        const float tmin = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
        const float tfar = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
        const bool hit = r[c] != k_empty_child && tmin <= tfar * 1.0000004f;
        d[c] = hit ? tmin : __builtin_inff();
        nhit += hit ? 1u : 0u;
    }
    auto cswap = [&](int i, int j) {
        const bool sw = d[j] < d[i];
        const float di = d[i], dj = d[j];
        const int32_t ri = r[i], rj = r[j];
        d[i] = sw ? dj : di;
        d[j] = sw ? di : dj;
        r[i] = sw ? rj : ri;
        r[j] = sw ? ri : rj;
    };
    cswap(0, 1);
    cswap(2, 3);
    cswap(0, 2);
    cswap(1, 3);
    cswap(1, 2);
    if (nhit == 0) return false;
    cur = r[0];
    if (nhit > 3) push((uint32_t)r[3]);
    if (nhit > 2) push((uint32_t)r[2]);
    if (nhit > 1) push((uint32_t)r[1]);
    return true;
}

// closest hit (or closest visible hit) through the four-wide BVH: the walk of traverse2
// with four child boxes per node (one 128-byte line), hit children visited nearest
// first and the others pushed far to near; postponed leaves as there
template <bool VISIBLE_ONLY, bool ANY_HIT, bool COUNT, bool LDS>
__device__ __forceinline__ hitrec traverse4(const dscene& S, f3 o_w, f3 d_w, float len_w, uint32_t tri_test,
                                            uint32_t target_entity, trace_counters& cnt) {
    const local_ray lr = make_local_ray(o_w, d_w);
    wt_setup wt;
    if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
    trav_state st;
    st.best.prim = RBW_MISS;
    st.best.t = 0.0f;
    st.best.b1 = st.best.b2 = 0.0f;
    st.best_inst_entity = RBW_NONE;
    st.len_w = len_w;
    st.thr_local = len_w * lr.L0;
    uint32_t stack[RBW_STACK];
    int sp = 0;
    uint32_t* ring = rbw_lds_stack + threadIdx.x;
    const uint32_t ring_stride = blockDim.x;
    auto push = [&](uint32_t v) {
        if (LDS) {
            uint32_t* e = ring + (uint32_t)(sp & (RBW_LDS_STACK - 1)) * ring_stride;
            if (sp >= RBW_LDS_STACK) stack[sp - RBW_LDS_STACK] = *e;
            *e = v;
        } else {
            stack[sp] = v;
        }
        sp++;
    };
    auto pop = [&]() -> uint32_t {
        sp--;
        if (LDS) {
            uint32_t* e = ring + (uint32_t)(sp & (RBW_LDS_STACK - 1)) * ring_stride;
            const uint32_t v = *e;
            if (sp >= RBW_LDS_STACK) *e = stack[sp - RBW_LDS_STACK];
            return v;
        }
        return stack[sp];
    };
    int32_t cur = (int32_t)S.root4;
    const f3 o = lr.o;
    const f3 inv = lr.inv;
    const f3 oi = f3{o.x * inv.x, o.y * inv.y, o.z * inv.z};
    bool done = false, has_leaf = false;
    int32_t leaf = 0;
    auto advance = [&]() {
        if (sp == 0) done = true;
        else cur = (int32_t)pop();
    };
    for (;;) {
        for (;;) {
            if (!done && !has_leaf && cur < 0) {
                if (cur != k_empty_child) {
                    leaf = cur;
                    has_leaf = true;
                }
                advance();
            }
            const bool walk = !done && cur >= 0;
            if (__ballot(walk) == 0) break;
            if (walk) {
                if (COUNT) cnt.nodes++;
                if (!visit4(S, cur, inv, oi, st.thr_local * 1.0000002f, push)) advance();
            }
            if (__ballot(!(has_leaf || done)) == 0) break;
        }
        while (has_leaf) {
            has_leaf = false;
            const uint32_t first = (uint32_t)leaf & 0x07FFFFFFu;
            const uint32_t count = (((uint32_t)leaf >> 27) & 15u) + 1u;
            if (test_leaf<VISIBLE_ONLY, ANY_HIT, COUNT>(S, S.prim_a, S.prim_b, S.prim_c, first, count, lr, o_w, wt,
                                                        tri_test, target_entity, cnt, st))
                return st.best;
            if (!done && cur < 0) {
                if (cur != k_empty_child) {
                    leaf = cur;
                    has_leaf = true;
                }
                advance();
            }
        }
        if (done) break;
    }
    return st.best;
}

// W: the BVH width the kernel is instantiated for (each kernel carries one walk)
// LDS: the kernel provides the walk's LDS stack (lds_stack_bytes of dynamic LDS)
template <int W, bool VISIBLE_ONLY, bool ANY_HIT, bool COUNT, bool LDS = false>
__device__ __forceinline__ hitrec traverse(const dscene& S, f3 o_w, f3 d_w, float len_w, uint32_t tri_test,
                                           uint32_t target_entity, trace_counters& cnt) {
    if (W == 8) return traverse8<VISIBLE_ONLY, ANY_HIT, COUNT, LDS>(S, o_w, d_w, len_w, tri_test, target_entity, cnt);
    if (W == 4) return traverse4<VISIBLE_ONLY, ANY_HIT, COUNT, LDS>(S, o_w, d_w, len_w, tri_test, target_entity, cnt);
    return traverse2<VISIBLE_ONLY, ANY_HIT, COUNT, LDS>(S, o_w, d_w, len_w, tri_test, target_entity, cnt);
}

// find_emitter's walk (integrator.cpp:49-82): the closest visible hit of a BSDF-sampled
// ray, or, while the scene has few emitter triangles (k_lights_few), the nearest emitter
// triangle hit when no other visible surface lies before it (RBW_MISS otherwise) -- the
// same answer for the caller, which only uses hits on emitters
template <int W, bool COUNT, bool LDS>
__device__ __forceinline__ hitrec find_emitter_walk(const dscene& S, f3 o, f3 d, uint32_t tri_test,
                                                    trace_counters& cnt) {
    if (S.light_flags & k_lights_few) {
        hitrec h = nearest_emitter_hit(S, S.emitters, S.num_emitters, o, d, tri_test);
        if (h.prim != RBW_MISS) {
            const uint32_t le = S.tri_entity[h.prim];
            const hitrec occ = traverse<W, true, true, COUNT, LDS>(S, o, d, h.t * make_local_ray(o, d).L1, tri_test,
                                                                   le, cnt);
            if (occ.prim != RBW_MISS) h.prim = RBW_MISS;
        }
        return h;
    }
    return traverse<W, true, false, COUNT, LDS>(S, o, d, __builtin_inff(), tri_test, RBW_NONE, cnt);
}

// ------------------------------------------------------------------ wave helpers

// slot in an append queue for every lane with `pred` set; call with the whole wave converged
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t prefix = __popcll(mask & lt);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + prefix;
}

__device__ __forceinline__ void wave_add64(unsigned long long* counter, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (__lane_id() == 0 && x) atomicAdd(counter, x);
}

// Statistics counters striped over RBW_STRIPES cache lines (RBW_STRIPE_WORDS words
// apart, picked by block): atomics of many waves on one address serialise in one L2
// channel.  The host sums the stripes.
constexpr uint32_t RBW_STRIPES = 64;
constexpr uint32_t RBW_STRIPE_WORDS = 16;
__device__ __forceinline__ void stripe_add64(unsigned long long* counters, uint32_t idx, uint32_t v) {
    wave_add64(counters + (blockIdx.x & (RBW_STRIPES - 1)) * RBW_STRIPE_WORDS + idx, v);
}

// Reserves n consecutive entries of *counter for this thread, with one atomic per
// block (a queue's tail is a single address).  Every thread of the block must call
// it; blockDim.x is a multiple of 64 and at most 1024.
__device__ __forceinline__ uint32_t block_reserve(uint32_t* counter, uint32_t n) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_base;
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    uint32_t x = n;  // inclusive scan over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if ((int)lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < nwaves; w++) {
            const uint32_t c = s_wave[w];
            s_wave[w] = t;
            t += c;
        }
        s_base = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    const uint32_t r = s_base + s_wave[wave] + x - n;
    __syncthreads();  // the next call reuses s_wave and s_base
    return r;
}

// block_reserve of one entry where pred holds
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool pred) {
    return block_reserve(counter, pred ? 1u : 0u);
}

// ------------------------------------------------------------------ path state

enum fidx : uint32_t {
    F_OX, F_OY, F_OZ, F_DX, F_DY, F_DZ,          // path ray (world)
    F_BX, F_BY, F_BZ, F_VX, F_VY, F_VZ,          // beta, value
    F_PX, F_PY,                                  // raster position of the camera sample
    F_HT,                                        // extension hit t (local)
    F_SOX, F_SOY, F_SOZ, F_SDX, F_SDY, F_SDZ, F_SLEN,  // shadow ray
    F_SFX, F_SFY, F_SFZ,                         // f * Le * w / pdf of the light sample
    F_MPX, F_MPY, F_MPZ, F_MNX, F_MNY, F_MNZ,    // vertex point and normal (MIS ray origin)
    F_MWX, F_MWY, F_MWZ, F_MFX, F_MFY, F_MFZ, F_MPDF,  // BSDF sample direction, f*|cos|, pdf
    F_MLEN,                                      // world length of the MIS ray to its nearest emitter triangle
    F_CBX, F_CBY, F_CBZ,                         // beta at the vertex
    F_COUNT
};
// U_DIMS: the stratified sampler's draws made by the sample (1D | 2D << 16)
// U_MEDIUM: the medium the path's ray travels in (volume path integrator; RBW_NONE = vacuum)
// U_MPRIM: the MIS ray's nearest visible emitter triangle (RBW_MISS: none), with F_MLEN
// (batched jobs' persistent occlusion walk)
enum uidx : uint32_t { U_PIXEL, U_SAMPLE, U_BOUNCE, U_HPRIM, U_CFLAGS, U_SENT, U_PENDING, U_JOB, U_STATE, U_DIMS, U_MEDIUM,
                       U_MPRIM, U_COUNT };

// connect flags (bits 0-7 of U_CFLAGS; the bounce above): the shadow and MIS rays to trace;
// the answers of the persistent occlusion walks: C_LIT, the shadow ray reaches the light;
// C_MOCC, something is in the way of the MIS ray before its nearest emitter triangle
constexpr uint32_t C_SHADOW = 1u, C_MIS = 2u, C_LIT = 4u, C_MOCC = 8u;

struct dpaths {
    float* f;      // F_COUNT * cap
    uint32_t* u;   // U_COUNT * cap
    uint64_t* rng; // pcg32 state, or the replay cursor
    uint32_t cap;
    __device__ float& F(uint32_t k, uint32_t i) const { return f[(size_t)k * cap + i]; }
    __device__ uint32_t& U(uint32_t k, uint32_t i) const { return u[(size_t)k * cap + i]; }
    __device__ f3 F3(uint32_t k, uint32_t i) const { return f3{F(k, i), F(k + 1, i), F(k + 2, i)}; }
    __device__ void setF3(uint32_t k, uint32_t i, f3 v) const {
        F(k, i) = v.x;
        F(k + 1, i) = v.y;
        F(k + 2, i) = v.z;
    }
};

struct dcamera {
    float r2c[16];
    float c2w[16];
    uint32_t width, height;
    int32_t bx0, by0, bx1, by1;
    float lens_radius, focus;
};

// guided continuation: per guide-queue row the surface record, MLP query and
// head outputs; per (slot, bounce) the training record of the sampled direction
enum gsurf : uint32_t { G_PX, G_PY, G_PZ, G_NX, G_NY, G_NZ, G_WX, G_WY, G_WZ, G_XX, G_XY, G_XZ, G_YX, G_YY, G_YZ,
                        G_ZX, G_ZY, G_ZZ, G_COUNT };
enum grec : uint32_t { R_Q0 = 0, R_DX = 9, R_DY, R_DZ, R_PDF, R_BX, R_BY, R_BZ, R_MX, R_MY, R_MZ, R_COUNT };
// a training record's floats (R_COUNT and one pad: five 16-byte words, R_PDF and R_MX on word boundaries)
constexpr uint32_t R_STRIDE = 20;
static_assert(R_COUNT + 1 == R_STRIDE && R_PDF % 4 == 0 && R_MX % 4 == 0, "record words");

struct dguide {
    uint32_t enable;
    float alpha;                 // probability of sampling the guide
    uint32_t train;
    f3 smin, sinv;               // position normalisation: (p - smin) * sinv
    float* surf;                 // G_COUNT * cap (SoA)
    uint32_t* surf_mat;          // material of the row
    uint32_t* gq;                // guide queue: slot of each row
    uint32_t* gq_count;
    float* query;                // 9 per row
    float* mix;                  // 48 per row (head outputs)
    float* rec;                  // R_STRIDE floats per record, record slot * rec_depth + bounce
    uint32_t* rec_flag;          // 0 none, 1 sampled, 2 closed
    uint32_t rec_depth;          // records per slot (at least every path's max depth)
    uint32_t cap;
    // training batch
    float* t_in;                 // 9 per sample
    float* t_dir;                // 3 per sample
    float* t_c;                  // radiance / pdf
    uint32_t* t_count;
    uint32_t t_cap;
    uint32_t light_records;      // also train on unoccluded light samples
    uint32_t depth;              // vertices below this bounce are guided (>= max_depth: all)
    uint32_t product;            // records weigh f |cos| L_i / pdf (else L_i / pdf)
};

struct dparams {
    uint32_t spp, max_depth;
    float rr_threshold;
    uint32_t seed;
    uint32_t sampler;
    uint32_t tri_test;
    uint32_t shadow_mode;
    const uint64_t* replay_offsets;
    const float* replay_values;
    uint32_t* debug;  // device debug hit buffer or null
    uint32_t strata_x, strata_y, dims_1d, dims_2d;  // RBW_SAMPLER_STRATIFIED
    uint32_t emitter_samples, bsdf_samples;          // RBW_INTEGRATOR_DIRECT
    uint32_t medium_mis;                             // RBW_INTEGRATOR_VOLUME_PATH: phase-sampled emitter hits count
    dguide g;
};

// ---- the stratified sampler's strata and jitter (no tables)
// permute(i, l, p): a pseudo-random permutation of [0, l) keyed by p, and
// randfloat(i, p): a hashed float in [0, 1); both from A. Kensler, "Correlated
// Multi-Jittered Sampling", Pixar Technical Memo 13-01 (2013), section 3.
__device__ __forceinline__ uint32_t cmj_permute(uint32_t i, uint32_t l, uint32_t p) {
    uint32_t w = l - 1u;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {  // cycle-walking: at most a few rounds, as w < 2l
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

__device__ __forceinline__ float cmj_randfloat(uint32_t i, uint32_t p) {
    i ^= p;
    i ^= i >> 17;
    i ^= i >> 10;
    i *= 0xb36534e5u;
    i ^= i >> 12;
    i ^= i >> 21;
    i *= 0x93fc4795u;
    i ^= 0xdf6e307fu;
    i ^= i >> 17;
    i *= 1u | p >> 18;
    return (float)i * (1.0f / 4294967808.0f);
}

// the pattern key of (seed, pixel, dimension): SplitMix64's finaliser over a mixed word
__device__ __forceinline__ uint32_t stratum_key(uint32_t seed, uint32_t pixel, uint32_t dim) {
    uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)pixel * 0xD1B54A32D192ED03ull +
                 (uint64_t)(dim + 1u) * 0x8CB92BA72F3D8DD7ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

// glm mat4 * vec4 (xyz): (m0*x + m1*y) + (m2*z + m3*w)
__device__ __forceinline__ f3 mat_mul(const float* m, f3 v, float w) {
    f3 r;
    r.x = (m[0] * v.x + m[4] * v.y) + (m[8] * v.z + m[12] * w);
    r.y = (m[1] * v.x + m[5] * v.y) + (m[9] * v.z + m[13] * w);
    r.z = (m[2] * v.x + m[6] * v.y) + (m[10] * v.z + m[14] * w);
    return r;
}

struct sampler_state {
    uint64_t state;
    uint64_t inc;
    uint64_t end;
    const float* replay;
    // stratified: the first d1 1D and d2 2D draws (d1 = d2 = 0 otherwise)
    uint32_t d1 = 0, d2 = 0, sx = 1, sy = 1, sseed = 0;
    uint32_t pixel = 0, sample = 0, dims = 0;
    __device__ void stratify(const dparams& prm, uint32_t px, uint32_t s, uint32_t drawn) {
        d1 = prm.dims_1d;
        d2 = prm.dims_2d;
        sx = prm.strata_x;
        sy = prm.strata_y;
        sseed = prm.seed;
        pixel = px;
        sample = s;
        dims = drawn;
    }
    __device__ float next() {
        if ((dims & 0xFFFFu) < d1) {
            const uint32_t k = dims & 0xFFFFu;
            dims += 1u;
            const uint32_t n = sx * sy;
            const uint32_t key = stratum_key(sseed, pixel, 2u * k);
            const float x = ((float)cmj_permute(sample, n, key) + cmj_randfloat(sample, key * 0x68bc21ebu)) / (float)n;
            return (x < k_one_minus_epsilon) ? x : k_one_minus_epsilon;
        }
        return uniform();
    }
    __device__ f2 next2() {
        if ((dims >> 16) < d2) {
            const uint32_t k = dims >> 16;
            dims += 1u << 16;
            const uint32_t key = stratum_key(sseed, pixel, 2u * k + 1u);
            const uint32_t c = cmj_permute(sample, sx * sy, key);
            const float x = ((float)(c % sx) + cmj_randfloat(sample, key * 0xa399d265u)) / (float)sx;
            const float y = ((float)(c / sx) + cmj_randfloat(sample, key * 0x711ad6a5u)) / (float)sy;
            return f2{(x < k_one_minus_epsilon) ? x : k_one_minus_epsilon, (y < k_one_minus_epsilon) ? y : k_one_minus_epsilon};
        }
        const float a = uniform();
        const float b = uniform();
        return f2{a, b};
    }
    // one draw of the replayed or per-sample pcg32 stream
    __device__ float uniform() {
        if (replay) return state < end ? replay[state++] : 0.0f;
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        const uint32_t r = (xs >> rot) | (xs << ((~rot + 1u) & 31));
        const float x = (float)r * 2.3283064365386963e-10f;
        return (x < k_one_minus_epsilon) ? x : k_one_minus_epsilon;
    }
};

__device__ __forceinline__ void store_sampler(const dpaths& P, uint32_t slot, const sampler_state& smp) {
    P.rng[slot] = smp.state;
    if (smp.d1 | smp.d2) P.U(U_DIMS, slot) = smp.dims;
}

__device__ __forceinline__ sampler_state load_sampler(const dpaths& P, const dparams& prm, uint32_t slot) {
    sampler_state s;
    const uint64_t sid = (uint64_t)P.U(U_PIXEL, slot) * prm.spp + P.U(U_SAMPLE, slot);
    s.state = P.rng[slot];
    if (prm.sampler == RBW_SAMPLER_REPLAY) {
        s.replay = prm.replay_values;
        s.end = prm.replay_offsets[sid + 1];
        s.inc = 0;
    } else {
        s.replay = nullptr;
        s.end = 0;
        s.inc = (sample_stream(prm.seed, sid) << 1u) | 1u;
        if (prm.sampler == RBW_SAMPLER_STRATIFIED)
            s.stratify(prm, P.U(U_PIXEL, slot), P.U(U_SAMPLE, slot), P.U(U_DIMS, slot));
    }
    return s;
}

__device__ __forceinline__ uint32_t* debug_slot(const dpaths& P, const dparams& prm, uint32_t slot, int bounce) {
    if (!prm.debug || bounce < 0 || (uint32_t)bounce >= prm.max_depth) return nullptr;
    const uint64_t sid = (uint64_t)P.U(U_PIXEL, slot) * prm.spp + P.U(U_SAMPLE, slot);
    return prm.debug + (sid * prm.max_depth + (uint32_t)bounce) * 3;
}

// surface interaction at the extension hit: entity::intersect's local ray and
// local -> world mapping (entity.cpp:18-58) for a mesh (identity) or sphere (translation)
__device__ __forceinline__ surf hit_surface(const dscene& S, uint32_t prim, f3 o_w, f3 d_w, uint32_t& entity) {
    const local_ray lr = make_local_ray(o_w, d_w);
    if (prim < S.num_tris) {
        entity = S.tri_entity[prim];
        const dentity E = S.ents[entity];
        const uint32_t i0 = S.idx[3 * prim], i1 = S.idx[3 * prim + 1], i2 = S.idx[3 * prim + 2];
        const f3 p0 = load3(S.pos, i0), p1 = load3(S.pos, i1), p2 = load3(S.pos, i2);
        float t, b1, b2;
        mt_test(p0, p1, p2, lr.o, lr.d, __builtin_inff(), t, b1, b2);
        const float b0 = 1.0f - b1 - b2;
        const surf l = triangle_interaction(S, E, prim, lr.d, b0, b1, b2);
        return surf_to_world(l, false, f3{0, 0, 0});
    }
    const uint32_t sph = prim - S.num_tris;
    entity = S.sphere_entity[sph];
    const dentity E = S.ents[entity];
    const float4 sp = S.spheres[sph];
    const f3 c = f3{sp.x, sp.y, sp.z};
    const f3 ol = translate(o_w, f3{-c.x, -c.y, -c.z});
    float t = 0.0f;
    sphere_test(sp.w, ol, lr.d, __builtin_inff(), t);
    const surf l = sphere_interaction(sp.w, (E.flags & RBW_ENTITY_REVERSE_ORIENTATION) != 0, ol, lr.d, t);
    return surf_to_world(l, true, c);
}

// sampler_integrator.cpp:98-127 + perspective_camera.cpp:78-99: the camera sample s of
// pixel (x, y) into path slot `slot`
__device__ __forceinline__ void raygen_path(const dcamera& cam, const dparams& prm, const dpaths& P, uint32_t slot,
                                            int32_t x, int32_t y, uint32_t s) {
    const uint32_t pixel = (uint32_t)y * cam.width + (uint32_t)x;
    const uint64_t sid = (uint64_t)pixel * prm.spp + s;
    sampler_state smp;
    if (prm.sampler == RBW_SAMPLER_REPLAY) {
        smp.replay = prm.replay_values;
        smp.state = prm.replay_offsets[sid];
        smp.end = prm.replay_offsets[sid + 1];
        smp.inc = 0;
    } else {
        pcg32 g;
        g.seed(sample_stream(prm.seed, sid));
        smp.state = g.state;
        smp.inc = g.inc;
        smp.replay = nullptr;
        smp.end = 0;
        if (prm.sampler == RBW_SAMPLER_STRATIFIED) smp.stratify(prm, pixel, s, 0u);
    }
    const f2 j = smp.next2();
    const float px = (float)x + j.x, py = (float)y + j.y;
    const f2 lu = smp.next2();
    // perspective_camera::sample(position, sample)
    const f3 target = mat_mul(cam.r2c, f3{px, py, 0.0f}, 1.0f);
    f3 d = normalize(normalize(target));
    f3 o = f3{0.0f, 0.0f, 0.0f};
    if (cam.lens_radius > 0.0f) {
        const f3 focus = o + d * (cam.focus / d.z);
        const f2 disk = concentric_sample_disk(lu);
        o = f3{disk.x * cam.lens_radius, disk.y * cam.lens_radius, 0.0f};
        d = normalize(focus - o);
    }
    // transform::operator()(ray) with CameraToWorld (transform.cpp:94-105)
    const f3 dv = mat_mul(cam.c2w, d, 0.0f);
    const f3 dw = normalize(normalize(dv));
    const f3 ow = mat_mul(cam.c2w, o, 1.0f);
    P.setF3(F_OX, slot, ow);
    P.setF3(F_DX, slot, dw);
    P.setF3(F_BX, slot, f3{1.0f, 1.0f, 1.0f});
    P.setF3(F_VX, slot, f3{0.0f, 0.0f, 0.0f});
    P.F(F_PX, slot) = px;
    P.F(F_PY, slot) = py;
    P.U(U_PIXEL, slot) = pixel;
    P.U(U_SAMPLE, slot) = s;
    P.U(U_BOUNCE, slot) = 0;
    P.U(U_PENDING, slot) = RBW_NONE;
    store_sampler(P, slot, smp);
}

// closest hit of the path ray (scene::intersect, scene.cpp:82-95)
template <int W, bool COUNT, bool LDS = false>
__device__ __forceinline__ void extend_path(const dscene& S, const dpaths& P, const dparams& prm, uint32_t slot,
                                            trace_counters& cnt) {
    const f3 o = P.F3(F_OX, slot), d = P.F3(F_DX, slot);
    const hitrec h = traverse<W, false, false, COUNT, LDS>(S, o, d, __builtin_inff(), prm.tri_test, RBW_NONE, cnt);
    P.U(U_HPRIM, slot) = h.prim;
    P.F(F_HT, slot) = h.t;
    if (uint32_t* dbg = debug_slot(P, prm, slot, (int)P.U(U_BOUNCE, slot))) dbg[0] = h.prim;
}

// sample_surface_interaction (integrator.cpp:329-404) up to the light and BSDF
// samples of next-event estimation, then path_integrator.cpp:35-43, for the path in
// `slot`.  GUIDED: vertices with a BSDF are handed back in gsi/guide_mat (to_guide)
// for the guided vertex stage instead.
template <bool GUIDED>
__device__ __forceinline__ void shade_path(const dscene& S, const dpaths& P, const dparams& prm, uint32_t slot,
                                           bool& alive, bool& need_connect, bool& to_guide, surf& gsi,
                                           uint32_t& guide_mat) {
    const uint32_t prim = P.U(U_HPRIM, slot);
    int bounce = (int)P.U(U_BOUNCE, slot);
    f3 value = P.F3(F_VX, slot);
    if (GUIDED) {
        // the radiance carried back along the previous vertex's sampled direction is
        // everything the path gathers from here on
        const uint32_t pend = P.U(U_PENDING, slot);
        if (pend != RBW_NONE) {
            const size_t ri = (size_t)slot * prm.g.rec_depth + pend;
            *reinterpret_cast<float4*>(prm.g.rec + ri * R_STRIDE + R_MX) = make_float4(value.x, value.y, value.z, 0.0f);
            prm.g.rec_flag[ri] = 2;
            P.U(U_PENDING, slot) = RBW_NONE;
        }
    }
    if (prim != RBW_MISS) {
        const f3 o_w = P.F3(F_OX, slot), d_w = P.F3(F_DX, slot);
        uint32_t eid;
        const surf si = hit_surface(S, prim, o_w, d_w, eid);
        const dentity E = S.ents[eid];
        f3 beta = P.F3(F_BX, slot);
        sampler_state smp = load_sampler(P, prm, slot);
        bool cont = true;
        // emission at the first vertex (integrator.cpp:350-351; no specular lobes exist)
        if (bounce == 0 && (E.flags & RBW_ENTITY_EMITTER)) value = value + beta * light_evaluate(E, si.normal, -d_w);
        // a vertex is guided below prm.g.depth; deeper ones shade as the plain estimator
        const bool guide_here = GUIDED && (uint32_t)bounce < prm.g.depth;
        if (GUIDED && bounce > 0 && (uint32_t)bounce <= prm.g.depth && (E.flags & RBW_ENTITY_EMITTER)) {
            // the previous vertex's guided continuation is the second technique of its next-event estimation
            const f3 Le = light_evaluate(E, si.normal, -d_w);
            if (!is_black(Le)) {
                const f3 mp = P.F3(F_MPX, slot), mn = P.F3(F_MNX, slot), mw = P.F3(F_MWX, slot);
                const float lp = light_pdf(S, E, mp, mn, mw) / (float)S.num_emitters;
                value = value + beta * Le * power_heuristic(P.F(F_MPDF, slot), lp);
            }
        }
        if (E.material == RBW_NONE) {
            // invisible entity: straight through, the bounce does not count (integrator.cpp:355-363)
            f3 no, nd;
            spawn_ray(si.point, si.normal, d_w, no, nd);
            P.setF3(F_OX, slot, no);
            P.setF3(F_DX, slot, nd);
            bounce -= 1;
        } else {
            const bsdf fn = make_bsdf(S, E.material, prim);
            uint32_t cflags = 0;
            if (!guide_here && fn.kind != LOBE_NONE) {
                // uniform_sample_one_emitter (integrator.cpp:101-161): light sample
                const uint32_t ne = S.num_emitters;
                const float u1 = smp.next();
                uint32_t which = (uint32_t)floorf(u1 * (float)ne);
                if (which > ne - 1) which = ne - 1;
                const float pick = 1.0f / (float)ne;
                const uint32_t lid = S.emitters[which];
                const dentity L = S.ents[lid];
                light_sample ls = sample_light(S, L, si.point, si.normal, smp.next2());
                ls.pdf = ls.pdf * pick;
                if (!is_black(ls.L) && ls.pdf > 0.0f) {
                    const f3 wi = world_to_local(si.sh, ls.wi);
                    const f3 wo = world_to_local(si.sh, si.wo);
                    f3 f = fn.evaluate(wo, wi);
                    const float fp = fn.pdf(wo, wi);
                    f = f * fabsf(dot(ls.wi, si.sh.z));
                    if (!is_black(f) && fp > 0.0f) {
                        f3 so, sd;
                        float slen;
                        spawn_ray_to(si.point, si.normal, ls.point, so, sd, slen);
                        const float w = power_heuristic(ls.pdf, fp);
                        const f3 contrib = f * ls.L * w / ls.pdf;
                        P.setF3(F_SOX, slot, so);
                        P.setF3(F_SDX, slot, sd);
                        P.F(F_SLEN, slot) = slen;
                        P.setF3(F_SFX, slot, contrib);
                        P.U(U_SENT, slot) = lid;
                        cflags |= C_SHADOW;
                    }
                }
                // BSDF sample for MIS (integrator.cpp:163-198)
                f3 mwi, mf;
                float mpdf = 0.0f;
                if (fn.sample(si, smp.next2(), mwi, mf, mpdf)) {
                    mf = mf * fabsf(dot(mwi, si.sh.z));
                    if (!is_black(mf) && mpdf > 0.0f) {
                        P.setF3(F_MPX, slot, si.point);
                        P.setF3(F_MNX, slot, si.normal);
                        P.setF3(F_MWX, slot, mwi);
                        P.setF3(F_MFX, slot, mf);
                        P.F(F_MPDF, slot) = mpdf;
                        cflags |= C_MIS;
                    }
                }
                if (cflags) {
                    P.setF3(F_CBX, slot, beta);
                    P.U(U_CFLAGS, slot) = cflags | ((uint32_t)bounce << 8);
                    need_connect = true;
                }
            }
            if (guide_here && fn.kind != LOBE_NONE) {
                to_guide = true;
                gsi = si;
                guide_mat = (S.tri_material && prim < S.num_tris && S.tri_material[prim] != RBW_NONE)
                                ? S.tri_material[prim] : E.material;
            } else if (guide_here) {
                cont = false;  // an empty lobe list ends the path (integrator.cpp:380-382)
            } else {
                // continuation sample (integrator.cpp:380-397)
                f3 wi, f;
                float pdf = 0.0f;
                const bool ok = fn.sample(si, smp.next2(), wi, f, pdf);
                if (!ok || is_black(f) || pdf == 0.0f) {
                    cont = false;
                } else {
                    beta = beta * (f * fabsf(dot(wi, si.sh.z)) / pdf);
                    f3 no, nd;
                    spawn_ray(si.point, si.normal, wi, no, nd);
                    P.setF3(F_OX, slot, no);
                    P.setF3(F_DX, slot, nd);
                }
            }
        }
        if (!to_guide) {
            if (cont) {
                // Russian roulette (path_integrator.cpp:35-43)
                const float mc = spectrum_max(beta * 1.0f);
                if (mc < prm.rr_threshold && bounce > 3) {
                    const float qq = gmax(0.05f, 1.0f - mc);
                    if (smp.next() < qq) cont = false;
                    else beta = beta / (1.0f - qq);
                }
            }
            bounce += 1;
            alive = cont && (uint32_t)bounce < prm.max_depth;
            P.U(U_BOUNCE, slot) = (uint32_t)bounce;
        }
        P.setF3(F_BX, slot, beta);
        P.setF3(F_VX, slot, value);
        store_sampler(P, slot, smp);
    }
}

// shadow ray, MIS ray and light pdf of next-event estimation for the path in `slot`;
// value += beta * L (integrator.cpp:101-201).  FAST: the scene's lights are planar and
// few (k_lights_planar | k_lights_few) and no debug hits are recorded, so both rays are
// occlusion queries (any-hit walks) and the closest-hit walks are not instantiated.  The
// MIS ray's vertex (point, normal, direction) is read again after its walk rather than
// held in registers across it.  WALKED (with FAST): a persistent walk ran both occlusion
// queries (C_LIT, U_MPRIM and C_MOCC hold their answers), and no walk is instantiated here.
template <int W, bool COUNT, bool LDS = false, bool FAST = false, bool WALKED = false>
__device__ __forceinline__ void connect_path(const dscene& S, const dpaths& P, const dparams& prm, uint32_t slot,
                                             trace_counters& cnt, uint32_t& nshadow, uint32_t& nmis,
                                             uint32_t& nlight, bool* lit = nullptr) {
    const uint32_t flags = P.U(U_CFLAGS, slot);
    const int bounce = (int)(flags >> 8);
    uint32_t* dbg = FAST ? nullptr : debug_slot(P, prm, slot, bounce);
    f3 L = f3{0.0f, 0.0f, 0.0f};
    if (WALKED && (flags & C_SHADOW)) {
        nshadow = 1;
        if (flags & C_LIT) L = L + P.F3(F_SFX, slot);
        if (lit) *lit = (flags & C_LIT) != 0u;
    } else if (flags & C_SHADOW) {
        nshadow = 1;
        const uint32_t lid = P.U(U_SENT, slot);
        const f3 so = P.F3(F_SOX, slot), sd = P.F3(F_SDX, slot);
        const float slen = P.F(F_SLEN, slot);
        // a planar light cannot be hit before the sample point on it, so the closest visible
        // hit is the light or none exactly when no other visible surface is in the way
        const bool any =
            FAST || prm.shadow_mode == RBW_SHADOW_ANY_HIT || (!dbg && (S.light_flags & k_lights_planar));
        hitrec h;
        if (any)
            h = traverse<W, true, true, COUNT, LDS>(S, so, sd, slen, prm.tri_test, lid, cnt);
        else
            h = traverse<W, true, false, COUNT, LDS>(S, so, sd, slen, prm.tri_test, RBW_NONE, cnt);
        if (dbg) dbg[1] = h.prim;
        bool unoccluded = h.prim == RBW_MISS;
        if (!unoccluded && !any) {
            const uint32_t he = h.prim < S.num_tris ? S.tri_entity[h.prim] : S.sphere_entity[h.prim - S.num_tris];
            unoccluded = he == lid;
        }
        if (unoccluded) L = L + P.F3(F_SFX, slot);
        if (lit) *lit = unoccluded;
    }
    if (flags & C_MIS) {
        nmis = 1;
        // find_emitter (integrator.cpp:49-82)
        uint32_t prim;
        if (WALKED) {
            prim = (flags & C_MOCC) ? RBW_MISS : P.U(U_MPRIM, slot);
        } else {
            const f3 mp = P.F3(F_MPX, slot), mn = P.F3(F_MNX, slot), mw = P.F3(F_MWX, slot);
            f3 eo, ed;
            spawn_ray(mp, mn, mw, eo, ed);
            hitrec h;
            if (FAST || (!dbg && (S.light_flags & k_lights_few))) {
                // the closest visible hit is an emitter exactly when the nearest visible
                // emitter triangle along the ray has no other visible surface before it: test
                // the few emitter triangles, then walk the tree only for rays that reach one,
                // as an occlusion query up to it
                h = nearest_emitter_hit(S, S.emitters, S.num_emitters, eo, ed, prm.tri_test);
                if (h.prim != RBW_MISS) {
                    const uint32_t le = S.tri_entity[h.prim];
                    const local_ray lr = make_local_ray(eo, ed);
                    const hitrec o =
                        traverse<W, true, true, COUNT, LDS>(S, eo, ed, h.t * lr.L1, prm.tri_test, le, cnt);
                    if (o.prim != RBW_MISS) h.prim = RBW_MISS;
                }
            } else {
                h = traverse<W, true, false, COUNT, LDS>(S, eo, ed, __builtin_inff(), prm.tri_test, RBW_NONE, cnt);
            }
            prim = h.prim;
        }
        if (dbg) dbg[2] = prim;
        if (prim != RBW_MISS) {
            // the vertex again from the path state (a compiler barrier keeps the loads here)
            __asm__ volatile("" ::: "memory");
            const f3 mp = P.F3(F_MPX, slot), mn = P.F3(F_MNX, slot), mw = P.F3(F_MWX, slot);
            f3 eo, ed;
            spawn_ray(mp, mn, mw, eo, ed);
            uint32_t he;
            const surf hs = hit_surface(S, prim, eo, ed, he);
            const dentity E = S.ents[he];
            if (E.flags & RBW_ENTITY_EMITTER) {
                nlight = 1;
                const float pick = 1.0f / (float)S.num_emitters;
                const f3 Le = light_evaluate(E, hs.normal, -mw);
                const float lp = light_pdf(S, E, mp, mn, mw) * pick;
                if (!is_black(Le) && lp > 0.0f) {
                    const float mpdf = P.F(F_MPDF, slot);
                    const float w = power_heuristic(mpdf, lp);
                    L = L + P.F3(F_MFX, slot) * Le * w / mpdf;
                }
            }
        }
    }
    const f3 beta = P.F3(F_CBX, slot);
    P.setF3(F_VX, slot, P.F3(F_VX, slot) + beta * L);
}

// ------------------------------------------------------------------ film

// Film sums are 64-bit fixed point (2^-32 radiance units): integer sums do not
// depend on the order in which samples arrive, so a render is reproducible bit
// for bit whether its samples are accumulated in sample order (single render)
// or as paths retire (batched jobs).  |value| is clamped to 2^30; a non-finite
// sample marks its channel, which then resolves to NaN (the reference's film
// sum turns NaN there, film.cpp:138).  A pixel's sum holds |sum| < 2^31
// radiance units; an add that would leave that range marks the channel
// (FILM_OVER_POS / FILM_OVER_NEG), which then resolves to +inf / -inf instead of
// a wrapped value.
struct dfilm {
    unsigned long long* sum;  // 3 per pixel
    uint32_t* weight;         // samples per pixel (box filter weight 1)
    uint32_t* nan;            // per pixel: channel flags (non-finite sample, sum overflow)
};

constexpr uint32_t FILM_OVER_POS = 8u;    // << channel
constexpr uint32_t FILM_OVER_NEG = 64u;   // << channel

__device__ __forceinline__ long long film_fix(float v) {
    const double x = fmin(fmax((double)v, -1073741824.0), 1073741824.0) * 4294967296.0;
    return __double2ll_rn(x);
}

// the overflow flag of adding q to a sum that held `old`
__device__ __forceinline__ uint32_t film_overflow(long long old, long long q, int ch) {
    if (q > 0 && old > __LONG_LONG_MAX__ - q) return FILM_OVER_POS << ch;
    if (q < 0 && old < -__LONG_LONG_MAX__ - 1 - q) return FILM_OVER_NEG << ch;
    return 0u;
}

// one channel's fixed-point add with the overflow check (the sum is flagged, not wrapped silently)
__device__ __forceinline__ void film_add_fixed(const dfilm& film, size_t p, int ch, long long q) {
    const long long old = (long long)atomicAdd(&film.sum[3 * p + ch], (unsigned long long)q);
    const uint32_t over = film_overflow(old, q, ch);
    if (over) atomicOr(&film.nan[p], over);
}

// film_tile::add_sample footprint (film.cpp:57-92, box filter radius 0.5) clamped to the pixel bounds
__device__ __forceinline__ void footprint(const dcamera& cam, float px, float py, int32_t& x0, int32_t& y0,
                                          int32_t& x1, int32_t& y1) {
    const float dx = px - 0.5f, dy = py - 0.5f;
    x0 = (int32_t)ceilf(dx - 0.5f);
    y0 = (int32_t)ceilf(dy - 0.5f);
    x1 = (int32_t)floorf(dx + 0.5f) + 1;
    y1 = (int32_t)floorf(dy + 0.5f) + 1;
    x0 = x0 > cam.bx0 ? x0 : cam.bx0;
    y0 = y0 > cam.by0 ? y0 : cam.by0;
    x1 = x1 < cam.bx1 ? x1 : cam.bx1;
    y1 = y1 < cam.by1 ? y1 : cam.by1;
}

// adds a sample to every pixel of [x0,x1) x [y0,y1) with atomics
__device__ __forceinline__ void film_add_span(const dcamera& cam, const dfilm& film, int32_t x0, int32_t y0, int32_t x1,
                                              int32_t y1, f3 v) {
    const float c[3] = {v.x, v.y, v.z};
    long long q[3];
    uint32_t bad = 0;
    for (int ch = 0; ch < 3; ch++) {
        q[ch] = isfinite(c[ch]) ? film_fix(c[ch]) : 0;
        if (!isfinite(c[ch])) bad |= 1u << ch;
    }
    for (int32_t yy = y0; yy < y1; yy++)
        for (int32_t xx = x0; xx < x1; xx++) {
            const size_t p = (size_t)yy * cam.width + (size_t)xx;
            for (int ch = 0; ch < 3; ch++)
                if (q[ch]) film_add_fixed(film, p, ch, q[ch]);
            atomicAdd(&film.weight[p], 1u);
            if (bad) atomicOr(&film.nan[p], bad);
        }
}

// One sample of own pixel (ox, oy) into the film with atomics.  Every sample's box
// footprint holds its own pixel; with count_own = false that pixel's weight is
// left out (the caller knows it: spp per pixel) and only spills into neighbours
// are counted.
__device__ __forceinline__ void film_add_sample(const dcamera& cam, const dfilm& film, int32_t ox, int32_t oy, float px,
                                                float py, f3 v, bool count_own) {
    int32_t x0, y0, x1, y1;
    footprint(cam, px, py, x0, y0, x1, y1);
    const float c[3] = {v.x, v.y, v.z};
    long long q[3];
    uint32_t bad = 0;
    for (int ch = 0; ch < 3; ch++) {
        q[ch] = isfinite(c[ch]) ? film_fix(c[ch]) : 0;
        if (!isfinite(c[ch])) bad |= 1u << ch;
    }
    for (int32_t yy = y0; yy < y1; yy++)
        for (int32_t xx = x0; xx < x1; xx++) {
            const size_t p = (size_t)yy * cam.width + (size_t)xx;
            for (int ch = 0; ch < 3; ch++)
                if (q[ch]) film_add_fixed(film, p, ch, q[ch]);
            if (count_own || xx != ox || yy != oy) atomicAdd(&film.weight[p], 1u);
            if (bad) atomicOr(&film.nan[p], bad);
        }
}

// pixel::spectrum (film.cpp:20-25): sum / (base_weight + weight); pixels outside the bounds are 0
__device__ __forceinline__ void film_resolve_pixel(const dcamera& cam, const dfilm& film, uint32_t p, float* rgb,
                                                   uint32_t base_weight = 0) {
    const int32_t x = (int32_t)(p % cam.width), y = (int32_t)(p / cam.width);
    const uint32_t w = base_weight + film.weight[p];
    const bool in = x >= cam.bx0 && x < cam.bx1 && y >= cam.by0 && y < cam.by1 && w != 0;
    const uint32_t bad = in ? film.nan[p] : 0u;
    for (int ch = 0; ch < 3; ch++) {
        float v = 0.0f;
        if (in) v = (float)((double)(long long)film.sum[3 * p + ch] * 2.3283064365386963e-10 / (double)w);
        if (bad & (FILM_OVER_POS << ch)) v = __builtin_huge_valf();
        if (bad & (FILM_OVER_NEG << ch)) v = -__builtin_huge_valf();
        if (bad & (1u << ch)) v = __builtin_nanf("");
        rgb[3 * p + ch] = v;
    }
}

// host: the device camera record of a camera descriptor
inline dcamera make_dcamera(const rbw_camera_desc* c) {
    dcamera d;
    std::memcpy(d.r2c, c->raster_to_camera, sizeof(d.r2c));
    std::memcpy(d.c2w, c->camera_to_world, sizeof(d.c2w));
    d.width = c->width;
    d.height = c->height;
    d.bx0 = c->bound_min[0];
    d.by0 = c->bound_min[1];
    d.bx1 = c->bound_max[0];
    d.by1 = c->bound_max[1];
    d.lens_radius = c->lens_radius;
    d.focus = c->focus_distance;
    return d;
}

}  // namespace rbw
