// Wavefront path tracer for MI355X (gfx950) behind the C ABI of include/rbw.h.
//
// One render = passes of (pixel, sample) paths held in SoA state in HBM.
// Each bounce runs three kernels over compacted queues:
//   extend  : closest-hit BVH traversal of the path rays
//   shade   : surface interaction, emission, material, light and BSDF samples
//             for next-event estimation, continuation sample, Russian roulette
//   connect : the shadow ray and the MIS BSDF ray of next-event estimation,
//             the light's own pdf cast, and the MIS-weighted contribution
// then film kernels accumulate the pass into the film in sample order.
// The arithmetic follows rainbow-cpu's path integrator
// (integrators/path_integrator.cpp:15-47, integrators/integrator.cpp:101-404).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rbw.h"
#include "bvh.hpp"
#include "device.hpp"
#include "guiding.hpp"
#include "shading.hpp"

using namespace rbw;

#define RBW_STACK 64

// ------------------------------------------------------------------ traversal

// acceptance bound of the next candidate: inside the same instanced mesh the
// reference compares against the mesh-local length directly (mesh.cpp:431);
// any other candidate sees the world length mapped into its local space.
__device__ __forceinline__ float thr_from(uint32_t best_inst, uint32_t cand_inst, float best_t, float len_w, float L0) {
    if (cand_inst != RBW_NONE && cand_inst == best_inst) return best_t;
    return len_w * L0;
}


template <bool VISIBLE_ONLY, bool ANY_HIT, bool COUNT>
__device__ __forceinline__ hitrec traverse(const dscene& S, f3 o_w, f3 d_w, float len_w, uint32_t tri_test,
                                           uint32_t target_entity, trace_counters& cnt) {
    const local_ray lr = make_local_ray(o_w, d_w);
    wt_setup wt;
    if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
    hitrec best;
    best.prim = RBW_MISS;
    best.t = 0.0f;
    best.b1 = best.b2 = 0.0f;
    uint32_t best_inst_entity = RBW_NONE;  // entity of the best hit when it belongs to an instanced mesh
    float thr_local = len_w * lr.L0;       // current acceptance bound in local units
    uint32_t stack[RBW_STACK];
    int sp = 0;
    int32_t cur = 0;
    const f3 o = lr.o;
    const f3 inv = lr.inv;
    const f3 oi = f3{o.x * inv.x, o.y * inv.y, o.z * inv.z};
    for (;;) {
        if (cur >= 0) {
            const float4* n = S.nodes + 4 * (size_t)cur;
            const float4 a = n[0], b = n[1], c = n[2];
            const int4 ch = *reinterpret_cast<const int4*>(n + 3);
            if (COUNT) cnt.nodes++;
            const float tmax = thr_local * 1.0000002f;
            // left
            // slab distances (lo - o) / d as fma(lo, 1/d, -o/d); the bounds below are widened
            // by a few ulps so that rounding never culls a box that holds an accepted hit
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
                stack[sp++] = (uint32_t)(lfirst ? ch.y : ch.x);
                continue;
            } else if (hl) {
                cur = ch.x;
                continue;
            } else if (hr) {
                cur = ch.y;
                continue;
            }
        } else if (cur != k_empty_child) {
            const uint32_t first = (uint32_t)cur & 0x07FFFFFFu;
            const uint32_t count = (((uint32_t)cur >> 27) & 15u) + 1u;
            for (uint32_t k = 0; k < count; k++) {
                const uint32_t slot = first + k;
                const float4 pa = S.prim_a[slot];
                const uint32_t word = __float_as_uint(pa.w);
                if (VISIBLE_ONLY && (word & k_prim_invisible)) continue;
                if (COUNT) cnt.prims++;
                const uint32_t id = word & k_prim_id_mask;
                float t = 0.0f, b1 = 0.0f, b2 = 0.0f;
                bool hit;
                uint32_t inst_entity = RBW_NONE;
                if (word & k_prim_sphere) {
                    const float4 pb = S.prim_b[slot];
                    const float thr = thr_from(best_inst_entity, RBW_NONE, best.t, len_w, lr.L0);
                    const f3 ol = translate(o_w, f3{-pa.x, -pa.y, -pa.z});
                    hit = sphere_test(pb.x, ol, lr.d, thr, t);
                } else {
                    const float4 pb = S.prim_b[slot], pc = S.prim_c[slot];
                    if (word & k_prim_instanced) inst_entity = S.tri_entity[id];
                    const float thr = thr_from(best_inst_entity, inst_entity, best.t, len_w, lr.L0);
                    const f3 p0 = f3{pa.x, pa.y, pa.z}, p1 = f3{pb.x, pb.y, pb.z}, p2 = f3{pc.x, pc.y, pc.z};
                    if (tri_test == RBW_TRI_WATERTIGHT)
                        hit = wt_test(p0, p1, p2, o, lr.d, thr, wt.kz, wt.Sx, wt.Sy, wt.Sz, t, b1, b2);
                    else
                        hit = mt_test(p0, p1, p2, o, lr.d, thr, t, b1, b2);
                }
                if (hit) {
                    if (ANY_HIT) {
                        const uint32_t ent =
                            (word & k_prim_sphere) ? S.sphere_entity[id - S.num_tris] : S.tri_entity[id];
                        if (ent != target_entity) {
                            best.prim = id;
                            best.t = t;
                            return best;
                        }
                        continue;
                    }
                    best.prim = id;
                    best.t = t;
                    best.b1 = b1;
                    best.b2 = b2;
                    best_inst_entity = inst_entity;
                    len_w = t * lr.L1;  // entity.cpp:384
                    thr_local = (inst_entity != RBW_NONE) ? t : len_w * lr.L0;
                }
            }
        }
        if (sp == 0) break;
        cur = (int32_t)stack[--sp];
    }
    return best;
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
    F_CBX, F_CBY, F_CBZ,                         // beta at the vertex
    F_COUNT
};
enum uidx : uint32_t { U_PIXEL, U_SAMPLE, U_BOUNCE, U_HPRIM, U_CFLAGS, U_SENT, U_PENDING, U_COUNT };

constexpr uint32_t C_SHADOW = 1u, C_MIS = 2u;

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
    float* rec;                  // R_COUNT * (cap * max_depth) (SoA)
    uint32_t* rec_flag;          // 0 none, 1 sampled, 2 closed
    uint32_t cap;
    // training batch
    float* t_in;                 // 9 per sample
    float* t_dir;                // 3 per sample
    float* t_c;                  // radiance / pdf
    uint32_t* t_count;
    uint32_t t_cap;
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
    dguide g;
};

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
    __device__ float next() {
        if (replay) return state < end ? replay[state++] : 0.0f;
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        const uint32_t r = (xs >> rot) | (xs << ((~rot + 1u) & 31));
        const float x = (float)r * 2.3283064365386963e-10f;
        return (x < k_one_minus_epsilon) ? x : k_one_minus_epsilon;
    }
    __device__ f2 next2() {
        const float a = next();
        const float b = next();
        return f2{a, b};
    }
};

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
    }
    return s;
}

__device__ __forceinline__ uint32_t* debug_slot(const dpaths& P, const dparams& prm, uint32_t slot, int bounce) {
    if (!prm.debug || bounce < 0 || (uint32_t)bounce >= prm.max_depth) return nullptr;
    const uint64_t sid = (uint64_t)P.U(U_PIXEL, slot) * prm.spp + P.U(U_SAMPLE, slot);
    return prm.debug + (sid * prm.max_depth + (uint32_t)bounce) * 3;
}

// ------------------------------------------------------------------ kernels

// sampler_integrator.cpp:98-127 + perspective_camera.cpp:78-99: one camera sample per slot
__global__ void k_raygen(dscene S, dpaths P, dparams prm, dcamera cam, uint32_t n, uint32_t s0, uint32_t* queue,
                         uint32_t* queue_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bw = (uint32_t)(cam.bx1 - cam.bx0), bh = (uint32_t)(cam.by1 - cam.by0);
    const uint32_t npix = bw * bh;
    const uint32_t li = i % npix, sl = i / npix;
    const int32_t x = cam.bx0 + (int32_t)(li % bw), y = cam.by0 + (int32_t)(li / bw);
    const uint32_t pixel = (uint32_t)y * cam.width + (uint32_t)x;
    const uint32_t s = s0 + sl;
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
    P.setF3(F_OX, i, ow);
    P.setF3(F_DX, i, dw);
    P.setF3(F_BX, i, f3{1.0f, 1.0f, 1.0f});
    P.setF3(F_VX, i, f3{0.0f, 0.0f, 0.0f});
    P.F(F_PX, i) = px;
    P.F(F_PY, i) = py;
    P.U(U_PIXEL, i) = pixel;
    P.U(U_SAMPLE, i) = s;
    P.U(U_BOUNCE, i) = 0;
    P.U(U_PENDING, i) = RBW_NONE;
    P.rng[i] = smp.state;
    queue[i] = i;
    if (i == 0) *queue_count = n;
}

template <bool COUNT>
__global__ void k_extend(dscene S, dpaths P, dparams prm, const uint32_t* queue, const uint32_t* queue_count,
                         unsigned long long* counters) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *queue_count;
    trace_counters cnt{0, 0};
    if (q < count) {
        const uint32_t slot = queue[q];
        const f3 o = P.F3(F_OX, slot), d = P.F3(F_DX, slot);
        const hitrec h = traverse<false, false, COUNT>(S, o, d, __builtin_inff(), prm.tri_test, RBW_NONE, cnt);
        P.U(U_HPRIM, slot) = h.prim;
        P.F(F_HT, slot) = h.t;
        if (uint32_t* dbg = debug_slot(P, prm, slot, (int)P.U(U_BOUNCE, slot))) dbg[0] = h.prim;
    }
    if (COUNT) {
        wave_add64(counters + 0, cnt.nodes);
        wave_add64(counters + 1, cnt.prims);
    }
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

// sample_surface_interaction (integrator.cpp:329-404) up to the light and BSDF
// samples of next-event estimation, then path_integrator.cpp:35-43.  GUIDED:
// vertices with a BSDF go to the guide queue instead; k_shade_guided samples
// them once the MLP has produced their mixtures.
template <bool GUIDED>
__global__ void k_shade(dscene S, dpaths P, dparams prm, const uint32_t* qin, const uint32_t* qin_count,
                        uint32_t* qout, uint32_t* qout_count, uint32_t* cq, uint32_t* cq_count) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *qin_count;
    bool alive = false, need_connect = false, to_guide = false;
    uint32_t slot = 0, guide_mat = 0;
    surf gsi;
    if (q < count) {
        slot = qin[q];
        const uint32_t prim = P.U(U_HPRIM, slot);
        int bounce = (int)P.U(U_BOUNCE, slot);
        f3 value = P.F3(F_VX, slot);
        if (GUIDED) {
            // the radiance carried back along the previous vertex's sampled direction is
            // everything the path gathers from here on
            const uint32_t pend = P.U(U_PENDING, slot);
            if (pend != RBW_NONE) {
                const size_t ri = (size_t)slot * prm.max_depth + pend;
                const size_t rc = (size_t)P.cap * prm.max_depth;
                prm.g.rec[R_MX * rc + ri] = value.x;
                prm.g.rec[R_MY * rc + ri] = value.y;
                prm.g.rec[R_MZ * rc + ri] = value.z;
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
            if (GUIDED && bounce > 0 && (E.flags & RBW_ENTITY_EMITTER)) {
                // the guided continuation is the second technique of next-event estimation
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
                if (!GUIDED && fn.kind != LOBE_NONE) {
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
                if (GUIDED && fn.kind != LOBE_NONE) {
                    to_guide = true;
                    gsi = si;
                    guide_mat = (S.tri_material && prim < S.num_tris && S.tri_material[prim] != RBW_NONE)
                                    ? S.tri_material[prim] : E.material;
                } else if (GUIDED) {
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
            P.rng[slot] = smp.state;
        }
    }
    const uint32_t a = wave_append(qout_count, alive);
    if (alive) qout[a] = slot;
    const uint32_t c = wave_append(cq_count, need_connect);
    if (need_connect) cq[c] = slot;
    if (GUIDED) {
        const uint32_t g = wave_append(prm.g.gq_count, to_guide);
        if (to_guide) {
            const dguide& G = prm.g;
            const size_t cap = G.cap;
            const float v[G_COUNT] = {gsi.point.x, gsi.point.y, gsi.point.z, gsi.normal.x, gsi.normal.y, gsi.normal.z,
                                      gsi.wo.x, gsi.wo.y, gsi.wo.z, gsi.sh.x.x, gsi.sh.x.y, gsi.sh.x.z,
                                      gsi.sh.y.x, gsi.sh.y.y, gsi.sh.y.z, gsi.sh.z.x, gsi.sh.z.y, gsi.sh.z.z};
#pragma unroll
            for (int k = 0; k < G_COUNT; k++) G.surf[k * cap + g] = v[k];
            G.surf_mat[g] = guide_mat;
            G.gq[g] = slot;
            // MLP query: position in the scene box, shading normal and wo mapped to [0,1]
            const f3 pn = (gsi.point - G.smin) * G.sinv;
            const float qv[9] = {gclamp(pn.x, 0.0f, 1.0f), gclamp(pn.y, 0.0f, 1.0f), gclamp(pn.z, 0.0f, 1.0f),
                                 0.5f * (gsi.sh.z.x + 1.0f), 0.5f * (gsi.sh.z.y + 1.0f), 0.5f * (gsi.sh.z.z + 1.0f),
                                 0.5f * (gsi.wo.x + 1.0f), 0.5f * (gsi.wo.y + 1.0f), 0.5f * (gsi.wo.z + 1.0f)};
#pragma unroll
            for (int k = 0; k < 9; k++) G.query[(size_t)g * 9 + k] = qv[k];
        }
    }
}

// the vMF-mixture head of the guiding MLP (same parameterisation as the training loss)
struct vmf_mixture {
    f3 mu[guide::N_LOBES];
    float kappa[guide::N_LOBES], w[guide::N_LOBES], norm[guide::N_LOBES];
    __device__ void load(const float* o) {
        float amax = -3.4e38f;
        for (int j = 0; j < guide::N_LOBES; j++) amax = fmaxf(amax, o[5 * j + 4]);
        float ws = 0.0f;
        for (int j = 0; j < guide::N_LOBES; j++) {
            w[j] = expf(o[5 * j + 4] - amax);
            ws += w[j];
            const f3 m = f3{o[5 * j], o[5 * j + 1], o[5 * j + 2]};
            const float nm = fmaxf(sqrtf(m.x * m.x + m.y * m.y + m.z * m.z), 1e-12f);
            mu[j] = f3{m.x / nm, m.y / nm, m.z / nm};
            const float x = o[5 * j + 3];
            const float sp = x > 20.0f ? x : log1pf(expf(x));
            kappa[j] = fminf(sp + guide::KAPPA_MIN, guide::KAPPA_MAX);
            norm[j] = kappa[j] / (2.0f * 3.14159265f * (-expm1f(-2.0f * kappa[j])));
        }
        for (int j = 0; j < guide::N_LOBES; j++) w[j] /= ws;
    }
    __device__ float pdf(f3 d) const {
        float q = 0.0f;
        for (int j = 0; j < guide::N_LOBES; j++)
            q += w[j] * norm[j] * expf(kappa[j] * ((mu[j].x * d.x + mu[j].y * d.y + mu[j].z * d.z) - 1.0f));
        return q;
    }
    // pick a lobe with u.x, then Jakob's vMF sampler around its mean
    __device__ f3 sample(f2 u) const {
        int j = 0;
        float acc = 0.0f;
        for (; j < guide::N_LOBES - 1; j++) {
            if (u.x < acc + w[j]) break;
            acc += w[j];
        }
        const float ux = fminf(fmaxf((u.x - acc) / fmaxf(w[j], 1e-20f), 0.0f), 0.99999994f);
        const float k = kappa[j];
        const float W = 1.0f + logf(ux + (1.0f - ux) * expf(-2.0f * k)) / k;
        const float sr = sqrtf(fmaxf(0.0f, 1.0f - W * W));
        const float phi = 2.0f * 3.14159265f * u.y;
        const frame fr = frame_from_normal(mu[j]);
        return local_to_world(fr, f3{sr * cosf(phi), sr * sinf(phi), W});
    }
};

// sample from the guided technique: the vMF mixture with probability alpha, else
// the BSDF; returns false when no direction results
__device__ __forceinline__ bool guided_sample(const bsdf& fn, const surf& si, const vmf_mixture& mix, float alpha,
                                              float sel, f2 u, f3& wi) {
    if (sel < alpha) {
        wi = mix.sample(u);
        return true;
    }
    f3 f0;
    float p0;
    return fn.sample(si, u, wi, f0, p0);
}

// Guided vertex: next-event estimation whose second technique is the path's own
// continuation, drawn from the guided density p_G = alpha q + (1 - alpha) p_bsdf.
// The light sample is weighted by the power heuristic against p_G (the structure
// of integrator.cpp:101-161 with the BSDF technique replaced); the continuation
// ray, when it reaches an emitter, adds the emission weighted against the light
// pdf (k_shade, bounce > 0).  At the last vertex there is no continuation and the
// light sample takes weight 1.  Russian roulette as path_integrator.cpp:35-43.
// The continuation direction becomes a training record, closed at the next vertex:
// its target is everything the continuation brings back (emission weighted by its
// MIS weight, plus all later bounces), so the guide learns the radiance that light
// sampling does not already cover.
__global__ void k_shade_guided(dscene S, dpaths P, dparams prm, uint32_t* qout, uint32_t* qout_count, uint32_t* cq,
                               uint32_t* cq_count) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const dguide& G = prm.g;
    const uint32_t count = *G.gq_count;
    bool alive = false, need_connect = false;
    uint32_t slot = 0;
    if (g < count) {
        slot = G.gq[g];
        const size_t cap = G.cap;
        float v[G_COUNT];
#pragma unroll
        for (int k = 0; k < G_COUNT; k++) v[k] = G.surf[k * cap + g];
        surf si;
        si.point = f3{v[G_PX], v[G_PY], v[G_PZ]};
        si.normal = f3{v[G_NX], v[G_NY], v[G_NZ]};
        si.wo = f3{v[G_WX], v[G_WY], v[G_WZ]};
        si.sh.x = f3{v[G_XX], v[G_XY], v[G_XZ]};
        si.sh.y = f3{v[G_YX], v[G_YY], v[G_YZ]};
        si.sh.z = f3{v[G_ZX], v[G_ZY], v[G_ZZ]};
        const bsdf fn = make_bsdf(S, G.surf_mat[g], RBW_NONE);
        vmf_mixture mix;
        mix.load(G.mix + (size_t)g * guide::N_OUT_PAD);
        sampler_state smp = load_sampler(P, prm, slot);
        int bounce = (int)P.U(U_BOUNCE, slot);
        f3 beta = P.F3(F_BX, slot);
        const f3 wo_s = world_to_local(si.sh, si.wo);
        const float a = G.alpha;
        const bool last = (uint32_t)bounce + 1 >= prm.max_depth;
        auto pdf_g = [&](f3 w, f3 w_s) { return a * mix.pdf(w) + (1.0f - a) * fn.pdf(wo_s, w_s); };
        // light sample
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
            f3 f = fn.evaluate(wo_s, wi);
            f = f * fabsf(dot(ls.wi, si.sh.z));
            if (!is_black(f)) {
                f3 so, sd;
                float slen;
                spawn_ray_to(si.point, si.normal, ls.point, so, sd, slen);
                const float w = last ? 1.0f : power_heuristic(ls.pdf, pdf_g(ls.wi, wi));
                P.setF3(F_SOX, slot, so);
                P.setF3(F_SDX, slot, sd);
                P.F(F_SLEN, slot) = slen;
                P.setF3(F_SFX, slot, f * ls.L * w / ls.pdf);
                P.U(U_SENT, slot) = lid;
                P.setF3(F_CBX, slot, beta);
                P.U(U_CFLAGS, slot) = C_SHADOW | ((uint32_t)bounce << 8);
                need_connect = true;
            }
        }
        // continuation from the guided technique
        bool cont = !last;
        f3 wi = f3{0.0f, 0.0f, 0.0f};
        float pdf = 0.0f;
        if (cont) {
            const float sel = smp.next();
            const f2 u = smp.next2();
            cont = guided_sample(fn, si, mix, a, sel, u, wi);
        }
        if (cont) {
            const f3 wi_s = world_to_local(si.sh, wi);
            const f3 f = fn.evaluate(wo_s, wi_s);
            pdf = pdf_g(wi, wi_s);
            if (is_black(f) || !(pdf > 0.0f)) {
                cont = false;
            } else {
                beta = beta * (f * fabsf(dot(wi, si.sh.z)) / pdf);
                f3 no, nd;
                spawn_ray(si.point, si.normal, wi, no, nd);
                P.setF3(F_OX, slot, no);
                P.setF3(F_DX, slot, nd);
                // what the next vertex needs to weight emission it finds against light sampling
                P.setF3(F_MPX, slot, si.point);
                P.setF3(F_MNX, slot, si.normal);
                P.setF3(F_MWX, slot, wi);
                P.F(F_MPDF, slot) = pdf;
            }
        }
        if (cont) {
            const float mc = spectrum_max(beta * 1.0f);
            if (mc < prm.rr_threshold && bounce > 3) {
                const float qq = gmax(0.05f, 1.0f - mc);
                if (smp.next() < qq) cont = false;
                else beta = beta / (1.0f - qq);
            }
        }
        const int b0 = bounce;
        bounce += 1;
        alive = cont;
        if (alive && G.train) {
            const size_t ri = (size_t)slot * prm.max_depth + (uint32_t)b0;
            const size_t rc = (size_t)P.cap * prm.max_depth;
            for (int k = 0; k < 9; k++) G.rec[(R_Q0 + k) * rc + ri] = G.query[(size_t)g * 9 + k];
            G.rec[R_DX * rc + ri] = wi.x;
            G.rec[R_DY * rc + ri] = wi.y;
            G.rec[R_DZ * rc + ri] = wi.z;
            G.rec[R_PDF * rc + ri] = pdf;
            G.rec[R_BX * rc + ri] = beta.x;
            G.rec[R_BY * rc + ri] = beta.y;
            G.rec[R_BZ * rc + ri] = beta.z;
            G.rec_flag[ri] = 1;
            P.U(U_PENDING, slot) = (uint32_t)b0;
        }
        P.setF3(F_BX, slot, beta);
        P.U(U_BOUNCE, slot) = (uint32_t)bounce;
        P.rng[slot] = smp.state;
    }
    const uint32_t ai = wave_append(qout_count, alive);
    if (alive) qout[ai] = slot;
    const uint32_t ci = wave_append(cq_count, need_connect);
    if (need_connect) cq[ci] = slot;
}

// after a pass: closed training records -> (query, direction, radiance / pdf)
__global__ void k_collect(dpaths P, dparams prm, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const dguide& G = prm.g;
    const size_t rc = (size_t)P.cap * prm.max_depth;
    bool take = false;
    float c = 0.0f;
    size_t ri = 0;
    if (i < n * prm.max_depth) {
        const uint32_t slot = i / prm.max_depth;
        ri = i;
        if (G.rec_flag[ri] == 2) {
            const f3 vf = P.F3(F_VX, slot);
            const f3 m = f3{G.rec[R_MX * rc + ri], G.rec[R_MY * rc + ri], G.rec[R_MZ * rc + ri]};
            const f3 b = f3{G.rec[R_BX * rc + ri], G.rec[R_BY * rc + ri], G.rec[R_BZ * rc + ri]};
            const f3 d = vf - m;
            const float Lx = b.x > 0.0f ? d.x / b.x : 0.0f, Ly = b.y > 0.0f ? d.y / b.y : 0.0f,
                        Lz = b.z > 0.0f ? d.z / b.z : 0.0f;
            const float lum = 0.2126f * Lx + 0.7152f * Ly + 0.0722f * Lz;
            const float pdf = G.rec[R_PDF * rc + ri];
            c = pdf > 0.0f ? lum / pdf : 0.0f;
            take = c > 0.0f && isfinite(c);
        }
        G.rec_flag[ri] = 0;
    }
    const uint32_t k = wave_append(G.t_count, take);
    if (take && k < G.t_cap) {
        for (int j = 0; j < 9; j++) G.t_in[(size_t)k * 9 + j] = G.rec[(R_Q0 + j) * rc + ri];
        G.t_dir[(size_t)k * 3 + 0] = G.rec[R_DX * rc + ri];
        G.t_dir[(size_t)k * 3 + 1] = G.rec[R_DY * rc + ri];
        G.t_dir[(size_t)k * 3 + 2] = G.rec[R_DZ * rc + ri];
        G.t_c[k] = c;
    }
}

// shadow ray, MIS ray and light pdf of next-event estimation; value += beta * L
template <bool COUNT>
__global__ void k_connect(dscene S, dpaths P, dparams prm, const uint32_t* cq, const uint32_t* cq_count,
                          unsigned long long* counters) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *cq_count;
    trace_counters cnt{0, 0};
    uint32_t nshadow = 0, nmis = 0, nlight = 0;
    if (q < count) {
        const uint32_t slot = cq[q];
        const uint32_t flags = P.U(U_CFLAGS, slot);
        const int bounce = (int)(flags >> 8);
        uint32_t* dbg = debug_slot(P, prm, slot, bounce);
        f3 L = f3{0.0f, 0.0f, 0.0f};
        if (flags & C_SHADOW) {
            nshadow = 1;
            const uint32_t lid = P.U(U_SENT, slot);
            const f3 so = P.F3(F_SOX, slot), sd = P.F3(F_SDX, slot);
            const float slen = P.F(F_SLEN, slot);
            hitrec h;
            if (prm.shadow_mode == RBW_SHADOW_ANY_HIT)
                h = traverse<true, true, COUNT>(S, so, sd, slen, prm.tri_test, lid, cnt);
            else
                h = traverse<true, false, COUNT>(S, so, sd, slen, prm.tri_test, RBW_NONE, cnt);
            if (dbg) dbg[1] = h.prim;
            bool unoccluded = h.prim == RBW_MISS;
            if (!unoccluded && prm.shadow_mode != RBW_SHADOW_ANY_HIT) {
                const uint32_t he = h.prim < S.num_tris ? S.tri_entity[h.prim] : S.sphere_entity[h.prim - S.num_tris];
                unoccluded = he == lid;
            }
            if (unoccluded) L = L + P.F3(F_SFX, slot);
        }
        if (flags & C_MIS) {
            nmis = 1;
            // find_emitter (integrator.cpp:49-82)
            const f3 mp = P.F3(F_MPX, slot), mn = P.F3(F_MNX, slot), mw = P.F3(F_MWX, slot);
            f3 eo, ed;
            spawn_ray(mp, mn, mw, eo, ed);
            const hitrec h = traverse<true, false, COUNT>(S, eo, ed, __builtin_inff(), prm.tri_test, RBW_NONE, cnt);
            if (dbg) dbg[2] = h.prim;
            if (h.prim != RBW_MISS) {
                uint32_t he;
                const surf hs = hit_surface(S, h.prim, eo, ed, he);
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
    wave_add64(counters + 2, nshadow);
    wave_add64(counters + 3, nmis);
    wave_add64(counters + 4, nlight);
    if (COUNT) {
        wave_add64(counters + 5, cnt.nodes);
        wave_add64(counters + 6, cnt.prims);
    }
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

// one thread per pixel: the pass's samples whose footprint is exactly the own pixel, in sample order
__global__ void k_film_own(dpaths P, dcamera cam, uint32_t k, float* sum, float* weight) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t bw = (uint32_t)(cam.bx1 - cam.bx0), bh = (uint32_t)(cam.by1 - cam.by0);
    const uint32_t npix = bw * bh;
    if (li >= npix) return;
    const int32_t x = cam.bx0 + (int32_t)(li % bw), y = cam.by0 + (int32_t)(li / bw);
    const size_t p = (size_t)y * cam.width + (size_t)x;
    f3 acc = f3{sum[3 * p], sum[3 * p + 1], sum[3 * p + 2]};
    float w = weight[p];
    for (uint32_t s = 0; s < k; s++) {
        const uint32_t slot = s * npix + li;
        int32_t x0, y0, x1, y1;
        footprint(cam, P.F(F_PX, slot), P.F(F_PY, slot), x0, y0, x1, y1);
        if (x0 == x && y0 == y && x1 == x + 1 && y1 == y + 1) {
            acc = acc + P.F3(F_VX, slot) * 1.0f;
            w = w + 1.0f;
        }
    }
    sum[3 * p] = acc.x;
    sum[3 * p + 1] = acc.y;
    sum[3 * p + 2] = acc.z;
    weight[p] = w;
}

// samples on a pixel boundary (a jitter of exactly 0, or x + u rounding up to x + 1) touch up to 4 pixels
__global__ void k_film_edge(dpaths P, dcamera cam, uint32_t n, float* sum, float* weight) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    const uint32_t bw = (uint32_t)(cam.bx1 - cam.bx0);
    const uint32_t npix = bw * (uint32_t)(cam.by1 - cam.by0);
    const uint32_t li = slot % npix;
    const int32_t x = cam.bx0 + (int32_t)(li % bw), y = cam.by0 + (int32_t)(li / bw);
    int32_t x0, y0, x1, y1;
    footprint(cam, P.F(F_PX, slot), P.F(F_PY, slot), x0, y0, x1, y1);
    if (x0 == x && y0 == y && x1 == x + 1 && y1 == y + 1) return;
    const f3 v = P.F3(F_VX, slot) * 1.0f;
    for (int32_t yy = y0; yy < y1; yy++)
        for (int32_t xx = x0; xx < x1; xx++) {
            const size_t p = (size_t)yy * cam.width + (size_t)xx;
            atomicAdd(&sum[3 * p], v.x);
            atomicAdd(&sum[3 * p + 1], v.y);
            atomicAdd(&sum[3 * p + 2], v.z);
            atomicAdd(&weight[p], 1.0f);
        }
}

// pixel::spectrum (film.cpp:20-25): sum / weight; pixels outside the bounds are 0
__global__ void k_film_resolve(dcamera cam, const float* sum, const float* weight, float* rgb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cam.width * cam.height) return;
    const int32_t x = (int32_t)(p % cam.width), y = (int32_t)(p / cam.width);
    const bool in = x >= cam.bx0 && x < cam.bx1 && y >= cam.by0 && y < cam.by1 && weight[p] != 0.0f;
    const float w = weight[p];
    rgb[3 * p] = in ? sum[3 * p] / w : 0.0f;
    rgb[3 * p + 1] = in ? sum[3 * p + 1] / w : 0.0f;
    rgb[3 * p + 2] = in ? sum[3 * p + 2] / w : 0.0f;
}

// scene::intersect / intersect_with_shadow_ray for caller-given rays
__global__ void k_trace_rays(dscene S, uint32_t n, const float* o, const float* d, const float* tmax, uint32_t visible,
                             uint32_t tri_test, rbw_hit* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 ro = f3{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
    const f3 rd = f3{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
    trace_counters cnt{0, 0};
    const hitrec h = visible ? traverse<true, false, false>(S, ro, rd, tmax[i], tri_test, RBW_NONE, cnt)
                             : traverse<false, false, false>(S, ro, rd, tmax[i], tri_test, RBW_NONE, cnt);
    rbw_hit r;
    r.prim = h.prim;
    r.entity = h.prim == RBW_MISS ? RBW_NONE
                                  : (h.prim < S.num_tris ? S.tri_entity[h.prim] : S.sphere_entity[h.prim - S.num_tris]);
    r.t = h.t;
    r.b1 = h.b1;
    r.b2 = h.b2;
    out[i] = r;
}

// ================================================================== host side

namespace {

thread_local std::string g_last_error;

struct rbw_error : std::runtime_error {
    rbw_status code;
    rbw_error(rbw_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                                      \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess)                                                                                \
            throw rbw_error(_e == hipErrorOutOfMemory ? RBW_ERR_OUT_OF_MEMORY : RBW_ERR_DEVICE,              \
                            std::string(#expr) + ": " + hipGetErrorString(_e));                              \
    } while (0)

template <typename F>
rbw_status guarded(F&& f) {
    try {
        f();
        return RBW_OK;
    } catch (const rbw_error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return RBW_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return RBW_ERR_DEVICE;
    }
}

void require(bool cond, rbw_status code, const std::string& msg) {
    if (!cond) throw rbw_error(code, msg);
}

struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        HIP_CHECK(hipMalloc(&p, n ? n : 4));
        bytes = n;
    }
    template <typename T>
    void upload(const T* src, size_t count, hipStream_t s) {
        ensure(count * sizeof(T));
        if (count) HIP_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

constexpr uint32_t k_instanced_threshold = 1u << 17;  // scene.cpp:43
constexpr uint32_t k_max_light_triangles = 4096;
constexpr uint32_t k_train_pool = 8;  // training samples kept per pass, in batches

struct kernel_timer {
    struct rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<rec> recs;
    size_t used = 0;
    hipEvent_t get() {
        if (used < pool.size()) return pool[used++];
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        pool.push_back(e);
        used++;
        return e;
    }
    void reset() {
        recs.clear();
        used = 0;
    }
    void begin(int kind, hipStream_t s) {
        rec r{kind, get(), nullptr};
        HIP_CHECK(hipEventRecord(r.a, s));
        recs.push_back(r);
    }
    void end(hipStream_t s) {
        recs.back().b = get();
        HIP_CHECK(hipEventRecord(recs.back().b, s));
    }
    ~kernel_timer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace

struct rbw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool has_scene = false;
    uint32_t num_tris = 0, num_spheres = 0, num_entities = 0, num_materials = 0;
    std::vector<dentity> ents;
    std::vector<uint32_t> emitters;
    uint32_t bvh_depth = 0, bvh_nodes = 0;
    // scene buffers
    dbuf nodes, prim_a, prim_b, prim_c, pos, nrm, uv, idx, tri_entity, tri_material, spheres, sphere_entity,
        d_ents, d_mats, d_emitters;
    bool has_tri_material = false, has_uv = false;
    // render buffers
    dbuf pf, pu, prng, q0, q1, cq, counts, film_sum, film_w, rgb, dbg, replay_off, replay_val, counters;
    uint32_t path_cap = 0;
    uint32_t* host_count = nullptr;  // pinned
    kernel_timer timer;
    // guiding
    f3 scene_min{0, 0, 0}, scene_max{0, 0, 0};
    guide::model gm;
    bool guide_ready = false;
    dbuf g_surf, g_mat, g_q, g_query, g_mix, g_rec, g_flag, g_tin, g_tdir, g_tc, g_io_in, g_io_out, g_io_dir, g_io_c;
    uint32_t g_cap = 0, g_rec_cap = 0, g_t_cap = 0;
    void ensure_guide() {
        if (!guide_ready) {
            guide::model_init(gm, 0x5eed, stream);
            guide_ready = true;
        }
    }

    dscene scene_view() const {
        dscene S;
        S.nodes = nodes.as<float4>();
        S.prim_a = prim_a.as<float4>();
        S.prim_b = prim_b.as<float4>();
        S.prim_c = prim_c.as<float4>();
        S.pos = pos.as<float>();
        S.nrm = nrm.as<float>();
        S.uv = has_uv ? uv.as<float>() : nullptr;
        S.idx = idx.as<uint32_t>();
        S.tri_entity = tri_entity.as<uint32_t>();
        S.tri_material = has_tri_material ? tri_material.as<uint32_t>() : nullptr;
        S.spheres = spheres.as<float4>();
        S.sphere_entity = sphere_entity.as<uint32_t>();
        S.ents = d_ents.as<dentity>();
        S.mats = d_mats.as<dmaterial>();
        S.emitters = d_emitters.as<uint32_t>();
        S.num_emitters = (uint32_t)emitters.size();
        S.num_tris = num_tris;
        S.num_spheres = num_spheres;
        return S;
    }

    ~rbw_ctx() {
        (void)hipSetDevice(device);
        for (dbuf* b : {&nodes, &prim_a, &prim_b, &prim_c, &pos, &nrm, &uv, &idx, &tri_entity, &tri_material, &spheres,
                        &sphere_entity, &d_ents, &d_mats, &d_emitters, &pf, &pu, &prng, &q0, &q1, &cq, &counts,
                        &film_sum, &film_w, &rgb, &dbg, &replay_off, &replay_val, &counters, &g_surf, &g_mat, &g_q,
                        &g_query, &g_mix, &g_rec, &g_flag, &g_tin, &g_tdir, &g_tc, &g_io_in, &g_io_out, &g_io_dir,
                        &g_io_c})
            b->release();
        guide::model_free(gm);
        if (host_count) (void)hipHostFree(host_count);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline float expand_lo(float v) { return v - (fabsf(v) * 4.8e-7f + 1e-30f); }
inline float expand_hi(float v) { return v + (fabsf(v) * 4.8e-7f + 1e-30f); }

void upload_scene(rbw_ctx* ctx, const rbw_scene_desc* d) {
    require(d != nullptr, RBW_ERR_INVALID_ARGUMENT, "scene is null");
    require(d->num_entities > 0 && d->entities, RBW_ERR_INVALID_ARGUMENT, "scene has no entities");
    require(d->num_triangles == 0 || (d->positions && d->indices && d->triangle_entity), RBW_ERR_INVALID_ARGUMENT,
            "triangle arrays missing");
    require(d->num_spheres == 0 || (d->spheres && d->sphere_entity), RBW_ERR_INVALID_ARGUMENT, "sphere arrays missing");
    require(d->num_triangles + d->num_spheres > 0, RBW_ERR_INVALID_ARGUMENT, "scene has no primitives");
    require(uint64_t(d->num_triangles) + d->num_spheres < (1ull << 27), RBW_ERR_UNSUPPORTED, "too many primitives");
    require(d->num_materials > 0 && d->materials, RBW_ERR_INVALID_ARGUMENT, "scene has no materials");
    for (uint32_t i = 0; i < d->num_materials; i++) {
        const rbw_material& m = d->materials[i];
        require(m.type == RBW_MATERIAL_MATTE || m.type == RBW_MATERIAL_METAL, RBW_ERR_INVALID_ARGUMENT,
                "unknown material type");
        require(m.type != RBW_MATERIAL_MATTE || m.sigma == 0.0f, RBW_ERR_UNSUPPORTED,
                "matte sigma != 0 (Oren-Nayar) is not supported");
        require(m.type != RBW_MATERIAL_METAL || (m.alpha_u > 0.0f && m.alpha_v > 0.0f), RBW_ERR_INVALID_ARGUMENT,
                "metal alpha must be positive");
    }
    const uint32_t nt = d->num_triangles, ns = d->num_spheres, nv = d->num_vertices;
    for (uint64_t i = 0; i < 3ull * nt; i++)
        require(d->indices[i] < nv, RBW_ERR_INVALID_ARGUMENT, "triangle index out of range");
    std::vector<dentity> ents(d->num_entities);
    std::vector<uint32_t> emitters;
    for (uint32_t e = 0; e < d->num_entities; e++) {
        const rbw_entity& in = d->entities[e];
        dentity& o = ents[e];
        std::memset(&o, 0, sizeof(o));
        o.shape = in.shape;
        o.first = in.first;
        o.count = in.count;
        o.material = in.material;
        o.flags = in.flags;
        o.radiance[0] = in.radiance[0];
        o.radiance[1] = in.radiance[1];
        o.radiance[2] = in.radiance[2];
        require(in.material == RBW_NONE || in.material < d->num_materials, RBW_ERR_INVALID_ARGUMENT,
                "entity material out of range");
        if (in.shape == RBW_SHAPE_MESH) {
            require(uint64_t(in.first) + in.count <= nt && in.count > 0, RBW_ERR_INVALID_ARGUMENT,
                    "mesh entity triangle range out of bounds");
            require(!(in.flags & RBW_ENTITY_HAS_NORMALS) || d->normals, RBW_ERR_INVALID_ARGUMENT,
                    "entity has normals but the scene has none");
            require(!(in.flags & RBW_ENTITY_HAS_UVS) || d->uvs, RBW_ERR_INVALID_ARGUMENT,
                    "entity has uvs but the scene has none");
            for (uint32_t i = 0; i < in.count; i++)
                require(d->triangle_entity[in.first + i] == e, RBW_ERR_INVALID_ARGUMENT,
                        "triangle_entity disagrees with the entity ranges");
            // mesh::area(transform) with the identity transform (mesh.cpp:133-152)
            float area = 0.0f;
            for (uint32_t i = 0; i < in.count; i++) {
                const uint32_t t = in.first + i;
                const f3 p0 = ident(load3(d->positions, d->indices[3 * t]));
                const f3 p1 = ident(load3(d->positions, d->indices[3 * t + 1]));
                const f3 p2 = ident(load3(d->positions, d->indices[3 * t + 2]));
                area = area + 0.5f * length(cross(p2 - p0, p2 - p1));
            }
            o.area = area;
        } else if (in.shape == RBW_SHAPE_SPHERE) {
            require(in.first < ns, RBW_ERR_INVALID_ARGUMENT, "sphere entity index out of bounds");
            require(d->sphere_entity[in.first] == e, RBW_ERR_INVALID_ARGUMENT,
                    "sphere_entity disagrees with the entity");
            o.center[0] = d->spheres[4 * in.first];
            o.center[1] = d->spheres[4 * in.first + 1];
            o.center[2] = d->spheres[4 * in.first + 2];
            o.radius = d->spheres[4 * in.first + 3];
            require(o.radius > 0.0f, RBW_ERR_INVALID_ARGUMENT, "sphere radius must be positive");
        } else {
            require(in.shape == RBW_SHAPE_NONE, RBW_ERR_INVALID_ARGUMENT, "unknown shape kind");
        }
        if (in.flags & RBW_ENTITY_EMITTER) {
            require(in.shape == RBW_SHAPE_MESH, RBW_ERR_UNSUPPORTED, "only mesh area lights are supported");
            require(in.count <= k_max_light_triangles, RBW_ERR_UNSUPPORTED, "area light mesh has too many triangles");
            emitters.push_back(e);
        }
    }
    // integrator.cpp:89,121 dereference the sampled emitter: a scene needs one
    require(!emitters.empty(), RBW_ERR_INVALID_ARGUMENT, "scene has no emitter");
    if (d->triangle_material)
        for (uint32_t t = 0; t < nt; t++)
            require(d->triangle_material[t] == RBW_NONE || d->triangle_material[t] < d->num_materials,
                    RBW_ERR_INVALID_ARGUMENT, "triangle material out of range");

    // primitive boxes and the BVH over all entities' primitives
    std::vector<aabb> boxes(size_t(nt) + ns);
    for (uint32_t t = 0; t < nt; t++) {
        aabb& b = boxes[t];
        for (int a = 0; a < 3; a++) {
            b.lo[a] = 3.4e38f;
            b.hi[a] = -3.4e38f;
        }
        for (int k = 0; k < 3; k++) {
            const float* p = d->positions + 3 * size_t(d->indices[3 * t + k]);
            for (int a = 0; a < 3; a++) {
                b.lo[a] = std::min(b.lo[a], p[a]);
                b.hi[a] = std::max(b.hi[a], p[a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            b.lo[a] = expand_lo(b.lo[a]);
            b.hi[a] = expand_hi(b.hi[a]);
        }
    }
    for (uint32_t s = 0; s < ns; s++) {
        aabb& b = boxes[size_t(nt) + s];
        const float r = d->spheres[4 * s + 3] * 1.000001f;
        for (int a = 0; a < 3; a++) {
            b.lo[a] = expand_lo(d->spheres[4 * s + a] - r);
            b.hi[a] = expand_hi(d->spheres[4 * s + a] + r);
        }
    }
    f3 smin{3.4e38f, 3.4e38f, 3.4e38f}, smax{-3.4e38f, -3.4e38f, -3.4e38f};
    for (const aabb& b : boxes) {
        smin = f3{std::min(smin.x, b.lo[0]), std::min(smin.y, b.lo[1]), std::min(smin.z, b.lo[2])};
        smax = f3{std::max(smax.x, b.hi[0]), std::max(smax.y, b.hi[1]), std::max(smax.z, b.hi[2])};
    }
    ctx->scene_min = smin;
    ctx->scene_max = smax;
    bvh_build_result bvh;
    build_bvh(boxes, bvh, 4);
    require(bvh.depth < RBW_STACK, RBW_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    const size_t nslots = bvh.order.size();
    std::vector<float4> pa(nslots), pb(nslots), pc(nslots);
    for (size_t sl = 0; sl < nslots; sl++) {
        const uint32_t prim = bvh.order[sl];
        uint32_t word = prim;
        if (prim < nt) {
            const dentity& E = ents[d->triangle_entity[prim]];
            if (E.material == RBW_NONE) word |= k_prim_invisible;
            if (E.count > k_instanced_threshold) word |= k_prim_instanced;
            const float* p0 = d->positions + 3 * size_t(d->indices[3 * prim]);
            const float* p1 = d->positions + 3 * size_t(d->indices[3 * prim + 1]);
            const float* p2 = d->positions + 3 * size_t(d->indices[3 * prim + 2]);
            float wf;
            std::memcpy(&wf, &word, 4);
            pa[sl] = make_float4(p0[0], p0[1], p0[2], wf);
            pb[sl] = make_float4(p1[0], p1[1], p1[2], 0.0f);
            pc[sl] = make_float4(p2[0], p2[1], p2[2], 0.0f);
        } else {
            const uint32_t s = prim - nt;
            const dentity& E = ents[d->sphere_entity[s]];
            word |= k_prim_sphere;
            if (E.material == RBW_NONE) word |= k_prim_invisible;
            float wf;
            std::memcpy(&wf, &word, 4);
            pa[sl] = make_float4(d->spheres[4 * s], d->spheres[4 * s + 1], d->spheres[4 * s + 2], wf);
            pb[sl] = make_float4(d->spheres[4 * s + 3], 0.0f, 0.0f, 0.0f);
            pc[sl] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    std::vector<dmaterial> mats(d->num_materials);
    for (uint32_t i = 0; i < d->num_materials; i++) {
        const rbw_material& m = d->materials[i];
        dmaterial& o = mats[i];
        o.type = m.type;
        for (int c = 0; c < 3; c++) {
            o.diffuse[c] = m.diffuse[c];
            o.eta[c] = m.eta[c];
            o.k[c] = m.k[c];
        }
        o.alpha_u = m.alpha_u;
        o.alpha_v = m.alpha_v;
    }
    hipStream_t s = ctx->stream;
    ctx->nodes.upload(bvh.nodes2.data(), bvh.nodes2.size(), s);
    ctx->prim_a.upload(pa.data(), pa.size(), s);
    ctx->prim_b.upload(pb.data(), pb.size(), s);
    ctx->prim_c.upload(pc.data(), pc.size(), s);
    ctx->pos.upload(d->positions, size_t(nv) * 3, s);
    if (d->normals) ctx->nrm.upload(d->normals, size_t(nv) * 3, s);
    else ctx->nrm.ensure(4);
    ctx->has_uv = d->uvs != nullptr;
    if (d->uvs) ctx->uv.upload(d->uvs, size_t(nv) * 2, s);
    ctx->idx.upload(d->indices, size_t(nt) * 3, s);
    ctx->tri_entity.upload(d->triangle_entity, nt, s);
    ctx->has_tri_material = d->triangle_material != nullptr;
    if (d->triangle_material) ctx->tri_material.upload(d->triangle_material, nt, s);
    std::vector<float4> sph(ns);
    for (uint32_t i = 0; i < ns; i++)
        sph[i] = make_float4(d->spheres[4 * i], d->spheres[4 * i + 1], d->spheres[4 * i + 2], d->spheres[4 * i + 3]);
    ctx->spheres.upload(sph.data(), sph.size(), s);
    ctx->sphere_entity.upload(d->sphere_entity ? d->sphere_entity : emitters.data(), ns, s);
    ctx->d_ents.upload(ents.data(), ents.size(), s);
    ctx->d_mats.upload(mats.data(), mats.size(), s);
    ctx->d_emitters.upload(emitters.data(), emitters.size(), s);
    HIP_CHECK(hipStreamSynchronize(s));
    ctx->num_tris = nt;
    ctx->num_spheres = ns;
    ctx->num_entities = d->num_entities;
    ctx->num_materials = d->num_materials;
    ctx->ents = ents;
    ctx->emitters = emitters;
    ctx->bvh_depth = bvh.depth;
    ctx->bvh_nodes = (uint32_t)bvh.nodes2.size();
    ctx->has_scene = true;
}

dcamera make_dcamera(const rbw_camera_desc* c) {
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

inline uint32_t blocks(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

void render(rbw_ctx* ctx, const rbw_camera_desc* cam, const rbw_render_params* prm_in, float* out_rgb,
            rbw_stats* stats) {
    require(ctx->has_scene, RBW_ERR_NO_SCENE, "no scene uploaded");
    require(cam && prm_in, RBW_ERR_INVALID_ARGUMENT, "camera or params is null");
    const rbw_render_params& pr = *prm_in;
    require(pr.spp > 0 && pr.max_depth > 0, RBW_ERR_INVALID_ARGUMENT, "spp and max_depth must be positive");
    require(pr.max_depth < (1u << 20), RBW_ERR_INVALID_ARGUMENT, "max_depth too large");
    require(cam->width > 0 && cam->height > 0, RBW_ERR_INVALID_ARGUMENT, "empty film");
    require(cam->bound_min[0] >= 0 && cam->bound_min[1] >= 0 && cam->bound_max[0] <= (int32_t)cam->width &&
                cam->bound_max[1] <= (int32_t)cam->height && cam->bound_min[0] < cam->bound_max[0] &&
                cam->bound_min[1] < cam->bound_max[1],
            RBW_ERR_INVALID_ARGUMENT, "pixel bounds outside the film");
    require(pr.sampler == RBW_SAMPLER_PER_SAMPLE_PCG || pr.sampler == RBW_SAMPLER_REPLAY, RBW_ERR_INVALID_ARGUMENT,
            "unknown sampler");
    require(pr.triangle_test == RBW_TRI_MOLLER_TRUMBORE || pr.triangle_test == RBW_TRI_WATERTIGHT,
            RBW_ERR_INVALID_ARGUMENT, "unknown triangle test");
    require(pr.shadow_mode == RBW_SHADOW_CLOSEST_VISIBLE || pr.shadow_mode == RBW_SHADOW_ANY_HIT,
            RBW_ERR_INVALID_ARGUMENT, "unknown shadow mode");
    const uint32_t W = cam->width, H = cam->height;
    const uint32_t bw = uint32_t(cam->bound_max[0] - cam->bound_min[0]);
    const uint32_t bh = uint32_t(cam->bound_max[1] - cam->bound_min[1]);
    const uint64_t npix = uint64_t(bw) * bh;
    const uint64_t nsamples_total = uint64_t(W) * H * pr.spp;
    require(npix <= (1u << 28), RBW_ERR_UNSUPPORTED, "film too large");
    const rbw_guiding_opts* go = pr.guiding;
    const bool guided = go && go->enable;
    if (guided) {
        require(go->mis_fraction >= 0.0f && go->mis_fraction <= 1.0f, RBW_ERR_INVALID_ARGUMENT,
                "guiding mis_fraction must be in [0, 1]");
        require(!go->train || (go->batch > 0 && go->batch <= (1u << 20)), RBW_ERR_INVALID_ARGUMENT,
                "guiding batch must be in [1, 2^20]");
        require(pr.sampler == RBW_SAMPLER_PER_SAMPLE_PCG, RBW_ERR_INVALID_ARGUMENT,
                "guiding draws a different number of samples per vertex than the replayed reference stream");
    }
    uint32_t k = pr.samples_per_pass;
    // guided renders train between passes: smaller passes give the network more updates
    const uint64_t per_pass = guided ? (1u << 18) : (1u << 21);
    if (k == 0) k = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(pr.spp, per_pass / npix));
    k = std::min(k, pr.spp);
    while (uint64_t(k) * npix > (1ull << 30) && k > 1) k--;
    const uint32_t cap = uint32_t(uint64_t(k) * npix);
    hipStream_t s = ctx->stream;
    if (ctx->path_cap < cap) {
        ctx->pf.ensure(size_t(F_COUNT) * cap * sizeof(float));
        ctx->pu.ensure(size_t(U_COUNT) * cap * sizeof(uint32_t));
        ctx->prng.ensure(size_t(cap) * sizeof(uint64_t));
        ctx->q0.ensure(size_t(cap) * 4);
        ctx->q1.ensure(size_t(cap) * 4);
        ctx->cq.ensure(size_t(cap) * 4);
        ctx->path_cap = cap;
    }
    ctx->counts.ensure(16 * sizeof(uint32_t));
    HIP_CHECK(hipMemsetAsync(ctx->counts.p, 0, 16 * sizeof(uint32_t), s));
    ctx->counters.ensure(16 * sizeof(unsigned long long));
    ctx->film_sum.ensure(size_t(W) * H * 3 * sizeof(float));
    ctx->film_w.ensure(size_t(W) * H * sizeof(float));
    ctx->rgb.ensure(size_t(W) * H * 3 * sizeof(float));
    HIP_CHECK(hipMemsetAsync(ctx->film_sum.p, 0, size_t(W) * H * 3 * sizeof(float), s));
    HIP_CHECK(hipMemsetAsync(ctx->film_w.p, 0, size_t(W) * H * sizeof(float), s));
    HIP_CHECK(hipMemsetAsync(ctx->counters.p, 0, 16 * sizeof(unsigned long long), s));
    dparams dp;
    dp.spp = pr.spp;
    dp.max_depth = pr.max_depth;
    dp.rr_threshold = pr.rr_threshold;
    dp.seed = pr.seed;
    dp.sampler = pr.sampler;
    dp.tri_test = pr.triangle_test;
    dp.shadow_mode = pr.shadow_mode;
    dp.replay_offsets = nullptr;
    dp.replay_values = nullptr;
    dp.debug = nullptr;
    if (pr.sampler == RBW_SAMPLER_REPLAY) {
        require(pr.replay_offsets && pr.replay_values, RBW_ERR_INVALID_ARGUMENT, "replay buffers missing");
        for (uint64_t i = 0; i < nsamples_total; i++)
            require(pr.replay_offsets[i] <= pr.replay_offsets[i + 1], RBW_ERR_INVALID_ARGUMENT,
                    "replay offsets must not decrease");
        require(pr.replay_offsets[nsamples_total] <= pr.replay_count, RBW_ERR_INVALID_ARGUMENT,
                "replay offsets exceed replay_count");
        ctx->replay_off.upload(pr.replay_offsets, nsamples_total + 1, s);
        ctx->replay_val.upload(pr.replay_values, pr.replay_count, s);
        dp.replay_offsets = ctx->replay_off.as<uint64_t>();
        dp.replay_values = ctx->replay_val.as<float>();
    }
    const size_t ndbg = pr.debug_hits ? size_t(nsamples_total) * pr.max_depth * 3 : 0;
    if (ndbg) {
        ctx->dbg.ensure(ndbg * sizeof(uint32_t));
        HIP_CHECK(hipMemsetAsync(ctx->dbg.p, 0xFF, ndbg * sizeof(uint32_t), s));
        dp.debug = ctx->dbg.as<uint32_t>();
    }
    std::memset(&dp.g, 0, sizeof(dp.g));
    if (guided) {
        ctx->ensure_guide();
        const uint32_t gcap = ctx->path_cap;
        if (ctx->g_cap < gcap) {
            ctx->g_surf.ensure(size_t(G_COUNT) * gcap * sizeof(float));
            ctx->g_mat.ensure(size_t(gcap) * 4);
            ctx->g_q.ensure(size_t(gcap) * 4);
            ctx->g_query.ensure(size_t(gcap) * 9 * sizeof(float));
            ctx->g_mix.ensure(size_t(gcap + 16) * guide::N_OUT_PAD * sizeof(float));
            ctx->g_cap = gcap;
        }
        dp.g.enable = 1;
        dp.g.alpha = go->mis_fraction;
        dp.g.train = go->train ? 1u : 0u;
        f3 smin = ctx->scene_min, smax = ctx->scene_max;
        if (go->scene_min[0] != 0.0f || go->scene_min[1] != 0.0f || go->scene_min[2] != 0.0f ||
            go->scene_max[0] != 0.0f || go->scene_max[1] != 0.0f || go->scene_max[2] != 0.0f) {
            smin = f3{go->scene_min[0], go->scene_min[1], go->scene_min[2]};
            smax = f3{go->scene_max[0], go->scene_max[1], go->scene_max[2]};
        }
        dp.g.smin = smin;
        dp.g.sinv = f3{1.0f / std::max(smax.x - smin.x, 1e-6f), 1.0f / std::max(smax.y - smin.y, 1e-6f),
                       1.0f / std::max(smax.z - smin.z, 1e-6f)};
        dp.g.surf = ctx->g_surf.as<float>();
        dp.g.surf_mat = ctx->g_mat.as<uint32_t>();
        dp.g.gq = ctx->g_q.as<uint32_t>();
        dp.g.gq_count = ctx->counts.as<uint32_t>() + 3;
        dp.g.query = ctx->g_query.as<float>();
        dp.g.mix = ctx->g_mix.as<float>();
        dp.g.cap = ctx->path_cap;
        if (go->train) {
            const size_t nrec = size_t(ctx->path_cap) * pr.max_depth;
            if (ctx->g_rec_cap < nrec) {
                ctx->g_rec.ensure(size_t(R_COUNT) * nrec * sizeof(float));
                ctx->g_flag.ensure(nrec * 4);
                ctx->g_rec_cap = uint32_t(nrec);
            }
            HIP_CHECK(hipMemsetAsync(ctx->g_flag.p, 0, nrec * 4, s));
            // a pass closes up to one continuation record per path vertex; it keeps up to
            // k_train_pool batches of them and each step draws a batch from that pool
            const uint32_t tcap = go->batch * k_train_pool;
            if (ctx->g_t_cap < tcap) {
                ctx->g_tin.ensure(size_t(tcap) * 9 * sizeof(float));
                ctx->g_tdir.ensure(size_t(tcap) * 3 * sizeof(float));
                ctx->g_tc.ensure(size_t(tcap) * sizeof(float));
                ctx->g_t_cap = tcap;
            }
            dp.g.rec = ctx->g_rec.as<float>();
            dp.g.rec_flag = ctx->g_flag.as<uint32_t>();
            dp.g.t_in = ctx->g_tin.as<float>();
            dp.g.t_dir = ctx->g_tdir.as<float>();
            dp.g.t_c = ctx->g_tc.as<float>();
            dp.g.t_count = ctx->counts.as<uint32_t>() + 4;
            dp.g.t_cap = tcap;
        }
    }
    const dscene S = ctx->scene_view();
    dpaths P;
    P.f = ctx->pf.as<float>();
    P.u = ctx->pu.as<uint32_t>();
    P.rng = ctx->prng.as<uint64_t>();
    P.cap = ctx->path_cap;
    const dcamera dc = make_dcamera(cam);
    uint32_t* counts = ctx->counts.as<uint32_t>();  // [0] queue A, [1] queue B, [2] connect queue
    unsigned long long* counters = ctx->counters.as<unsigned long long>();
    const bool count = pr.count_visits != 0;
    const bool timed = stats != nullptr;
    ctx->timer.reset();
    hipEvent_t ev0 = ctx->timer.get(), ev1 = ctx->timer.get();
    HIP_CHECK(hipEventRecord(ev0, s));
    uint64_t ext_rays = 0, guide_samples = 0, guide_steps = 0;
    const uint32_t BS = 256;
    uint32_t* qa = ctx->q0.as<uint32_t>();
    uint32_t* qb = ctx->q1.as<uint32_t>();
    uint32_t* cqp = ctx->cq.as<uint32_t>();
    for (uint32_t s0 = 0; s0 < pr.spp; s0 += k) {
        const uint32_t kk = std::min(k, pr.spp - s0);
        const uint32_t n = uint32_t(uint64_t(kk) * npix);
        // the pass's training samples are its closed continuation records
        if (guided && go->train) HIP_CHECK(hipMemsetAsync(counts + 4, 0, 4, s));
        if (timed) ctx->timer.begin(RBW_KERNEL_RAYGEN, s);
        hipLaunchKernelGGL(k_raygen, dim3(blocks(n, BS)), dim3(BS), 0, s, S, P, dp, dc, n, s0, qa, counts + 0);
        if (timed) ctx->timer.end(s);
        HIP_CHECK(hipGetLastError());
        uint32_t active = n;
        uint32_t* qin = qa;
        uint32_t* qout = qb;
        int cin = 0, cout = 1;
        // each iteration ends a bounce of every live path; only invisible (material-less)
        // entities repeat a bounce index, so the loop is bounded for any finite scene
        uint32_t iterations = 0;
        const uint32_t max_iterations = 4 * pr.max_depth + 256;
        while (active > 0) {
            require(++iterations <= max_iterations, RBW_ERR_DEVICE, "path loop did not terminate");
            ext_rays += active;
            HIP_CHECK(hipMemsetAsync(counts + cout, 0, 4, s));
            HIP_CHECK(hipMemsetAsync(counts + 2, 0, 4, s));
            if (timed) ctx->timer.begin(RBW_KERNEL_EXTEND, s);
            if (count)
                hipLaunchKernelGGL(k_extend<true>, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, qin,
                                   counts + cin, counters);
            else
                hipLaunchKernelGGL(k_extend<false>, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, qin,
                                   counts + cin, counters);
            if (timed) ctx->timer.end(s);
            if (timed) ctx->timer.begin(RBW_KERNEL_SHADE, s);
            if (guided) {
                HIP_CHECK(hipMemsetAsync(counts + 3, 0, 4, s));
                hipLaunchKernelGGL(k_shade<true>, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, qin, counts + cin,
                                   qout, counts + cout, cqp, counts + 2);
            } else {
                hipLaunchKernelGGL(k_shade<false>, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, qin,
                                   counts + cin, qout, counts + cout, cqp, counts + 2);
            }
            if (timed) ctx->timer.end(s);
            if (guided) {
                HIP_CHECK(hipMemcpyAsync(ctx->host_count + 1, counts + 3, 4, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                const uint32_t ng = ctx->host_count[1];
                if (ng > 0) {
                    if (timed) ctx->timer.begin(RBW_KERNEL_GUIDE_INFER, s);
                    guide::forward(ctx->gm, dp.g.query, ng, dp.g.mix, s);
                    if (timed) ctx->timer.end(s);
                    if (timed) ctx->timer.begin(RBW_KERNEL_SHADE_GUIDED, s);
                    hipLaunchKernelGGL(k_shade_guided, dim3(blocks(ng, BS)), dim3(BS), 0, s, S, P, dp, qout,
                                       counts + cout, cqp, counts + 2);
                    if (timed) ctx->timer.end(s);
                    HIP_CHECK(hipGetLastError());
                }
            }
            if (timed) ctx->timer.begin(RBW_KERNEL_CONNECT, s);
            if (count)
                hipLaunchKernelGGL(k_connect<true>, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, cqp,
                                   counts + 2, counters);
            else
                hipLaunchKernelGGL(k_connect<false>, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, cqp,
                                   counts + 2, counters);
            if (timed) ctx->timer.end(s);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(ctx->host_count, counts + cout, 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            active = *ctx->host_count;
            std::swap(qin, qout);
            std::swap(cin, cout);
        }
        if (guided && go->train) {
            if (timed) ctx->timer.begin(RBW_KERNEL_GUIDE_TRAIN, s);
            hipLaunchKernelGGL(k_collect, dim3(blocks(uint64_t(n) * pr.max_depth, BS)), dim3(BS), 0, s, P, dp, n);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(ctx->host_count + 2, counts + 4, 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            const uint32_t pool = std::min(ctx->host_count[2], dp.g.t_cap);
            const uint32_t nt = std::min(pool, go->batch);
            guide_samples += ctx->host_count[2];
            if (nt >= 256) {
                guide_steps += go->steps_per_pass;
                guide::adam_opts ao;
                ao.lr = go->learning_rate;
                for (uint32_t st = 0; st < go->steps_per_pass; st++) {
                    const guide::sample_map map = guide::make_sample_map(pool, (uint64_t(pr.seed) << 32) ^
                                                                                   (uint64_t(s0) << 8) ^ st);
                    guide::train_step(ctx->gm, dp.g.t_in, dp.g.t_dir, dp.g.t_c, nt, ao, true, s, map);
                }
            }
            if (timed) ctx->timer.end(s);
        }
        if (timed) ctx->timer.begin(RBW_KERNEL_FILM, s);
        hipLaunchKernelGGL(k_film_own, dim3(blocks(npix, BS)), dim3(BS), 0, s, P, dc, kk, ctx->film_sum.as<float>(),
                           ctx->film_w.as<float>());
        hipLaunchKernelGGL(k_film_edge, dim3(blocks(n, BS)), dim3(BS), 0, s, P, dc, n, ctx->film_sum.as<float>(),
                           ctx->film_w.as<float>());
        if (timed) ctx->timer.end(s);
        HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_film_resolve, dim3(blocks(uint64_t(W) * H, BS)), dim3(BS), 0, s, dc,
                       ctx->film_sum.as<float>(), ctx->film_w.as<float>(), ctx->rgb.as<float>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev1, s));
    if (out_rgb)
        HIP_CHECK(hipMemcpyAsync(out_rgb, ctx->rgb.p, size_t(W) * H * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    if (ndbg) HIP_CHECK(hipMemcpyAsync(pr.debug_hits, ctx->dbg.p, ndbg * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    unsigned long long hc[16];
    HIP_CHECK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    float gloss = 0.0f;
    if (guide_steps) HIP_CHECK(hipMemcpyAsync(&gloss, ctx->gm.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        stats->render_ms = ms;
        stats->samples = npix * pr.spp;
        stats->camera_rays = stats->samples;
        stats->extension_rays = ext_rays;
        stats->shadow_rays = hc[2];
        stats->mis_rays = hc[3];
        stats->light_casts = hc[4];
        stats->extend_node_visits = hc[0];
        stats->extend_prim_tests = hc[1];
        stats->node_visits = hc[0] + hc[5];
        stats->prim_tests = hc[1] + hc[6];
        for (const auto& r : ctx->timer.recs) {
            float kms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&kms, r.a, r.b));
            stats->kernel_ms[r.kind] += kms;
            stats->kernel_launches[r.kind] += 1;
        }
        stats->guide_samples = guide_samples;
        stats->guide_steps = guide_steps;
        stats->guide_loss = gloss;
        const double sec = ms * 1e-3;
        stats->msamples_per_s = sec > 0 ? stats->samples / sec * 1e-6 : 0.0;
        const uint64_t rays = stats->extension_rays + stats->shadow_rays + stats->mis_rays;
        stats->mrays_per_s = sec > 0 ? rays / sec * 1e-6 : 0.0;
    }
}

}  // namespace

extern "C" {

rbw_status rbw_create(const rbw_device_opts* opts, rbw_ctx** out) {
    return guarded([&] {
        require(out != nullptr, RBW_ERR_INVALID_ARGUMENT, "out is null");
        *out = nullptr;
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        const int dev = opts ? opts->device : 0;
        require(dev >= 0 && dev < ndev, RBW_ERR_INVALID_ARGUMENT, "no such HIP device");
        auto ctx = std::make_unique<rbw_ctx>();
        ctx->device = dev;
        HIP_CHECK(hipSetDevice(dev));
        HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_count), 64, hipHostMallocDefault));
        *out = ctx.release();
    });
}

void rbw_destroy(rbw_ctx* ctx) { delete ctx; }

const char* rbw_last_error(void) { return g_last_error.c_str(); }

rbw_status rbw_upload_scene(rbw_ctx* ctx, const rbw_scene_desc* scene) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->has_scene = false;
        upload_scene(ctx, scene);
    });
}

rbw_status rbw_render(rbw_ctx* ctx, const rbw_camera_desc* camera, const rbw_render_params* params, float* out_rgb,
                      rbw_stats* stats) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        HIP_CHECK(hipSetDevice(ctx->device));
        render(ctx, camera, params, out_rgb, stats);
    });
}

rbw_status rbw_trace_rays(rbw_ctx* ctx, uint32_t n, const float* origins, const float* directions, const float* tmax,
                          uint32_t visible_only, uint32_t triangle_test, rbw_hit* out) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(ctx->has_scene, RBW_ERR_NO_SCENE, "no scene uploaded");
        require(n == 0 || (origins && directions && tmax && out), RBW_ERR_INVALID_ARGUMENT, "null ray arrays");
        require(triangle_test == RBW_TRI_MOLLER_TRUMBORE || triangle_test == RBW_TRI_WATERTIGHT,
                RBW_ERR_INVALID_ARGUMENT, "unknown triangle test");
        if (n == 0) return;
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        dbuf o, d, t, h;
        try {
            o.upload(origins, size_t(n) * 3, s);
            d.upload(directions, size_t(n) * 3, s);
            t.upload(tmax, n, s);
            h.ensure(size_t(n) * sizeof(rbw_hit));
            hipLaunchKernelGGL(k_trace_rays, dim3(blocks(n, 256)), dim3(256), 0, s, ctx->scene_view(), n, o.as<float>(),
                               d.as<float>(), t.as<float>(), visible_only, triangle_test, h.as<rbw_hit>());
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(out, h.p, size_t(n) * sizeof(rbw_hit), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            o.release(); d.release(); t.release(); h.release();
            throw;
        }
        o.release(); d.release(); t.release(); h.release();
    });
}

rbw_status rbw_guide_reset(rbw_ctx* ctx, uint32_t seed) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        HIP_CHECK(hipSetDevice(ctx->device));
        guide::model_init(ctx->gm, seed, ctx->stream);
        ctx->guide_ready = true;
    });
}

rbw_status rbw_guide_get_params(rbw_ctx* ctx, float* params) {
    return guarded([&] {
        require(ctx != nullptr && params != nullptr, RBW_ERR_INVALID_ARGUMENT, "null argument");
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->ensure_guide();
        guide::model_get(ctx->gm, params, ctx->stream);
    });
}

rbw_status rbw_guide_set_params(rbw_ctx* ctx, const float* params) {
    return guarded([&] {
        require(ctx != nullptr && params != nullptr, RBW_ERR_INVALID_ARGUMENT, "null argument");
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->ensure_guide();
        guide::model_set(ctx->gm, params, ctx->stream);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

rbw_status rbw_guide_forward(rbw_ctx* ctx, uint32_t n, const float* inputs, float* out) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(n == 0 || (inputs && out), RBW_ERR_INVALID_ARGUMENT, "null arrays");
        if (n == 0) return;
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->ensure_guide();
        hipStream_t s = ctx->stream;
        ctx->g_io_in.upload(inputs, size_t(n) * 9, s);
        const uint32_t n_pad = (n + 15) & ~15u;
        ctx->g_io_out.ensure(size_t(n_pad) * guide::N_OUT_PAD * sizeof(float));
        guide::forward(ctx->gm, ctx->g_io_in.as<float>(), n, ctx->g_io_out.as<float>(), s);
        HIP_CHECK(hipMemcpyAsync(out, ctx->g_io_out.p, size_t(n) * guide::N_OUT_PAD * sizeof(float),
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    });
}

rbw_status rbw_guide_train(rbw_ctx* ctx, uint32_t n, const float* inputs, const float* dirs, const float* c,
                           float learning_rate, uint32_t apply, float* loss, float* grad) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(n > 0 && inputs && dirs && c, RBW_ERR_INVALID_ARGUMENT, "null arrays");
        require(n <= (1u << 22), RBW_ERR_INVALID_ARGUMENT, "batch too large");
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->ensure_guide();
        hipStream_t s = ctx->stream;
        ctx->g_io_in.upload(inputs, size_t(n) * 9, s);
        ctx->g_io_dir.upload(dirs, size_t(n) * 3, s);
        ctx->g_io_c.upload(c, n, s);
        HIP_CHECK(hipMemsetAsync(ctx->gm.grad, 0, sizeof(float) * guide::N_PARAMS, s));
        guide::adam_opts ao;
        ao.lr = learning_rate;
        guide::train_step(ctx->gm, ctx->g_io_in.as<float>(), ctx->g_io_dir.as<float>(), ctx->g_io_c.as<float>(), n,
                          ao, false, s);
        if (grad || loss) {
            std::vector<float> g(guide::N_PARAMS);
            float l = 0.0f;
            HIP_CHECK(hipMemcpyAsync(g.data(), ctx->gm.grad, sizeof(float) * guide::N_PARAMS, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(&l, ctx->gm.loss, sizeof(float), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (grad)
                for (int i = 0; i < guide::N_PARAMS; i++) grad[i] = g[i] / (float)n;
            if (loss) *loss = l;
        }
        if (apply) {
            // the Adam step of train_step(apply) on the gradient just computed
            ctx->gm.step++;
            const float c1 = 1.0f - std::pow(ao.b1, (float)ctx->gm.step);
            const float c2 = 1.0f - std::pow(ao.b2, (float)ctx->gm.step);
            guide::adam_apply(ctx->gm, n, ao, c1, c2, s);
        }
        HIP_CHECK(hipMemsetAsync(ctx->gm.grad, 0, sizeof(float) * guide::N_PARAMS, s));
        HIP_CHECK(hipStreamSynchronize(s));
    });
}

}  // extern "C"
