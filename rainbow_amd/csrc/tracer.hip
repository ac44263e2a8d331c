// Wavefront path tracer for MI355X (gfx950) behind the C ABI of include/rbw.h.
//
// This file holds the pass schedule and the C ABI.  A render in passes holds
// (pixel, sample) paths in SoA state in HBM; each bounce runs three kernels
// over compacted queues:
//   extend  : closest-hit BVH traversal of the path rays
//   shade   : surface interaction, emission, material, light and BSDF samples
//             for next-event estimation, continuation sample, Russian roulette
//   connect : the shadow ray and the MIS BSDF ray of next-event estimation,
//             the light's own pdf cast, and the MIS-weighted contribution
// then film kernels accumulate the pass into the film.  Replay of recorded
// draws, debug hits, visit counting and guided rendering use it; the plain
// estimator runs on the persistent-slot schedule of batch.hip.  The per-path
// stages live in path.hpp and follow rainbow-cpu's path integrator
// (integrators/path_integrator.cpp:15-47, integrators/integrator.cpp:101-404).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rbw.h"
#include "bvh.hpp"
#include "device.hpp"
#include "guided.hpp"
#include "guiding.hpp"
#include "volume.hpp"
#include "image.hpp"
#include "host.hpp"
#include "path.hpp"
#include "shading.hpp"

using namespace rbw;

// ------------------------------------------------------------------ kernels

// sampler_integrator.cpp:98-127 + perspective_camera.cpp:78-99: one camera sample per slot
// (with_medium: the volume path integrator's camera rays start in vacuum, volume_path_integrator.cpp:21-24)
__global__ void k_raygen(dscene S, dpaths P, dparams prm, dcamera cam, uint32_t n, uint32_t s0, uint32_t* queue,
                         uint32_t* queue_count, uint32_t with_medium) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bw = (uint32_t)(cam.bx1 - cam.bx0), bh = (uint32_t)(cam.by1 - cam.by0);
    const uint32_t npix = bw * bh;
    const uint32_t li = i % npix, sl = i / npix;
    raygen_path(cam, prm, P, i, cam.bx0 + (int32_t)(li % bw), cam.by0 + (int32_t)(li / bw), s0 + sl);
    if (with_medium) P.U(U_MEDIUM, i) = RBW_NONE;
    queue[i] = i;
    if (i == 0) *queue_count = n;
}

template <int W, bool COUNT>
__global__ RBW_EXTEND_OCCUPANCY void k_extend(dscene S, dpaths P, dparams prm, const uint32_t* queue, const uint32_t* queue_count,
                         unsigned long long* counters) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *queue_count;
    trace_counters cnt{0, 0};
    if (q < count) extend_path<W, COUNT, true>(S, P, prm, queue[q], cnt);
    if (COUNT) {
        stripe_add64(counters, 0, cnt.nodes);
        stripe_add64(counters, 1, cnt.prims);
    }
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
        shade_path<GUIDED>(S, P, prm, slot, alive, need_connect, to_guide, gsi, guide_mat);
    }
    const uint32_t a = block_append(qout_count, alive);
    if (alive) qout[a] = slot;
    const uint32_t c = block_append(cq_count, need_connect);
    if (need_connect) cq[c] = slot;
    if (GUIDED) {
        const uint32_t g = block_append(prm.g.gq_count, to_guide);
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
            float qv[9];
            guide_query(gsi, G.smin, G.sinv, qv);
#pragma unroll
            for (int k = 0; k < 9; k++) G.query[(size_t)g * 9 + k] = qv[k];
        }
    }
}

// the guided vertex (guided.hpp) of each guide-queue row, its surface record and network
// outputs stored by row
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
        guided_vertex(S, P, prm, slot, si, G.surf_mat[g], G.mix + (size_t)g * guide::N_OUT_PAD,
                      G.query + (size_t)g * 9, g, alive, need_connect);
    }
    const uint32_t ai = block_append(qout_count, alive);
    if (alive) qout[ai] = slot;
    const uint32_t ci = block_append(cq_count, need_connect);
    if (need_connect) cq[ci] = slot;
}

// after a pass: closed training records -> (query, direction, radiance / pdf).  A
// block takes COLLECT_PER_THREAD records per thread (coalesced) and appends its
// takes to the training pool with one reservation.
constexpr uint32_t COLLECT_PER_THREAD = 8;
__global__ __launch_bounds__(256) void k_collect(dpaths P, dparams prm, uint32_t n) {
    const dguide& G = prm.g;
    const size_t total = (size_t)n * G.rec_depth;
    const size_t first = (size_t)blockIdx.x * blockDim.x * COLLECT_PER_THREAD + threadIdx.x;
    float c[COLLECT_PER_THREAD];
    uint32_t takes = 0, mask = 0;
#pragma unroll
    for (uint32_t j = 0; j < COLLECT_PER_THREAD; j++) {
        const size_t ri = first + (size_t)j * blockDim.x;
        c[j] = 0.0f;
        if (ri < total && G.rec_flag[ri] == 2) {
            c[j] = record_target(P, G, (uint32_t)(ri / G.rec_depth), ri);
            if (c[j] > 0.0f) {
                mask |= 1u << j;
                takes++;
            }
        }
        if (ri < total) G.rec_flag[ri] = 0;
    }
    uint32_t k = block_reserve(G.t_count, takes);
#pragma unroll
    for (uint32_t j = 0; j < COLLECT_PER_THREAD; j++) {
        if (!((mask >> j) & 1u)) continue;
        const size_t ri = first + (size_t)j * blockDim.x;
        if (k < G.t_cap) record_to_row(P, G, ri, c[j], G.t_in, G.t_dir, G.t_c, k);
        k++;
    }
}

// shadow ray, MIS ray and light pdf of next-event estimation; value += beta * L
template <int W, bool COUNT>
__global__ RBW_CONNECT_OCCUPANCY void k_connect(dscene S, dpaths P, dparams prm, const uint32_t* cq,
                                                const uint32_t* cq_count, unsigned long long* counters) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *cq_count;
    trace_counters cnt{0, 0};
    uint32_t nshadow = 0, nmis = 0, nlight = 0;
    bool lit = false;
    uint32_t slot = 0;
    if (q < count) {
        slot = cq[q];
        connect_path<W, COUNT, true>(S, P, prm, slot, cnt, nshadow, nmis, nlight, &lit);
    }
    if (prm.g.train && prm.g.light_records) {
        // unoccluded guided light samples become training records (query row, direction, weight)
        const float c = lit ? P.F(F_MFX, slot) : 0.0f;
        const bool take = lit && c > 0.0f && isfinite(c);
        const uint32_t k = block_append(prm.g.t_count, take);
        if (take && k < prm.g.t_cap) {
            const uint32_t g = __float_as_uint(P.F(F_MFY, slot));
            for (int j = 0; j < 9; j++) prm.g.t_in[(size_t)k * 9 + j] = prm.g.query[(size_t)g * 9 + j];
            const f3 d = P.F3(F_SDX, slot);
            prm.g.t_dir[(size_t)k * 3 + 0] = d.x;
            prm.g.t_dir[(size_t)k * 3 + 1] = d.y;
            prm.g.t_dir[(size_t)k * 3 + 2] = d.z;
            prm.g.t_c[k] = c;
        }
    }
    stripe_add64(counters, 2, nshadow);
    stripe_add64(counters, 3, nmis);
    stripe_add64(counters, 4, nlight);
    if (COUNT) {
        stripe_add64(counters, 5, cnt.nodes);
        stripe_add64(counters, 6, cnt.prims);
    }
}

// one thread per pixel: the pass's samples whose footprint is exactly the own pixel, in sample order
__global__ void k_film_own(dpaths P, dcamera cam, uint32_t k, dfilm film) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t bw = (uint32_t)(cam.bx1 - cam.bx0), bh = (uint32_t)(cam.by1 - cam.by0);
    const uint32_t npix = bw * bh;
    if (li >= npix) return;
    const int32_t x = cam.bx0 + (int32_t)(li % bw), y = cam.by0 + (int32_t)(li / bw);
    const size_t p = (size_t)y * cam.width + (size_t)x;
    long long acc[3] = {(long long)film.sum[3 * p], (long long)film.sum[3 * p + 1], (long long)film.sum[3 * p + 2]};
    uint32_t w = film.weight[p], bad = 0;
    for (uint32_t s = 0; s < k; s++) {
        const uint32_t slot = s * npix + li;
        int32_t x0, y0, x1, y1;
        footprint(cam, P.F(F_PX, slot), P.F(F_PY, slot), x0, y0, x1, y1);
        if (x0 == x && y0 == y && x1 == x + 1 && y1 == y + 1) {
            const f3 v = P.F3(F_VX, slot);
            const float c[3] = {v.x, v.y, v.z};
            for (int ch = 0; ch < 3; ch++) {
                if (isfinite(c[ch])) {
                    const long long q = film_fix(c[ch]);
                    bad |= film_overflow(acc[ch], q, ch);
                    acc[ch] = (long long)((unsigned long long)acc[ch] + (unsigned long long)q);
                } else {
                    bad |= 1u << ch;
                }
            }
            w += 1;
        }
    }
    film.sum[3 * p] = (unsigned long long)acc[0];
    film.sum[3 * p + 1] = (unsigned long long)acc[1];
    film.sum[3 * p + 2] = (unsigned long long)acc[2];
    film.weight[p] = w;
    if (bad) film.nan[p] |= bad;
}

// samples on a pixel boundary (a jitter of exactly 0, or x + u rounding up to x + 1) touch up to 4 pixels
__global__ void k_film_edge(dpaths P, dcamera cam, uint32_t n, dfilm film) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    const uint32_t bw = (uint32_t)(cam.bx1 - cam.bx0);
    const uint32_t npix = bw * (uint32_t)(cam.by1 - cam.by0);
    const uint32_t li = slot % npix;
    const int32_t x = cam.bx0 + (int32_t)(li % bw), y = cam.by0 + (int32_t)(li / bw);
    int32_t x0, y0, x1, y1;
    footprint(cam, P.F(F_PX, slot), P.F(F_PY, slot), x0, y0, x1, y1);
    if (x0 == x && y0 == y && x1 == x + 1 && y1 == y + 1) return;
    film_add_span(cam, film, x0, y0, x1, y1, P.F3(F_VX, slot));
}

// pixel::spectrum (film.cpp:20-25): sum / weight; pixels outside the bounds are 0
__global__ void k_film_resolve(dcamera cam, dfilm film, float* rgb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cam.width * cam.height) return;
    film_resolve_pixel(cam, film, p, rgb);
}

// the sampler's draws for the samples of one pixel: per sample, n1 1D draws then n2 2D draws
__global__ void k_sampler_draws(dparams prm, uint32_t pixel, uint32_t n1, uint32_t n2, float* out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= prm.spp) return;
    pcg32 g;
    g.seed(sample_stream(prm.seed, (uint64_t)pixel * prm.spp + s));
    sampler_state smp;
    smp.state = g.state;
    smp.inc = g.inc;
    smp.replay = nullptr;
    smp.end = 0;
    if (prm.sampler == RBW_SAMPLER_STRATIFIED) smp.stratify(prm, pixel, s, 0u);
    float* o = out + (size_t)s * (n1 + 2 * n2);
    for (uint32_t k = 0; k < n1; k++) o[k] = smp.next();
    for (uint32_t k = 0; k < n2; k++) {
        const f2 v = smp.next2();
        o[n1 + 2 * k] = v.x;
        o[n1 + 2 * k + 1] = v.y;
    }
}

// direct_integrator::trace (direct_integrator.cpp:20-128) at the path's hit: emission;
// a surface without lobes lets the ray through (the path stays in the queue, its depth
// unchanged); otherwise emitter_samples light samples and bsdf_samples BSDF samples,
// each with its shadow or emitter-finding ray traced here, weighted by the power
// heuristic over the sample fractions, and the path ends.
template <int W>
__global__ __launch_bounds__(256) void k_direct(dscene S, dpaths P, dparams prm, const uint32_t* qin,
                                               const uint32_t* qin_count, uint32_t* qout, uint32_t* qout_count,
                                               unsigned long long* counters) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *qin_count;
    bool alive = false;
    uint32_t slot = 0, nshadow = 0, nmis = 0, nlight = 0;
    trace_counters cnt{0, 0};
    if (q < count) {
        slot = qin[q];
        const uint32_t prim = P.U(U_HPRIM, slot);
        if (prim != RBW_MISS && S.num_emitters > 0) {
            const f3 o_w = P.F3(F_OX, slot), d_w = P.F3(F_DX, slot);
            uint32_t eid;
            const surf si = hit_surface(S, prim, o_w, d_w, eid);
            const dentity E = S.ents[eid];
            // L + trace(next): the vertices passed through so far are a prefix added to this one's sum
            const f3 prefix = P.F3(F_VX, slot);
            f3 L = f3{0.0f, 0.0f, 0.0f};
            if (E.flags & RBW_ENTITY_EMITTER) L = L + light_evaluate(E, si.normal, -d_w);
            bsdf fn;
            fn.kind = LOBE_NONE;
            if (E.material != RBW_NONE) fn = make_bsdf(S, E.material, prim);
            if (fn.kind == LOBE_NONE) {
                // no scattering functions: straight on (direct_integrator.cpp:52-56)
                f3 no, nd;
                spawn_ray(si.point, si.normal, d_w, no, nd);
                P.setF3(F_OX, slot, no);
                P.setF3(F_DX, slot, nd);
                alive = true;
            } else {
                sampler_state smp = load_sampler(P, prm, slot);
                const uint32_t ne = prm.emitter_samples, nb = prm.bsdf_samples;
                const float fe = (float)ne / (float)(ne + nb), fb = (float)nb / (float)(ne + nb);
                const float we = 1.0f / (float)ne, wb = 1.0f / (float)nb;
                const uint32_t nem = S.num_emitters;
                const f3 wo = world_to_local(si.sh, si.wo);
                for (uint32_t i = 0; i < ne; i++) {  // :58-98
                    const float u1 = smp.next();
                    uint32_t which = (uint32_t)floorf(u1 * (float)nem);
                    if (which > nem - 1) which = nem - 1;
                    const float pick = 1.0f / (float)nem;
                    const uint32_t lid = S.emitters[which];
                    const dentity Lt = S.ents[lid];
                    light_sample ls = sample_light(S, Lt, si.point, si.normal, smp.next2());
                    ls.pdf = ls.pdf * pick;
                    if (is_black(ls.L) || !(ls.pdf > 0.0f)) continue;
                    const f3 wi = world_to_local(si.sh, ls.wi);
                    f3 f = fn.evaluate(wo, wi);
                    const float fp = fn.pdf(wo, wi);
                    f = f * fabsf(dot(ls.wi, si.sh.z));
                    if (is_black(f) || fp <= 0.0f) continue;
                    f3 so, sd;
                    float slen;
                    spawn_ray_to(si.point, si.normal, ls.point, so, sd, slen);
                    nshadow++;
                    hitrec h;
                    const bool any = prm.shadow_mode == RBW_SHADOW_ANY_HIT || (S.light_flags & k_lights_planar);
                    if (any)
                        h = traverse<W, true, true, false, true>(S, so, sd, slen, prm.tri_test, lid, cnt);
                    else
                        h = traverse<W, true, false, false, true>(S, so, sd, slen, prm.tri_test, RBW_NONE, cnt);
                    bool unoccluded = h.prim == RBW_MISS;
                    if (!unoccluded && !any)
                        unoccluded = (h.prim < S.num_tris ? S.tri_entity[h.prim] : S.sphere_entity[h.prim - S.num_tris]) == lid;
                    if (unoccluded) {
                        const float w = power_heuristic(ls.pdf * fe, fp * fb) * we;
                        L = L + f * ls.L * w / ls.pdf;
                    }
                }
                for (uint32_t i = 0; i < nb; i++) {  // :100-128
                    f3 mwi, mf;
                    float mpdf = 0.0f;
                    if (!fn.sample(si, smp.next2(), mwi, mf, mpdf)) continue;
                    mf = mf * fabsf(dot(mwi, si.sh.z));
                    if (is_black(mf) || !(mpdf > 0.0f)) continue;
                    // find_emitter (integrator.cpp:49-82)
                    f3 eo, ed;
                    spawn_ray(si.point, si.normal, mwi, eo, ed);
                    nmis++;
                    const hitrec h = find_emitter_walk<W, false, true>(S, eo, ed, prm.tri_test, cnt);
                    if (h.prim == RBW_MISS) continue;
                    uint32_t he;
                    const surf hs = hit_surface(S, h.prim, eo, ed, he);
                    const dentity Eh = S.ents[he];
                    if (!(Eh.flags & RBW_ENTITY_EMITTER)) continue;
                    nlight++;
                    const float pick = 1.0f / (float)nem;
                    const f3 Le = light_evaluate(Eh, hs.normal, -mwi);
                    const float lp = light_pdf(S, Eh, si.point, si.normal, mwi) * pick;
                    if (is_black(Le) || lp <= 0.0f) continue;
                    const float w = power_heuristic(mpdf * fb, lp * fe) * wb;
                    L = L + mf * Le * w / mpdf;
                }
                store_sampler(P, slot, smp);
            }
            P.setF3(F_VX, slot, prefix + L);
        }
    }
    const uint32_t a = block_append(qout_count, alive);
    if (alive) qout[a] = slot;
    stripe_add64(counters, 2, nshadow);
    stripe_add64(counters, 3, nmis);
    stripe_add64(counters, 4, nlight);
}

// volume_path_integrator::trace (volume_path_integrator.cpp:29-72), one bounce of the
// path in `slot` after its extension hit: distance sampling in the current medium
// (a medium vertex: next-event estimation with the phase function, then a phase
// sample), else the surface vertex of sample_surface_interaction (integrator.cpp:329-404,
// media = true) and the medium switch at entities with media (medium.cpp:24-33);
// Russian roulette as path_integrator.cpp:35-43.
template <int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_volume(dscene S, dpaths P, dparams prm, const uint32_t* qin,
                                               const uint32_t* qin_count, uint32_t* qout, uint32_t* qout_count,
                                               unsigned long long* counters) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *qin_count;
    bool alive = false;
    uint32_t slot = 0, nshadow = 0, nmis = 0, nlight = 0;
    trace_counters cnt{0, 0};
    if (q < count) {
        slot = qin[q];
        const uint32_t prim = P.U(U_HPRIM, slot);
        int bounce = (int)P.U(U_BOUNCE, slot);
        uint32_t medium = P.U(U_MEDIUM, slot);
        f3 value = P.F3(F_VX, slot), beta = P.F3(F_BX, slot);
        const f3 o_w = P.F3(F_OX, slot), d_w = P.F3(F_DX, slot);
        sampler_state smp = load_sampler(P, prm, slot);
        bool cont = true, at_surface = true;
        f3 no = o_w, nd = d_w;
        if (medium != RBW_NONE) {
            // the ray's world length after the hit: the local hit parameter mapped back (entity.cpp:384)
            const float len = prim == RBW_MISS ? __builtin_inff() : P.F(F_HT, slot) * make_local_ray(o_w, d_w).L1;
            const medium_event ev = medium_sample(S, S.media[medium], smp, o_w, d_w, len);
            beta = beta * ev.value;
            if (is_black(beta)) {
                cont = false;
                at_surface = false;
            } else if (ev.scattered) {
                // sample_medium_interaction (integrator.cpp:406-421)
                at_surface = false;
                surf mi;
                mi.point = o_w + d_w * ev.t;
                mi.normal = splat(0.0f);
                mi.wo = -d_w;
                const bsdf none{};
                value = value + beta * volume_nee<W, true>(S, prm, smp, mi, none, ev.g, medium, cnt, nshadow, nmis,
                                                           nlight);
                float pp;
                const f3 wi = hg_sample(mi.wo, ev.g, smp.next2(), pp);
                spawn_ray(mi.point, mi.normal, wi, no, nd);
            }
        }
        if (at_surface) {
            if (prim == RBW_MISS) {
                cont = false;  // no environment emitters
            } else {
                uint32_t eid;
                const surf si = hit_surface(S, prim, o_w, d_w, eid);
                const dentity E = S.ents[eid];
                // emission at the first vertex (integrator.cpp:350-351; no specular lobes exist)
                if (bounce == 0 && (E.flags & RBW_ENTITY_EMITTER)) value = value + beta * light_evaluate(E, si.normal, -d_w);
                if (E.material == RBW_NONE) {
                    // invisible entity: straight through, the bounce does not count (integrator.cpp:355-363)
                    spawn_ray(si.point, si.normal, d_w, no, nd);
                    bounce -= 1;
                } else {
                    const bsdf fn = make_bsdf(S, E.material, prim);
                    if (fn.kind != LOBE_NONE)
                        value = value + beta * volume_nee<W, false>(S, prm, smp, si, fn, 0.0f, medium, cnt, nshadow,
                                                                    nmis, nlight);
                    f3 wi, f;
                    float pdf = 0.0f;
                    const bool ok = fn.sample(si, smp.next2(), wi, f, pdf);
                    if (!ok || is_black(f) || pdf == 0.0f) {
                        cont = false;
                    } else {
                        beta = beta * (f * fabsf(dot(wi, si.sh.z)) / pdf);
                        spawn_ray(si.point, si.normal, wi, no, nd);
                    }
                }
                if (cont && (E.flags & RBW_ENTITY_HAS_MEDIA)) medium = medium_after(S, eid, si.normal, nd);
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
        bounce += 1;
        alive = cont && (uint32_t)bounce < prm.max_depth;
        P.setF3(F_VX, slot, value);
        if (alive) {
            P.setF3(F_BX, slot, beta);
            P.setF3(F_OX, slot, no);
            P.setF3(F_DX, slot, nd);
            P.U(U_BOUNCE, slot) = (uint32_t)bounce;
            P.U(U_MEDIUM, slot) = medium;
            store_sampler(P, slot, smp);
        }
    }
    const uint32_t a = block_append(qout_count, alive);
    if (alive) qout[a] = slot;
    stripe_add64(counters, 2, nshadow);
    stripe_add64(counters, 3, nmis);
    stripe_add64(counters, 4, nlight);
}

// scene::intersect / intersect_with_shadow_ray for caller-given rays
template <int W>
__global__ void k_trace_rays(dscene S, uint32_t n, const float* o, const float* d, const float* tmax, uint32_t visible,
                             uint32_t tri_test, rbw_hit* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 ro = f3{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
    const f3 rd = f3{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
    trace_counters cnt{0, 0};
    const hitrec h = visible ? traverse<W, true, false, false, true>(S, ro, rd, tmax[i], tri_test, RBW_NONE, cnt)
                             : traverse<W, false, false, false, true>(S, ro, rd, tmax[i], tri_test, RBW_NONE, cnt);
    rbw_hit r;
    r.prim = h.prim;
    r.entity = h.prim == RBW_MISS ? RBW_NONE
                                  : (h.prim < S.num_tris ? S.tri_entity[h.prim] : S.sphere_entity[h.prim - S.num_tris]);
    r.t = h.t;
    r.b1 = h.b1;
    r.b2 = h.b2;
    out[i] = r;
}

// ---------------------------------------------------- allocation-free ray queries
// rbw_trace_rays_static stages its rays and hit records here instead of in memory the
// library allocates: origins, directions and tmax of up to RBW_STATIC_MAX_RAYS rays, then
// their hit records (48 bytes per ray, 12 MiB).  Each device holds its own copy of the
// array from the module's load until the process ends.  Host code never passes `scratch`
// itself to a kernel or a copy (there it names a host-side stand-in): it uses the device
// address hipGetSymbolAddress returns for the current device (static_scratch_of).
constexpr size_t STATIC_O = 0;
constexpr size_t STATIC_D = STATIC_O + size_t(RBW_STATIC_MAX_RAYS) * 3 * sizeof(float);
constexpr size_t STATIC_T = STATIC_D + size_t(RBW_STATIC_MAX_RAYS) * 3 * sizeof(float);
constexpr size_t STATIC_H = STATIC_T + size_t(RBW_STATIC_MAX_RAYS) * sizeof(float);
constexpr size_t STATIC_BYTES = STATIC_H + size_t(RBW_STATIC_MAX_RAYS) * sizeof(rbw_hit);
static_assert(STATIC_D % 256 == 0 && STATIC_T % 256 == 0 && STATIC_H % 256 == 0, "aligned sections");
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

// ================================================================== host side

using namespace rbw::host;

namespace {

thread_local std::string g_last_error;

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

}  // namespace

namespace {

// device address of `scratch` per device, resolved once per device with that device
// current; the mutex serialises the calls of this process's threads that share it
struct static_scratch_slot {
    std::once_flag once;
    hipError_t status = hipSuccess;
    unsigned char* ptr = nullptr;
    std::mutex use;
};
constexpr int MAX_DEVICES = 64;
static_scratch_slot g_static_scratch[MAX_DEVICES];

// call with `device` current
static_scratch_slot& static_scratch_of(int device) {
    require(device >= 0 && device < MAX_DEVICES, RBW_ERR_UNSUPPORTED, "device ordinal above the static table");
    static_scratch_slot& sl = g_static_scratch[device];
    std::call_once(sl.once, [&] {
        void* p = nullptr;
        sl.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        sl.ptr = static_cast<unsigned char*>(p);
    });
    if (sl.status != hipSuccess || !sl.ptr)
        throw rbw_error(RBW_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(sl.status));
    return sl;
}

}  // namespace

void rbw_ctx::ensure_guide() {
    if (!guide_ready) {
        guide::model_init(gm, 0x5eed, stream);
        guide_ready = true;
    }
}

rbw_ctx::~rbw_ctx() {
    (void)hipSetDevice(device);
    batch_free(this);
    guide::model_free(gm);
    guide::model_free(gjobs);
    if (host_count) (void)hipHostFree(host_count);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

void upload_scene(rbw_ctx* ctx, const rbw_scene_desc* d) {
    host_scene h;
    build_host_scene(d, h);
    ctx->has_scene = false;
    ctx->main.upload(h, ctx->stream);
    ctx->scene_min = h.smin;
    ctx->scene_max = h.smax;
    ctx->bvh_depth = h.depth;
    ctx->bvh_nodes = uint32_t(h.nodes.size() / 4);
    ctx->has_scene = true;
}


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
    require(pr.sampler == RBW_SAMPLER_PER_SAMPLE_PCG || pr.sampler == RBW_SAMPLER_REPLAY ||
                pr.sampler == RBW_SAMPLER_STRATIFIED,
            RBW_ERR_INVALID_ARGUMENT, "unknown sampler");
    if (pr.sampler == RBW_SAMPLER_STRATIFIED)
        require(pr.strata_x > 0 && pr.strata_y > 0 && uint64_t(pr.strata_x) * pr.strata_y == pr.spp &&
                    pr.dims_1d <= 0xFFFFu && pr.dims_2d <= 0xFFFFu,
                RBW_ERR_INVALID_ARGUMENT, "stratified sampler: strata_x * strata_y must equal spp, dims below 65536");
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
        require(pr.sampler != RBW_SAMPLER_REPLAY, RBW_ERR_INVALID_ARGUMENT,
                "guiding draws a different number of samples per vertex than the replayed reference stream");
    }
    require(pr.integrator == RBW_INTEGRATOR_PATH || pr.integrator == RBW_INTEGRATOR_DIRECT ||
                pr.integrator == RBW_INTEGRATOR_VOLUME_PATH,
            RBW_ERR_INVALID_ARGUMENT, "unknown integrator");
    const bool direct = pr.integrator == RBW_INTEGRATOR_DIRECT;
    const bool volume = pr.integrator == RBW_INTEGRATOR_VOLUME_PATH;
    if (direct)
        require(!guided && !pr.debug_hits && !pr.count_visits, RBW_ERR_INVALID_ARGUMENT,
                "the direct integrator takes no guiding, debug hits or visit counts");
    if (volume)
        require(!guided && !pr.debug_hits && !pr.count_visits && pr.sampler != RBW_SAMPLER_REPLAY,
                RBW_ERR_INVALID_ARGUMENT,
                "the volume path integrator takes no guiding, replay, debug hits or visit counts");
    require(pr.schedule == RBW_SCHEDULE_AUTO || pr.schedule == RBW_SCHEDULE_PERSISTENT ||
                pr.schedule == RBW_SCHEDULE_PASSES,
            RBW_ERR_INVALID_ARGUMENT, "unknown schedule");
    // the plain estimator runs on persistent, in-place regenerated path slots
    const bool can_persist =
        pr.sampler != RBW_SAMPLER_REPLAY && !pr.debug_hits && !pr.count_visits && !guided && !direct && !volume;
    require(pr.schedule != RBW_SCHEDULE_PERSISTENT || can_persist, RBW_ERR_INVALID_ARGUMENT,
            "persistent path slots take the plain estimator only (no replay, debug hits, visit counts, guiding, "
            "direct or volume integrator)");
    // automatic: persistent slots from 2^22 camera samples on; below, passes with compacted
    // queues (a render that fills the pool once drains it in full-pool iterations; measured
    // on the 512^2 narrow-aperture room, 8 spp: passes 9.9 ms, persistent 11.1; 16 spp:
    // 19.7 and 18.6; 128 spp: 158 and 118)
    const bool persist = can_persist && (pr.schedule == RBW_SCHEDULE_PERSISTENT ||
                                         (pr.schedule == RBW_SCHEDULE_AUTO && npix * pr.spp >= (1ull << 22)));
    if (persist) {
        render_scheduled(ctx, cam, pr, out_rgb, stats);
        return;
    }
    uint32_t k = pr.samples_per_pass;
    // passes of ~2^21 paths keep every bounce's launches wide enough to hide the
    // traversal's memory latency (guided renders train once per pass)
    const uint64_t per_pass = 1u << 21;
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
    constexpr size_t n_counter_words = size_t(RBW_STRIPES) * RBW_STRIPE_WORDS;
    ctx->counters.ensure(n_counter_words * sizeof(unsigned long long));
    ctx->film_sum.ensure(size_t(W) * H * 3 * sizeof(unsigned long long));
    ctx->film_w.ensure(size_t(W) * H * sizeof(uint32_t));
    ctx->film_nan.ensure(size_t(W) * H * sizeof(uint32_t));
    ctx->rgb.ensure(size_t(W) * H * 3 * sizeof(float));
    HIP_CHECK(hipMemsetAsync(ctx->film_sum.p, 0, size_t(W) * H * 3 * sizeof(unsigned long long), s));
    HIP_CHECK(hipMemsetAsync(ctx->film_w.p, 0, size_t(W) * H * sizeof(uint32_t), s));
    HIP_CHECK(hipMemsetAsync(ctx->film_nan.p, 0, size_t(W) * H * sizeof(uint32_t), s));
    dfilm film;
    film.sum = ctx->film_sum.as<unsigned long long>();
    film.weight = ctx->film_w.as<uint32_t>();
    film.nan = ctx->film_nan.as<uint32_t>();
    HIP_CHECK(hipMemsetAsync(ctx->counters.p, 0, n_counter_words * sizeof(unsigned long long), s));
    dparams dp;
    dp.spp = pr.spp;
    dp.max_depth = pr.max_depth;
    dp.rr_threshold = pr.rr_threshold;
    dp.seed = pr.seed;
    dp.sampler = pr.sampler;
    dp.strata_x = pr.strata_x;
    dp.strata_y = pr.strata_y;
    dp.dims_1d = pr.dims_1d;
    dp.dims_2d = pr.dims_2d;
    dp.emitter_samples = pr.emitter_samples ? pr.emitter_samples : 1u;
    dp.bsdf_samples = pr.bsdf_samples ? pr.bsdf_samples : 1u;
    dp.medium_mis = pr.medium_mis ? 1u : 0u;
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
        dp.g.light_records = go->light_records ? 1u : 0u;
        dp.g.depth = go->guide_depth ? go->guide_depth : pr.max_depth;
        dp.g.product = go->product_targets ? 1u : 0u;
        dp.g.rec_depth = pr.max_depth;
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
                ctx->g_rec.ensure(size_t(R_STRIDE) * nrec * sizeof(float));
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
    const dscene S = ctx->main.view(resolve_width(pr.bvh_width));
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
        hipLaunchKernelGGL(k_raygen, dim3(blocks(n, BS)), dim3(BS), 0, s, S, P, dp, dc, n, s0, qa, counts + 0,
                           volume ? 1u : 0u);
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
            {
                auto kx = S.width == 8   ? (count ? k_extend<8, true> : k_extend<8, false>)
                          : S.width == 4 ? (count ? k_extend<4, true> : k_extend<4, false>)
                                         : (count ? k_extend<2, true> : k_extend<2, false>);
                hipLaunchKernelGGL(kx, dim3(blocks(active, BS)), dim3(BS), lds_stack_bytes(BS), s, S,
                                   P, dp, qin, counts + cin, counters);
            }
            if (timed) ctx->timer.end(s);
            if (direct || volume) {
                // the direct integrator's vertex, or the volume path integrator's bounce, with
                // their shadow, emitter-finding and beam rays in one kernel
                if (timed) ctx->timer.begin(RBW_KERNEL_CONNECT, s);
                if (direct)
                    hipLaunchKernelGGL(S.width == 8 ? k_direct<8> : (S.width == 4 ? k_direct<4> : k_direct<2>),
                                       dim3(blocks(active, BS)), dim3(BS),
                                       lds_stack_bytes(BS), s, S, P, dp, qin, counts + cin, qout, counts + cout,
                                       counters);
                else
                    hipLaunchKernelGGL(S.width == 8 ? k_volume<8> : (S.width == 4 ? k_volume<4> : k_volume<2>),
                                       dim3(blocks(active, BS)), dim3(BS),
                                       lds_stack_bytes(BS), s, S, P, dp, qin, counts + cin, qout, counts + cout,
                                       counters);
                if (timed) ctx->timer.end(s);
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipMemcpyAsync(ctx->host_count, counts + cout, 4, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                active = *ctx->host_count;
                std::swap(qin, qout);
                std::swap(cin, cout);
                continue;
            }
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
                // the guided rows (at most `active`) are counted on the device: the MLP and the
                // guided vertex kernel read the count there, no host round trip
                if (timed) ctx->timer.begin(RBW_KERNEL_GUIDE_INFER, s);
                guide::forward_dev(ctx->gm, dp.g.query, counts + 3, active, dp.g.mix, s);
                if (timed) ctx->timer.end(s);
                if (timed) ctx->timer.begin(RBW_KERNEL_SHADE_GUIDED, s);
                hipLaunchKernelGGL(k_shade_guided, dim3(blocks(active, BS)), dim3(BS), 0, s, S, P, dp, qout,
                                   counts + cout, cqp, counts + 2);
                if (timed) ctx->timer.end(s);
                HIP_CHECK(hipGetLastError());
            }
            if (timed) ctx->timer.begin(RBW_KERNEL_CONNECT, s);
            {
                auto kc = S.width == 8   ? (count ? k_connect<8, true> : k_connect<8, false>)
                          : S.width == 4 ? (count ? k_connect<4, true> : k_connect<4, false>)
                                         : (count ? k_connect<2, true> : k_connect<2, false>);
                hipLaunchKernelGGL(kc, dim3(blocks(active, BS)), dim3(BS), lds_stack_bytes(BS), s, S,
                                   P, dp, cqp, counts + 2, counters);
            }
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
            hipLaunchKernelGGL(k_collect, dim3(blocks(uint64_t(n) * pr.max_depth, 256 * COLLECT_PER_THREAD)), dim3(256), 0,
                               s, P, dp, n);
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
        hipLaunchKernelGGL(k_film_own, dim3(blocks(npix, BS)), dim3(BS), 0, s, P, dc, kk, film);
        hipLaunchKernelGGL(k_film_edge, dim3(blocks(n, BS)), dim3(BS), 0, s, P, dc, n, film);
        if (timed) ctx->timer.end(s);
        HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_film_resolve, dim3(blocks(uint64_t(W) * H, BS)), dim3(BS), 0, s, dc, film,
                       ctx->rgb.as<float>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev1, s));
    if (out_rgb)
        HIP_CHECK(hipMemcpyAsync(out_rgb, ctx->rgb.p, size_t(W) * H * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    if (ndbg) HIP_CHECK(hipMemcpyAsync(pr.debug_hits, ctx->dbg.p, ndbg * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> stripes(n_counter_words);
    HIP_CHECK(hipMemcpyAsync(stripes.data(), counters, n_counter_words * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, s));
    float gloss = 0.0f;
    if (guide_steps) HIP_CHECK(hipMemcpyAsync(&gloss, ctx->gm.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    unsigned long long hc[RBW_STRIPE_WORDS] = {};
    for (uint32_t t = 0; t < RBW_STRIPES; t++)
        for (uint32_t k = 0; k < RBW_STRIPE_WORDS; k++) hc[k] += stripes[t * RBW_STRIPE_WORDS + k];
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        stats->render_ms = ms;
        stats->bvh_width = S.width;
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
        (void)static_scratch_of(dev);  // the static ray buffers' device address, once per device
        *out = ctx.release();
    });
}

void rbw_destroy(rbw_ctx* ctx) { delete ctx; }

const char* rbw_last_error(void) { return g_last_error.c_str(); }

rbw_status rbw_sampler_draws(rbw_ctx* ctx, const rbw_render_params* params, uint32_t pixel, uint32_t n1,
                             uint32_t n2, float* out) {
    return guarded([&] {
        require(ctx && params && out, RBW_ERR_INVALID_ARGUMENT, "ctx, params and out are required");
        const rbw_render_params& pr = *params;
        require(pr.sampler == RBW_SAMPLER_PER_SAMPLE_PCG || pr.sampler == RBW_SAMPLER_STRATIFIED,
                RBW_ERR_INVALID_ARGUMENT, "sampler draws exist for the per-sample and stratified samplers");
        require(pr.spp > 0 && pr.spp <= (1u << 20) && n1 <= 0xFFFFu && n2 <= 0xFFFFu, RBW_ERR_INVALID_ARGUMENT,
                "spp or draw counts out of range");
        if (pr.sampler == RBW_SAMPLER_STRATIFIED)
            require(pr.strata_x > 0 && pr.strata_y > 0 && uint64_t(pr.strata_x) * pr.strata_y == pr.spp,
                    RBW_ERR_INVALID_ARGUMENT, "stratified sampler: strata_x * strata_y must equal spp");
        dparams dp;
        std::memset(&dp, 0, sizeof(dp));
        dp.spp = pr.spp;
        dp.seed = pr.seed;
        dp.sampler = pr.sampler;
        dp.strata_x = pr.strata_x;
        dp.strata_y = pr.strata_y;
        dp.dims_1d = pr.dims_1d;
        dp.dims_2d = pr.dims_2d;
        const size_t n = size_t(pr.spp) * (n1 + 2 * size_t(n2));
        dbuf d;
        d.ensure(std::max<size_t>(n, 1) * sizeof(float));
        hipStream_t s = ctx->stream;
        hipLaunchKernelGGL(k_sampler_draws, dim3(blocks(pr.spp, 256)), dim3(256), 0, s, dp, pixel, n1, n2,
                           d.as<float>());
        HIP_CHECK(hipGetLastError());
        if (n) HIP_CHECK(hipMemcpyAsync(out, d.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    });
}

rbw_status rbw_film_to_rgba8(const float* rgb, uint32_t width, uint32_t height, const int32_t* bound_min,
                             const int32_t* bound_max, float scale, uint8_t* out_rgba) {
    return guarded([&] {
        require(rgb && out_rgba, RBW_ERR_INVALID_ARGUMENT, "rgb and out_rgba are required");
        rbw::image::film_to_rgba8(rgb, width, height, bound_min, bound_max, scale, out_rgba);
    });
}

rbw_status rbw_write_png(const char* path, const float* rgb, uint32_t width, uint32_t height,
                         const int32_t* bound_min, const int32_t* bound_max, float scale) {
    return guarded([&] {
        require(path && rgb, RBW_ERR_INVALID_ARGUMENT, "path and rgb are required");
        require(width > 0 && height > 0 && width <= (1u << 16) && height <= (1u << 16), RBW_ERR_INVALID_ARGUMENT,
                "image size out of range");
        std::vector<uint8_t> rgba(size_t(width) * height * 4);
        rbw::image::film_to_rgba8(rgb, width, height, bound_min, bound_max, scale, rgba.data());
        rbw::image::write_png(path, rgba.data(), width, height);
    });
}

rbw_status rbw_write_pfm(const char* path, const float* rgb, uint32_t width, uint32_t height) {
    return guarded([&] {
        require(path && rgb, RBW_ERR_INVALID_ARGUMENT, "path and rgb are required");
        require(width > 0 && height > 0 && width <= (1u << 16) && height <= (1u << 16), RBW_ERR_INVALID_ARGUMENT,
                "image size out of range");
        rbw::image::write_pfm(path, rgb, width, height);
    });
}

rbw_status rbw_upload_scene(rbw_ctx* ctx, const rbw_scene_desc* scene) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->has_scene = false;
        upload_scene(ctx, scene);
    });
}

rbw_status rbw_add_scene(rbw_ctx* ctx, const rbw_scene_desc* scene, uint32_t* scene_id) {
    return guarded([&] {
        require(ctx != nullptr && scene_id != nullptr, RBW_ERR_INVALID_ARGUMENT, "context or scene_id is null");
        auto h = std::make_unique<host_scene>();
        build_host_scene(scene, *h);
        *scene_id = uint32_t(ctx->pool.size());
        ctx->pool.push_back(std::move(h));
    });
}

rbw_status rbw_remove_scene(rbw_ctx* ctx, uint32_t scene_id) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(scene_id < ctx->pool.size() && ctx->pool[scene_id], RBW_ERR_INVALID_ARGUMENT, "unknown scene id");
        ctx->pool[scene_id].reset();
    });
}

rbw_status rbw_submit_jobs(rbw_ctx* ctx, const rbw_job* jobs, uint32_t n, const rbw_batch_opts* opts,
                           rbw_job_stats* job_stats, rbw_batch_stats* stats) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        submit_jobs(ctx, jobs, n, opts, job_stats, stats);
    });
}

rbw_status rbw_autotune_jobs(rbw_ctx* ctx, const rbw_job* jobs, uint32_t n, rbw_batch_opts* opts, double* probe_ms) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        autotune_jobs(ctx, jobs, n, opts, probe_ms);
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
                          uint32_t visible_only, uint32_t triangle_test, uint32_t bvh_width, rbw_hit* out) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(ctx->has_scene, RBW_ERR_NO_SCENE, "no scene uploaded");
        require(n == 0 || (origins && directions && tmax && out), RBW_ERR_INVALID_ARGUMENT, "null ray arrays");
        require(triangle_test == RBW_TRI_MOLLER_TRUMBORE || triangle_test == RBW_TRI_WATERTIGHT,
                RBW_ERR_INVALID_ARGUMENT, "unknown triangle test");
        const uint32_t width = resolve_width(bvh_width);
        if (n == 0) return;
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        dbuf o, d, t, h;
        try {
            o.upload(origins, size_t(n) * 3, s);
            d.upload(directions, size_t(n) * 3, s);
            t.upload(tmax, n, s);
            h.ensure(size_t(n) * sizeof(rbw_hit));
            hipLaunchKernelGGL(width == 8 ? k_trace_rays<8> : (width == 4 ? k_trace_rays<4> : k_trace_rays<2>), dim3(blocks(n, 256)), dim3(256), lds_stack_bytes(256), s,
                               ctx->main.view(width), n, o.as<float>(),
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

rbw_status rbw_trace_rays_static(rbw_ctx* ctx, uint32_t n, const float* origins, const float* directions,
                                 const float* tmax, uint32_t visible_only, uint32_t triangle_test, uint32_t bvh_width,
                                 rbw_hit* out) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(ctx->has_scene, RBW_ERR_NO_SCENE, "no scene uploaded");
        require(n == 0 || (origins && directions && tmax && out), RBW_ERR_INVALID_ARGUMENT, "null ray arrays");
        require(n <= RBW_STATIC_MAX_RAYS, RBW_ERR_INVALID_ARGUMENT,
                "rbw_trace_rays_static takes at most RBW_STATIC_MAX_RAYS (262144) rays per call");
        require(triangle_test == RBW_TRI_MOLLER_TRUMBORE || triangle_test == RBW_TRI_WATERTIGHT,
                RBW_ERR_INVALID_ARGUMENT, "unknown triangle test");
        const uint32_t width = resolve_width(bvh_width);
        if (n == 0) return;
        HIP_CHECK(hipSetDevice(ctx->device));
        static_scratch_slot& sl = static_scratch_of(ctx->device);
        std::lock_guard<std::mutex> lock(sl.use);
        hipStream_t s = ctx->stream;
        float* o = reinterpret_cast<float*>(sl.ptr + STATIC_O);
        float* d = reinterpret_cast<float*>(sl.ptr + STATIC_D);
        float* t = reinterpret_cast<float*>(sl.ptr + STATIC_T);
        rbw_hit* h = reinterpret_cast<rbw_hit*>(sl.ptr + STATIC_H);
        try {
            HIP_CHECK(hipMemcpyAsync(o, origins, size_t(n) * 3 * sizeof(float), hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(d, directions, size_t(n) * 3 * sizeof(float), hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(t, tmax, size_t(n) * sizeof(float), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(width == 8 ? k_trace_rays<8> : (width == 4 ? k_trace_rays<4> : k_trace_rays<2>), dim3(blocks(n, 256)), dim3(256),
                               lds_stack_bytes(256), s, ctx->main.view(width), n, o, d, t, visible_only, triangle_test,
                               h);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(out, h, size_t(n) * sizeof(rbw_hit), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);  // the buffers are shared: nothing of this call stays in flight
            throw;
        }
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
        ctx->timer.reset();
        ctx->timer.begin(RBW_KERNEL_GUIDE_INFER, s);
        guide::forward(ctx->gm, ctx->g_io_in.as<float>(), n, ctx->g_io_out.as<float>(), s);
        ctx->timer.end(s);
        HIP_CHECK(hipMemcpyAsync(out, ctx->g_io_out.p, size_t(n) * guide::N_OUT_PAD * sizeof(float),
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        ctx->guide_last_ms = ctx->timer.elapsed_last();
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
        ctx->timer.reset();
        ctx->timer.begin(RBW_KERNEL_GUIDE_TRAIN, s);
        guide::adam_opts ao;
        ao.lr = learning_rate;
        // gradient (and, with apply, the Adam step) in three launches; the network's mean
        // gradient and loss stay in the model
        guide::train_step(ctx->gm, ctx->g_io_in.as<float>(), ctx->g_io_dir.as<float>(), ctx->g_io_c.as<float>(), n,
                          ao, apply != 0, s);
        if (grad) HIP_CHECK(hipMemcpyAsync(grad, ctx->gm.grad, sizeof(float) * guide::N_PARAMS, hipMemcpyDeviceToHost, s));
        if (loss) HIP_CHECK(hipMemcpyAsync(loss, ctx->gm.loss, sizeof(float), hipMemcpyDeviceToHost, s));
        ctx->timer.end(s);
        HIP_CHECK(hipStreamSynchronize(s));
        ctx->guide_last_ms = ctx->timer.elapsed_last();
    });
}

rbw_status rbw_guide_jobs_reset(rbw_ctx* ctx, uint32_t n_networks, uint32_t seed) {
    return guarded([&] {
        require(ctx != nullptr, RBW_ERR_INVALID_ARGUMENT, "context is null");
        require(n_networks >= 1 && n_networks <= guide::MAX_NETS, RBW_ERR_INVALID_ARGUMENT,
                "n_networks must be in [1, 32]");
        HIP_CHECK(hipSetDevice(ctx->device));
        guide::model_init(ctx->gjobs, seed, ctx->stream, n_networks);
    });
}

rbw_status rbw_guide_jobs_get_params(rbw_ctx* ctx, uint32_t network, float* params) {
    return guarded([&] {
        require(ctx != nullptr && params != nullptr, RBW_ERR_INVALID_ARGUMENT, "null argument");
        require(network < ctx->gjobs.nets, RBW_ERR_INVALID_ARGUMENT, "no such job network");
        HIP_CHECK(hipSetDevice(ctx->device));
        guide::model_get(ctx->gjobs, params, ctx->stream, network);
    });
}

rbw_status rbw_guide_jobs_set_params(rbw_ctx* ctx, uint32_t network, const float* params) {
    return guarded([&] {
        require(ctx != nullptr && params != nullptr, RBW_ERR_INVALID_ARGUMENT, "null argument");
        require(network < ctx->gjobs.nets, RBW_ERR_INVALID_ARGUMENT, "no such job network");
        HIP_CHECK(hipSetDevice(ctx->device));
        guide::model_set(ctx->gjobs, params, ctx->stream, network);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

rbw_status rbw_guide_jobs_train(rbw_ctx* ctx, uint32_t n_networks, const uint32_t* rows, const float* inputs,
                                const float* dirs, const float* c, float learning_rate, uint32_t apply, float* loss,
                                float* grad) {
    return guarded([&] {
        require(ctx != nullptr && rows != nullptr, RBW_ERR_INVALID_ARGUMENT, "null argument");
        require(n_networks >= 1 && n_networks <= ctx->gjobs.nets, RBW_ERR_INVALID_ARGUMENT,
                "n_networks must be in [1, the job networks]");
        std::vector<guide::train_rows> tr(ctx->gjobs.nets);
        uint64_t total = 0;
        for (uint32_t k = 0; k < n_networks; k++) {
            tr[k].base = uint32_t(total);
            tr[k].n = rows[k];
            total += rows[k];
        }
        require(total > 0 && total <= (1u << 22), RBW_ERR_INVALID_ARGUMENT, "1 to 2^22 rows in all");
        require(inputs && dirs && c, RBW_ERR_INVALID_ARGUMENT, "null arrays");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        ctx->g_io_in.upload(inputs, size_t(total) * 9, s);
        ctx->g_io_dir.upload(dirs, size_t(total) * 3, s);
        ctx->g_io_c.upload(c, total, s);
        ctx->timer.reset();
        ctx->timer.begin(RBW_KERNEL_GUIDE_TRAIN, s);
        guide::adam_opts ao;
        ao.lr = learning_rate;
        guide::train_step_batched(ctx->gjobs, ctx->g_io_in.as<float>(), ctx->g_io_dir.as<float>(),
                                  ctx->g_io_c.as<float>(), tr, ao, apply != 0, s);
        ctx->timer.end(s);
        if (grad)
            HIP_CHECK(hipMemcpyAsync(grad, ctx->gjobs.grad, sizeof(float) * guide::N_PARAMS * n_networks,
                                     hipMemcpyDeviceToHost, s));
        if (loss) HIP_CHECK(hipMemcpyAsync(loss, ctx->gjobs.loss, sizeof(float) * n_networks, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        ctx->guide_last_ms = ctx->timer.elapsed_last();
    });
}

rbw_status rbw_guide_last_ms(rbw_ctx* ctx, double* ms) {
    return guarded([&] {
        require(ctx != nullptr && ms != nullptr, RBW_ERR_INVALID_ARGUMENT, "null arguments");
        *ms = ctx->guide_last_ms;
    });
}

}  // extern "C"
