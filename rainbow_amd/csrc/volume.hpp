// The volume path integrator's vertex work (volume_path_integrator.cpp:13-79 with
// integrator.cpp:101-291, 329-421): beam transmittance through invisible medium
// boundaries (scene.cpp:114-148) and next-event estimation at surface and medium
// vertices, each ray traced where it is needed.
#pragma once

#include "media.hpp"
#include "path.hpp"

namespace rbw {

constexpr uint32_t k_beam_max_segments = 256;  // bound on boundary crossings of one beam

__device__ __forceinline__ uint32_t prim_entity(const dscene& S, uint32_t prim) {
    return prim < S.num_tris ? S.tri_entity[prim] : S.sphere_entity[prim - S.num_tris];
}

// scene::evaluate_media_beam (scene.cpp:114-148): the transmittance from (point, normal)
// to `to` through the media the beam crosses; 0 when a visible entity is in the way
template <int W>
__device__ __forceinline__ f3 media_beam(const dscene& S, const dparams& prm, sampler_state& smp, uint32_t medium,
                                         f3 point, f3 normal, f3 to, trace_counters& cnt, uint32_t& nrays) {
    f3 L = splat(1.0f);
    f3 o, d;
    float len;
    spawn_ray_to(point, normal, to, o, d, len);
    for (uint32_t seg = 0; seg < k_beam_max_segments; seg++) {
        nrays++;
        const hitrec h = traverse<W, false, false, false, true>(S, o, d, len, prm.tri_test, RBW_NONE, cnt);
        if (h.prim == RBW_MISS) break;
        // (the segment's world length: the local hit parameter mapped back, entity.cpp:384)
        if (medium != RBW_NONE) L = L * medium_transmittance(S, S.media[medium], smp, o, d, h.t * make_local_ray(o, d).L1);
        uint32_t eid;
        const surf hs = hit_surface(S, h.prim, o, d, eid);
        const dentity E = S.ents[eid];
        if (E.material != RBW_NONE) return splat(0.0f);
        spawn_ray_to(hs.point, hs.normal, to, o, d, len);
        if (E.flags & RBW_ENTITY_HAS_MEDIA) medium = medium_after(S, eid, hs.normal, d);
    }
    if (medium != RBW_NONE) L = L * medium_transmittance(S, S.media[medium], smp, o, d, len);
    return L;
}

// uniform_sample_one_emitter with media: the surface overload (integrator.cpp:101-201,
// media = true) when MEDIUM is false, the medium overload (integrator.cpp:203-291) with
// the Henyey-Greenstein phase function of asymmetry g when it is true.  `medium` is the
// medium the path travels in at the vertex.
template <int W, bool MEDIUM>
__device__ __forceinline__ f3 volume_nee(const dscene& S, const dparams& prm, sampler_state& smp, const surf& si,
                                         const bsdf& fn, float g, uint32_t medium, trace_counters& cnt,
                                         uint32_t& nshadow, uint32_t& nmis, uint32_t& nlight) {
    f3 L = splat(0.0f);
    const uint32_t ne = S.num_emitters;
    const float pick = 1.0f / (float)ne;
    {
        const float u1 = smp.next();
        uint32_t which = (uint32_t)floorf(u1 * (float)ne);
        if (which > ne - 1) which = ne - 1;
        const uint32_t lid = S.emitters[which];
        const dentity Lt = S.ents[lid];
        light_sample ls = sample_light(S, Lt, si.point, si.normal, smp.next2());
        ls.pdf = ls.pdf * pick;
        if (!is_black(ls.L) && ls.pdf > 0.0f) {
            f3 f;
            float fp;
            if (MEDIUM) {
                fp = hg_phase(dot(si.wo, ls.wi), g);
                f = splat(fp);
            } else {
                const f3 wi = world_to_local(si.sh, ls.wi);
                const f3 wo = world_to_local(si.sh, si.wo);
                f = fn.evaluate(wo, wi);
                fp = fn.pdf(wo, wi);
                f = f * fabsf(dot(ls.wi, si.sh.z));
            }
            if (!is_black(f) && fp > 0.0f) {
                f3 so, sd;
                float slen;
                spawn_ray_to(si.point, si.normal, ls.point, so, sd, slen);
                nshadow++;
                // (a planar light: the any-hit walk gives the closest-visible answer, as connect_path)
                const bool any = (S.light_flags & k_lights_planar) != 0;
                const hitrec h = any ? traverse<W, true, true, false, true>(S, so, sd, slen, prm.tri_test, lid, cnt)
                                     : traverse<W, true, false, false, true>(S, so, sd, slen, prm.tri_test, RBW_NONE, cnt);
                if (h.prim == RBW_MISS || (!any && prim_entity(S, h.prim) == lid)) {
                    const f3 beam = media_beam<W>(S, prm, smp, medium, si.point, si.normal, ls.point, cnt, nshadow);
                    const float w = power_heuristic(ls.pdf, fp);
                    if (MEDIUM) {
                        L = L + f * ls.L * beam * w / ls.pdf;
                    } else {
                        f = f * beam;
                        L = L + f * ls.L * w / ls.pdf;
                    }
                }
            }
        }
    }
    {
        f3 mwi, mf;
        float mpdf = 0.0f;
        bool ok;
        if (MEDIUM) {
            mwi = hg_sample(si.wo, g, smp.next2(), mpdf);
            mf = splat(mpdf);
            ok = true;
        } else {
            ok = fn.sample(si, smp.next2(), mwi, mf, mpdf);
            if (ok) mf = mf * fabsf(dot(mwi, si.sh.z));
        }
        // at a medium vertex with the reference's weighting an emitter hit never contributes
        // (its light pdf is not a number, see rbw_render_params.medium_mis): no walk needed
        const bool walk = !MEDIUM || prm.medium_mis;
        if (ok && !is_black(mf) && mpdf > 0.0f) nmis++;
        if (walk && ok && !is_black(mf) && mpdf > 0.0f) {
            // find_emitter (integrator.cpp:49-82)
            f3 eo, ed;
            spawn_ray(si.point, si.normal, mwi, eo, ed);
            const hitrec h = find_emitter_walk<W, false, true>(S, eo, ed, prm.tri_test, cnt);
            if (h.prim != RBW_MISS) {
                uint32_t he;
                const surf hs = hit_surface(S, h.prim, eo, ed, he);
                const dentity Eh = S.ents[he];
                if (Eh.flags & RBW_ENTITY_EMITTER) {
                    const f3 Le = light_evaluate(Eh, hs.normal, -mwi);
                    nlight++;
                    const float lp = light_pdf(S, Eh, si.point, si.normal, mwi, MEDIUM) * pick;
                    if (!is_black(Le) && lp > 0.0f) {
                        const f3 beam = media_beam<W>(S, prm, smp, medium, si.point, si.normal, hs.point, cnt, nshadow);
                        const float w = power_heuristic(mpdf, lp);
                        if (MEDIUM) {
                            L = L + mf * Le * beam * w / mpdf;
                        } else {
                            mf = mf * beam;
                            L = L + mf * Le * w / mpdf;
                        }
                    }
                }
            }
        }
    }
    return L;
}

}  // namespace rbw
