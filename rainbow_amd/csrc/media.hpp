// Participating media for the volume path integrator: distance sampling and
// transmittance of the homogeneous medium (media/homogeneous_medium.cpp:12-58)
// and of the density grid (media/heterogeneous_medium.cpp:24-120: delta tracking
// to sample, ratio tracking to evaluate, trilinear lookup as
// textures/detail/image_texture.hpp:119-162), and the Henyey-Greenstein phase
// function (shared/phases/henyey_greenstein.cpp:11-48, utilities.cpp:46-51).  log and exp
// are dmath.hpp's flog / fexp (one fixed float sequence, shared with the oracle), sin and
// cos the correctly rounded values via the double routines.
#pragma once

#include "device.hpp"
#include "shading.hpp"

namespace rbw {

constexpr float k_one_over_four_pi = 0.0795774715459476678844f;

// henyey_greenstein_phase_function (utilities.cpp:46-51)
RBW_HD float hg_phase(float cos_theta, float g) {
    const float den = 1.0f + g * g + 2.0f * g * cos_theta;
    return k_one_over_four_pi * (1.0f - g * g) / (den * sqrtf(den));
}

// henyey_greenstein::sample (henyey_greenstein.cpp:17-48): cos theta about -wo, the frame
// of coordinate_system(wo) with its z replaced by -wo; the pdf is the phase value
RBW_HD f3 hg_sample(f3 wo, float g, f2 u, float& pdf) {
    float ct;
    if (fabsf(g) >= 1e-3f) {
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u.x);
        ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    } else {
        ct = 1.0f - 2.0f * u.x;
    }
    const float st = sqrtf(gmax(0.0f, 1.0f - ct * ct));
    const float phi = k_two_pi * u.y;
    frame fr = frame_from_normal(wo);
    fr.z = -wo;
    pdf = hg_phase(-ct, g);
    return local_to_world(fr, mk(st * rcos(phi), st * rsin(phi), ct));
}

// the medium's local ray of a grid: direction not re-normalised (heterogeneous_medium.cpp:28-33)
RBW_HD f3 w2l_point(const dmedium& M, f3 p) {
    return mk(M.w2l[0] * p.x + M.w2l[1] * p.y + M.w2l[2] * p.z + M.w2l[3],
              M.w2l[4] * p.x + M.w2l[5] * p.y + M.w2l[6] * p.z + M.w2l[7],
              M.w2l[8] * p.x + M.w2l[9] * p.y + M.w2l[10] * p.z + M.w2l[11]);
}
RBW_HD f3 w2l_vector(const dmedium& M, f3 v) {
    return mk(M.w2l[0] * v.x + M.w2l[1] * v.y + M.w2l[2] * v.z, M.w2l[4] * v.x + M.w2l[5] * v.y + M.w2l[6] * v.z,
              M.w2l[8] * v.x + M.w2l[9] * v.y + M.w2l[10] * v.z);
}

// heterogeneous_medium::intersect (heterogeneous_medium.cpp:122-140): the unit cube's slab
// interval within [0, len]
RBW_HD bool unit_box_interval(f3 o, f3 d, float len, float& t0, float& t1) {
    t0 = 0.0f;
    t1 = len;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float inv = 1.0f / dd[k];
        float n = (0.0f - oo[k]) * inv, f = (1.0f - oo[k]) * inv;
        if (n > f) {
            const float tmp = n;
            n = f;
            f = tmp;
        }
        t0 = gmax(t0, n);
        t1 = gmin(t1, f);
        if (t0 > t1) return false;
    }
    return true;
}

// image_texture_t<3>::sample (image_texture.hpp:119-162): wrap into [0, 1), texel centres
// at (i + 0.5) / n, clamped neighbours, lerp s * (1 - x) + t * x along x, then y, then z
__device__ __forceinline__ float grid_density(const dscene& S, const dmedium& M, f3 p) {
    const int n[3] = {(int)M.nx, (int)M.ny, (int)M.nz};
    const float pp[3] = {p.x, p.y, p.z};
    int i0[3], i1[3];
    float fr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float m = pp[k] - floorf(pp[k]);  // glm::mod(x, 1)
        const float uv = m * (float)n[k] - 0.5f;
        int a = (int)floorf(uv);
        a = a < 0 ? 0 : (a > n[k] - 1 ? n[k] - 1 : a);
        int b = a + 1;
        b = b > n[k] - 1 ? n[k] - 1 : b;
        i0[k] = a;
        i1[k] = b;
        fr[k] = gmin(gmax(uv - (float)a, 0.0f), 1.0f);
    }
    const float* g = S.grid + M.grid_offset;
    const size_t sx = 1, sy = (size_t)n[0], sz = (size_t)n[0] * n[1];
    auto at = [&](int x, int y, int z) { return __ldg(g + (size_t)z * sz + (size_t)y * sy + (size_t)x * sx); };
    auto lerp = [](float s, float t, float x) { return s * (1.0f - x) + t * x; };
    const float v00 = lerp(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), fr[0]);
    const float v01 = lerp(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), fr[0]);
    const float v10 = lerp(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), fr[0]);
    const float v11 = lerp(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), fr[0]);
    const float v0 = lerp(v00, v10, fr[1]);
    const float v1 = lerp(v01, v11, fr[1]);
    return lerp(v0, v1, fr[2]);
}

// medium::sample: whether a particle was sampled before len, where, and the throughput factor
struct medium_event {
    bool scattered;
    float t;
    f3 value;
    float g;
};

// homogeneous_medium::sample (homogeneous_medium.cpp:20-58) and heterogeneous_medium::sample
// (heterogeneous_medium.cpp:70-120); draws come from the path's 1D sampler
template <class SMP>
__device__ __forceinline__ medium_event medium_sample(const dscene& S, const dmedium& M, SMP& smp, f3 o, f3 d,
                                                      float len) {
    medium_event ev;
    ev.scattered = false;
    ev.t = len;
    ev.value = splat(1.0f);
    ev.g = M.g;
    if (M.type == RBW_MEDIUM_HOMOGENEOUS) {
        const float u0 = smp.next();
        uint32_t ch = (uint32_t)(u0 * 3.0f);
        if (ch > 2) ch = 2;
        const float st[3] = {M.sigma_t[0], M.sigma_t[1], M.sigma_t[2]};
        const float u1 = smp.next();
        const float sc = ch == 0 ? st[0] : (ch == 1 ? st[1] : st[2]);
        const float t = gmin(-flog(1.0f - u1) / sc, len);
        const bool sampled = t < len;
        const float tc = gmin(t, k_flt_max);
        const f3 value = mk(fexp(-st[0] * tc), fexp(-st[1] * tc), fexp(-st[2] * tc));
        const f3 dens = sampled ? mk(st[0], st[1], st[2]) * value : value;
        float pdf = 0.0f;
        pdf = pdf + dens.x;
        pdf = pdf + dens.y;
        pdf = pdf + dens.z;
        pdf = pdf / 3.0f;
        if (pdf == 0.0f) pdf = 1.0f;
        ev.scattered = sampled;
        ev.t = t;
        ev.value = sampled ? value * mk(M.sigma_s[0], M.sigma_s[1], M.sigma_s[2]) / pdf : value / pdf;
        return ev;
    }
    const f3 lo = w2l_point(M, o), ld = w2l_vector(M, d);
    float t0, t1;
    if (!unit_box_interval(lo, ld, len, t0, t1)) return ev;
    float t = t0;
    while (true) {
        t = t - flog(1.0f - smp.next()) * M.inv_max_density;
        if (t >= t1) break;
        const float sigma_t = grid_density(S, M, lo + ld * t);
        if (sigma_t * M.inv_max_density > smp.next()) {
            ev.scattered = true;
            ev.t = t;
            ev.value = mk(M.albedo[0], M.albedo[1], M.albedo[2]);
            return ev;
        }
    }
    return ev;
}

// medium::evaluate on a segment of length len (homogeneous_medium.cpp:12-18,
// heterogeneous_medium.cpp:24-68 ratio tracking)
template <class SMP>
__device__ __forceinline__ f3 medium_transmittance(const dscene& S, const dmedium& M, SMP& smp, f3 o, f3 d, float len) {
    if (M.type == RBW_MEDIUM_HOMOGENEOUS) {
        const float tc = gmin(len, k_flt_max);
        return mk(fexp(-M.sigma_t[0] * tc), fexp(-M.sigma_t[1] * tc), fexp(-M.sigma_t[2] * tc));
    }
    const f3 lo = w2l_point(M, o), ld = w2l_vector(M, d);
    float t0, t1;
    if (!unit_box_interval(lo, ld, len, t0, t1)) return splat(1.0f);
    float value = 1.0f, t = t0;
    while (true) {
        t = t - flog(1.0f - smp.next()) * M.inv_max_density;
        if (t >= t1) break;
        const float sigma_t = grid_density(S, M, lo + ld * t);
        value = value * (1.0f - gmax(0.0f, sigma_t * M.inv_max_density));
    }
    return splat(value);
}

// medium_info(entity, normal, wi) (medium.cpp:24-33): outside when dot(normal, wi) > 0
__device__ __forceinline__ uint32_t medium_after(const dscene& S, uint32_t entity, f3 normal, f3 wi) {
    return dot(normal, wi) > 0.0f ? S.ent_media[2 * entity] : S.ent_media[2 * entity + 1];
}

}  // namespace rbw
