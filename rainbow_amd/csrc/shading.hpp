// Surface interaction, BSDFs, light sampling and ray spawning of the render
// path (device and host).  Each routine restates the cited reference code
// with its floating-point operation order.
#pragma once

#include "device.hpp"

namespace rbw {

struct frame {
    f3 x, y, z;
};

struct surf {
    f3 normal, point, wo;
    frame sh;
    f3 dpdu, dpdv;
};

// shared/coordinate_system.cpp:8-23
RBW_HD frame frame_from_normal(f3 n) {
    frame s;
    s.z = normalize(n);
    if (fabsf(s.z.x) > fabsf(s.z.y))
        s.x = mk(-s.z.z, 0.0f, s.z.x) / sqrtf(s.z.x * s.z.x + s.z.z * s.z.z);
    else
        s.x = mk(0.0f, s.z.z, -s.z.y) / sqrtf(s.z.y * s.z.y + s.z.z * s.z.z);
    s.y = cross(s.z, s.x);
    return s;
}

// shared/coordinate_system.cpp:64-77
RBW_HD f3 world_to_local(const frame& s, f3 v) { return mk(dot(v, s.x), dot(v, s.y), dot(v, s.z)); }
RBW_HD f3 local_to_world(const frame& s, f3 v) {
    return mk(s.x.x * v.x + s.y.x * v.y + s.z.x * v.z, s.x.y * v.x + s.y.y * v.y + s.z.y * v.z,
              s.x.z * v.x + s.y.z * v.y + s.z.z * v.z);
}

// rainbow-core/shading_function.cpp:77-80
RBW_HD f3 face_forward(f3 v, f3 f) { return dot(v, f) > 0.0f ? v : -v; }

RBW_HD f3 load3(const float* p, uint32_t i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

// shapes/mesh.cpp:354-441 (local-space interaction of an accepted triangle hit, no tangents)
RBW_HD surf triangle_interaction(const dscene& S, const dentity& E, uint32_t tri, f3 d, float b0, float b1, float b2) {
    const uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
    const f3 p0 = load3(S.pos, i0), p1 = load3(S.pos, i1), p2 = load3(S.pos, i2);
    const bool has_n = (E.flags & RBW_ENTITY_HAS_NORMALS) != 0;
    const bool reverse = (E.flags & RBW_ENTITY_REVERSE_ORIENTATION) != 0;
    f3 uv0 = mk(0, 0, 0), uv1 = mk(1, 0, 0), uv2 = mk(1, 1, 0);
    if ((E.flags & RBW_ENTITY_HAS_UVS) && S.uv) {
        uv0 = mk(S.uv[2 * i0], S.uv[2 * i0 + 1], 0);
        uv1 = mk(S.uv[2 * i1], S.uv[2 * i1 + 1], 0);
        uv2 = mk(S.uv[2 * i2], S.uv[2 * i2 + 1], 0);
    }
    const f3 e1 = p1 - p0;
    const f3 e2 = p2 - p0;
    surf s;
    s.point = p0 * b0 + p1 * b1 + p2 * b2;
    const f3 gn = reverse ? -normalize(cross(e1, e2)) : normalize(cross(e1, e2));
    const f3 duv02 = uv0 - uv2;
    const f3 duv12 = uv1 - uv2;
    const f3 dp02 = p0 - p2;
    const f3 dp12 = p1 - p2;
    const float det = duv02.x * duv12.y - duv02.y * duv12.x;
    const float inv_det = 1.0f / det;
    f3 dpdu = (+duv12.y * dp02 - duv02.y * dp12) * inv_det;
    f3 dpdv = (-duv12.x * dp02 + duv02.x * dp12) * inv_det;
    if ((double)fabsf(det) < 1e-8 || length(cross(dpdu, dpdv)) == 0.0f) {
        const frame f = frame_from_normal(gn);
        dpdu = f.x;
        dpdv = f.y;
    }
    frame sh;
    if (has_n) {
        const f3 n0 = load3(S.nrm, i0), n1 = load3(S.nrm, i1), n2 = load3(S.nrm, i2);
        sh.z = b0 * n0 + b1 * n1 + b2 * n2;
        if (reverse) sh.z = sh.z * -1.0f;
        if (length_squared(sh.z) <= 0.0f) sh.z = gn;
    } else
        sh.z = gn;
    sh.x = dpdu;
    sh.x = normalize(sh.x);
    sh.z = normalize(sh.z);
    sh.y = cross(sh.z, sh.x);
    if (length_squared(sh.y) > 0.0f) {
        sh.y = normalize(sh.y);
        sh.x = normalize(cross(sh.y, sh.z));
    } else
        sh = frame_from_normal(sh.z);
    s.sh = sh;
    s.dpdu = dpdu;
    s.dpdv = dpdv;
    s.normal = face_forward(gn, sh.z);
    s.wo = -d;
    return s;
}

// shapes/sphere.cpp:40-85 (local space, sphere at the origin)
RBW_HD surf sphere_interaction(float radius, bool reverse, f3 o, f3 d, float t_hit) {
    f3 p = o + d * t_hit;
    p = p * radius / length(p);
    if (p.x == 0.0f && p.y == 0.0f) p.x = 1e-5f * radius;
    const float phi_max = k_two_pi;
    const float theta_min = racos(-1.0f);
    const float theta_max = racos(1.0f);
    const float theta = racos(gclamp(p.z / radius, -1.0f, 1.0f));
    const float rr = sqrtf(p.x * p.x + p.y * p.y);
    const float inv_r = 1.0f / rr;
    const float cos_p = p.x * inv_r;
    const float sin_p = p.y * inv_r;
    surf s;
    s.dpdu = mk(-phi_max * p.y, phi_max * p.x, 0.0f);
    s.dpdv = mk(p.z * cos_p, p.z * sin_p, -radius * rsin(theta)) * (theta_max - theta_min);
    s.normal = reverse ? -normalize(cross(s.dpdu, s.dpdv)) : normalize(cross(s.dpdu, s.dpdv));
    s.point = p;
    s.wo = -d;
    // surface_interaction.cpp:19-57
    const f3 sx = normalize(s.dpdu);
    s.sh.x = sx;
    s.sh.y = cross(s.normal, sx);
    s.sh.z = s.normal;
    return s;
}

// local -> world for the identity (meshes) or a translation (spheres): shared/transform.cpp:38-79
RBW_HD surf surf_to_world(const surf& l, bool translate_by, f3 c) {
    surf w;
    const f3 x = ident(l.sh.x);
    f3 y = ident(l.sh.y);
    w.sh.z = normalize(cross(x, y));
    w.sh.x = normalize(x);
    w.sh.y = normalize(cross(w.sh.z, w.sh.x));
    w.dpdu = ident(l.dpdu);
    w.dpdv = ident(l.dpdv);
    w.normal = normalize(ident(l.normal));
    w.point = translate_by ? translate(l.point, c) : ident(l.point);
    w.wo = normalize(ident(l.wo));
    return w;
}

// shared/interactions/interaction.cpp:25-46
RBW_HD f3 offset_ray_origin(f3 point, f3 normal, f3 dir) {
    const float scale = k_ray_epsilon * (1.0f + max_component(abs3(point)));
    f3 off = scale * normal;
    if (dot(normal, dir) < 0.0f) off = -off;
    f3 o = point + off;
    o.x = next_after(o.x, off.x > 0.0f ? k_flt_max : k_flt_min);
    o.y = next_after(o.y, off.y > 0.0f ? k_flt_max : k_flt_min);
    o.z = next_after(o.z, off.z > 0.0f ? k_flt_max : k_flt_min);
    return o;
}

// interaction.cpp:12-15: ray(direction, origin) normalises the direction
RBW_HD void spawn_ray(f3 point, f3 normal, f3 d, f3& o_out, f3& d_out) {
    o_out = offset_ray_origin(point, normal, d);
    d_out = normalize(d);
}

// interaction.cpp:17-23
RBW_HD void spawn_ray_to(f3 point, f3 normal, f3 target, f3& o_out, f3& d_out, float& len_out) {
    const f3 o = offset_ray_origin(point, normal, normalize(target - point));
    const f3 dir = target - o;
    o_out = o;
    d_out = normalize(normalize(dir));
    len_out = length(dir) - k_shadow_epsilon;
}

// ------------------------------------------------------------------ warps (rainbow-core/sample_function.cpp)

RBW_HD f2 concentric_sample_disk(f2 u) {
    const f2 r{2.0f * u.x - 1.0f, 2.0f * u.y - 1.0f};
    if (r.x == 0.0f && r.y == 0.0f) return f2{0.0f, 0.0f};
    float theta, rad;
    if (fabsf(r.x) > fabsf(r.y)) {
        rad = r.x;
        theta = k_quarter_pi * (r.y / r.x);
    } else {
        rad = r.y;
        theta = k_half_pi - k_quarter_pi * (r.x / r.y);
    }
    return f2{rad * rcos(theta), rad * rsin(theta)};
}

RBW_HD f3 cosine_sample_hemisphere(f2 u) {
    const f2 d = concentric_sample_disk(u);
    const float z = sqrtf(gmax(0.0f, 1.0f - dot2(d, d)));
    return mk(d.x, d.y, z);
}

RBW_HD f2 uniform_sample_triangle(f2 u) {
    const float s = sqrtf(u.x);
    return f2{1.0f - s, u.y * s};
}

// ------------------------------------------------------------------ shading trig (shading_function.cpp)

RBW_HD float cos2_theta(f3 w) { return w.z * w.z; }
RBW_HD float sin2_theta(f3 w) { return gmax(0.0f, 1.0f - cos2_theta(w)); }
RBW_HD float sin_theta(f3 w) { return sqrtf(sin2_theta(w)); }
RBW_HD float tan_theta(f3 w) { return sin_theta(w) / w.z; }
RBW_HD float tan2_theta(f3 w) { return sin2_theta(w) / cos2_theta(w); }
RBW_HD float cos_phi(f3 w) {
    const float s = sin_theta(w);
    return s == 0.0f ? 1.0f : gclamp(w.x / s, -1.0f, 1.0f);
}
RBW_HD float sin_phi(f3 w) {
    const float s = sin_theta(w);
    return s == 0.0f ? 0.0f : gclamp(w.y / s, -1.0f, 1.0f);
}
RBW_HD float cos2_phi(f3 w) { return cos_phi(w) * cos_phi(w); }
RBW_HD float sin2_phi(f3 w) { return sin_phi(w) * sin_phi(w); }
RBW_HD bool same_hemisphere(f3 a, f3 b) { return a.z * b.z > 0.0f; }
RBW_HD f3 reflect(f3 wi, f3 n) { return -wi + 2.0f * dot(wi, n) * n; }

// ------------------------------------------------------------------ BSDFs

// scatterings/fresnel_effect.cpp:35-62 with eta_i = 1
RBW_HD f3 fresnel_conductor(float cos_i, f3 eta_o, f3 k) {
    cos_i = gclamp(cos_i, -1.0f, 1.0f);
    const f3 eta = div3(eta_o, splat(1.0f));
    const f3 eta_k = div3(k, splat(1.0f));
    const float cos2 = cos_i * cos_i;
    const float sin2 = 1.0f - cos2;
    const f3 eta2 = eta * eta;
    const f3 eta_k2 = eta_k * eta_k;
    const f3 t0 = eta2 - eta_k2 - splat(sin2);
    const f3 s = t0 * t0 + eta2 * eta_k2 * 4.0f;
    const f3 a2b2 = mk(sqrtf(s.x), sqrtf(s.y), sqrtf(s.z));
    const f3 t1 = a2b2 + splat(cos2);
    const f3 ah = (a2b2 + t0) * 0.5f;
    const f3 a = mk(sqrtf(ah.x), sqrtf(ah.y), sqrtf(ah.z));
    const f3 t2 = a * 2.0f * cos_i;
    const f3 t3 = a2b2 * cos2 + splat(sin2 * sin2);
    const f3 t4 = t2 * sin2;
    const f3 r0 = div3(t1 - t2, t1 + t2);
    const f3 r1 = div3(r0 * (t3 - t4), t3 + t4);
    return (r0 + r1) * 0.5f;
}

// trowbridge_reitz_distribution.cpp (sample_visible_area = true)
struct ggx {
    float ax, ay;
    RBW_HD float D(f3 wh) const {
        const float tan2 = tan2_theta(wh);
        if (isinf(tan2)) return 0.0f;
        const float cos4 = cos2_theta(wh) * cos2_theta(wh);
        const float e = (cos2_phi(wh) / (ax * ax) + sin2_phi(wh) / (ay * ay)) * tan2;
        return 1.0f / (k_pi * ax * ay * cos4 * (1.0f + e) * (1.0f + e));
    }
    RBW_HD float lambda(f3 w) const {
        const float abs_tan = fabsf(tan_theta(w));
        if (isinf(abs_tan)) return 0.0f;
        const float alpha = sqrtf(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
        const float a2 = (alpha * abs_tan) * (alpha * abs_tan);
        return (-1.0f + sqrtf(1.0f + a2)) / 2.0f;
    }
    RBW_HD float G1(f3 w) const { return 1.0f / (1.0f + lambda(w)); }
    RBW_HD float G(f3 wo, f3 wi) const { return 1.0f / (1.0f + lambda(wo) + lambda(wi)); }
    RBW_HD float pdf(f3 wo, f3 wh) const { return D(wh) * G1(wo) * fabsf(dot(wo, wh)) / fabsf(wo.z); }
    RBW_HD static f2 sample11(float cos_t, float u1, float u2) {
        if ((double)cos_t > 0.9999) {
            const float r = sqrtf(u1 / (1.0f - u1));
            const float phi = k_two_pi * u2;
            return f2{r * rcos(phi), r * rsin(phi)};
        }
        f2 slope;
        const float sin_t = sqrtf(gmax(0.0f, 1.0f - cos_t * cos_t));
        const float tan_t = sin_t / cos_t;
        const float a = 1.0f / tan_t;
        const float g1 = 2.0f / (1.0f + sqrtf(1.0f + 1.0f / (a * a)));
        const float A = 2.0f * u1 / g1 - 1.0f;
        float tmp = 1.0f / (A * A - 1.0f);
        if (tmp > 1e10f) tmp = 1e10f;
        const float B = tan_t;
        const float D = sqrtf(gmax(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.0f));
        const float sx1 = B * tmp - D;
        const float sx2 = B * tmp + D;
        slope.x = (A < 0.0f || sx2 > 1.0f / tan_t) ? sx1 : sx2;
        float S;
        if (u2 > 0.5f) {
            S = 1.0f;
            u2 = 2.0f * (u2 - 0.5f);
        } else {
            S = -1.0f;
            u2 = 2.0f * (0.5f - u2);
        }
        const float z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
                        (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        slope.y = S * z * sqrtf(1.0f + slope.x * slope.x);
        return slope;
    }
    RBW_HD f3 sample_visible(f3 wi, float u1, float u2) const {
        const f3 ws = normalize(mk(ax * wi.x, ay * wi.y, wi.z));
        f2 slope = sample11(ws.z, u1, u2);
        const float tmp = cos_phi(ws) * slope.x - sin_phi(ws) * slope.y;
        slope.y = sin_phi(ws) * slope.x + cos_phi(ws) * slope.y;
        slope.x = tmp;
        slope.x = ax * slope.x;
        slope.y = ay * slope.y;
        return normalize(mk(-slope.x, -slope.y, 1.0f));
    }
    RBW_HD f3 sample(f3 wo, f2 u) const {
        const bool flip = wo.z < 0.0f;
        const f3 wh = sample_visible(flip ? -wo : wo, u.x, u.y);
        return flip ? -wh : wh;
    }
};

enum : uint32_t { LOBE_NONE = 0, LOBE_LAMBERT = 1, LOBE_GGX = 2 };

// A one-lobe scattering_function_collection (scatterings/scattering_function_collection.cpp)
// scatterings/fresnel_effect.cpp:5-33
RBW_HD float fresnel_dielectric(float cos_i, float eta_i, float eta_o) {
    cos_i = gclamp(cos_i, -1.0f, 1.0f);
    if (cos_i < 0.0f) {  // wi on the eta_o side: the roles swap
        const float t = eta_i;
        eta_i = eta_o;
        eta_o = t;
        cos_i = fabsf(cos_i);
    }
    const float sin_i = sqrtf(gmax(0.0f, 1.0f - cos_i * cos_i));
    const float sin_o = sin_i * eta_i / eta_o;
    const float cos_o = sqrtf(gmax(0.0f, 1.0f - sin_o * sin_o));
    if (sin_o >= 1.0f) return 1.0f;
    const float r0 = (eta_o * cos_i - eta_i * cos_o) / (eta_o * cos_i + eta_i * cos_o);
    const float r1 = (eta_i * cos_i - eta_o * cos_o) / (eta_i * cos_i + eta_o * cos_o);
    return (r0 * r0 + r1 * r1) / 2.0f;
}

// A material's scattering_function_collection (scattering_function_collection.cpp):
// at most a Lambert lobe (index 0) and a GGX reflection lobe (index 1), every lobe a
// reflection lobe.  Matte: Lambert; metal: GGX with conductor Fresnel, reflectance 1;
// plastic: both, GGX with dielectric Fresnel (1, ior) and reflectance `Rs`.
struct bsdf {
    uint32_t kind;        // LOBE_NONE, or the lobes present (LOBE_LAMBERT | LOBE_GGX bits)
    f3 R;                 // Lambert reflectance
    f3 Rs;                // GGX reflectance
    ggx dist;
    f3 eta, k;            // conductor Fresnel
    float ior;            // > 0: dielectric Fresnel instead
    RBW_HD uint32_t count() const { return ((kind & LOBE_LAMBERT) ? 1u : 0u) + ((kind & LOBE_GGX) ? 1u : 0u); }

    // lambertian_reflection.cpp:10-13
    RBW_HD f3 lambert_eval() const { return R * k_one_over_pi; }
    // microfacet_reflection.cpp:13-40
    RBW_HD f3 ggx_eval(f3 wo, f3 wi) const {
        const float cos_o = fabsf(wo.z);
        const float cos_i = fabsf(wi.z);
        f3 wh = wi + wo;
        if (cos_o == 0.0f || cos_i == 0.0f) return splat(0.0f);
        if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return splat(0.0f);
        wh = face_forward(normalize(wh), mk(0.0f, 0.0f, 1.0f));
        const f3 F = ior > 0.0f ? splat(fresnel_dielectric(dot(wi, wh), 1.0f, ior))
                                : fresnel_conductor(fabsf(dot(wi, wh)), eta, k);
        const float G = dist.G(wo, wi);
        const float D = dist.D(wh);
        return Rs * (G * D) * F / (4.0f * cos_i * cos_o);
    }
    // reflection_function.cpp:27-30 / microfacet_reflection.cpp:66-73
    RBW_HD float lambert_pdf(f3 wo, f3 wi) const { return same_hemisphere(wo, wi) ? fabsf(wi.z) * k_one_over_pi : 0.0f; }
    RBW_HD float ggx_pdf(f3 wo, f3 wi) const {
        if (!same_hemisphere(wo, wi)) return 0.0f;
        const f3 wh = normalize(wo + wi);
        return dist.pdf(wo, wh) / (4.0f * dot(wo, wh));
    }
    // scattering_function_collection.cpp:29-46 (every lobe reflects)
    RBW_HD f3 evaluate(f3 wo, f3 wi) const {
        if (wo.z == 0.0f) return splat(0.0f);
        f3 f = splat(0.0f);
        if (kind != LOBE_NONE && same_hemisphere(wo, wi)) {
            if (kind & LOBE_LAMBERT) f = f + lambert_eval();
            if (kind & LOBE_GGX) f = f + ggx_eval(wo, wi);
        }
        return f;
    }
    // scattering_function_collection.cpp:131-148
    RBW_HD float pdf(f3 wo, f3 wi) const {
        if (kind == LOBE_NONE || wo.z == 0.0f) return 0.0f;
        float p = 0.0f;
        if (kind & LOBE_LAMBERT) p = p + lambert_pdf(wo, wi);
        if (kind & LOBE_GGX) p = p + ggx_pdf(wo, wi);
        return p / (float)count();
    }
    // scattering_function_collection.cpp:48-103: lobe `which` = floor(u.x * n), u.x remapped
    RBW_HD bool sample(const surf& s, f2 u, f3& wi_world, f3& f, float& pdf) const {
        if (kind == LOBE_NONE) return false;
        const uint32_t n = count();
        uint32_t which = (uint32_t)floorf(u.x * (float)n);
        if (which > n - 1u) which = n - 1u;
        const float a = u.x * (float)n - (float)which;
        const f2 rem{(k_one_minus_epsilon < a) ? k_one_minus_epsilon : a, u.y};
        const f3 wo = world_to_local(s.sh, s.wo);
        if (wo.z == 0.0f) return false;
        const bool pick_lambert = (kind & LOBE_LAMBERT) && which == 0u;
        f3 wi;
        if (pick_lambert) {  // reflection_function.cpp:12-25
            wi = cosine_sample_hemisphere(rem);
            if (wo.z < 0.0f) wi.z = -wi.z;
            pdf = lambert_pdf(wo, wi);
        } else {  // microfacet_reflection.cpp:42-64
            if (wo.z == 0.0f) return false;
            const f3 wh = dist.sample(wo, rem);
            if (dot(wo, wh) < 0.0f) return false;
            wi = reflect(wo, wh);
            if (!same_hemisphere(wo, wi)) return false;
            pdf = dist.pdf(wo, wh) / (4.0f * dot(wo, wh));
        }
        if (pdf == 0.0f) return false;
        // value and pdf over all lobes
        const bool refl = same_hemisphere(wo, wi);
        f = splat(0.0f);
        pdf = 0.0f;
        if (kind & LOBE_LAMBERT) {
            if (refl) f = f + lambert_eval();
            pdf = pdf + lambert_pdf(wo, wi);
        }
        if (kind & LOBE_GGX) {
            if (refl) f = f + ggx_eval(wo, wi);
            pdf = pdf + ggx_pdf(wo, wi);
        }
        pdf = pdf / (float)n;
        wi_world = local_to_world(s.sh, wi);
        return true;
    }
};

RBW_HD bsdf make_bsdf(const dscene& S, uint32_t entity_material, uint32_t prim) {
    bsdf b;
    b.kind = LOBE_NONE;
    b.R = splat(0.0f);
    b.Rs = splat(1.0f);
    b.dist = ggx{0.1f, 0.1f};
    b.eta = splat(1.0f);
    b.k = splat(0.0f);
    b.ior = 0.0f;
    uint32_t mid = entity_material;
    if (S.tri_material && prim < S.num_tris && S.tri_material[prim] != RBW_NONE) mid = S.tri_material[prim];
    const dmaterial m = S.mats[mid];
    if (m.type == RBW_MATERIAL_MATTE) {  // matte_material.cpp:13-29
        const f3 d = mk(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
        if (is_black(d)) return b;
        b.kind = LOBE_LAMBERT;
        b.R = d;
    } else if (m.type == RBW_MATERIAL_METAL) {  // metal_material.cpp:17-37
        b.kind = LOBE_GGX;
        b.dist = ggx{m.alpha_u, m.alpha_v};
        b.eta = mk(m.eta[0], m.eta[1], m.eta[2]);
        b.k = mk(m.k[0], m.k[1], m.k[2]);
    } else {  // plastic_material.cpp:30-56
        const f3 d = mk(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
        const f3 sp = mk(m.specular[0], m.specular[1], m.specular[2]);
        if (!is_black(d)) {
            b.kind |= LOBE_LAMBERT;
            b.R = d;
        }
        if (!is_black(sp)) {
            b.kind |= LOBE_GGX;
            b.Rs = sp;
            b.dist = ggx{m.alpha_u, m.alpha_v};
            b.ior = m.ior;
        }
    }
    return b;
}

// integrator.cpp:423-426
RBW_HD float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

// entity::sample<emitter> -> surface_light::sample -> shape::sample -> mesh::sample
// (detail/entity.hpp:32-41, surface_light.cpp:39-54, shape.cpp:26-44, mesh.cpp:84-116)
// for a mesh emitter with the identity transform.  pdf = 0 means the empty sample.
struct light_sample {
    f3 point, normal, wi, L;
    float pdf;
};

RBW_HD light_sample sample_light(const dscene& S, const dentity& E, f3 ref_point_w, f3 ref_normal_w, f2 u) {
    light_sample ls;
    ls.pdf = 0.0f;
    ls.L = splat(0.0f);
    // transform_interaction(to_local, ref): identity transform
    const f3 ref_point = ident(ref_point_w);
    (void)ref_normal_w;
    const uint32_t count = E.count;
    uint32_t which = (uint32_t)floorf(u.x * (float)count);
    if (which > count - 1) which = count - 1;
    const float rx0 = u.x * (float)count - (float)which;
    const f2 rem{(k_one_minus_epsilon < rx0) ? k_one_minus_epsilon : rx0, u.y};
    const uint32_t tri = E.first + which;
    const uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
    const f3 p0 = load3(S.pos, i0), p1 = load3(S.pos, i1), p2 = load3(S.pos, i2);
    const f2 b = uniform_sample_triangle(rem);
    const f3 point = p0 * b.x + p1 * b.y + p2 * (1.0f - b.x - b.y);
    f3 normal = normalize(cross(p1 - p0, p2 - p0));
    if (E.flags & RBW_ENTITY_HAS_NORMALS) {
        const f3 n0 = load3(S.nrm, i0), n1 = load3(S.nrm, i1), n2 = load3(S.nrm, i2);
        normal = n0 * b.x + n1 * b.y + n2 * (1.0f - b.x - b.y);
    }
    if (E.flags & RBW_ENTITY_REVERSE_ORIENTATION) normal = normal * -1.0f;
    float pdf = 1.0f / E.area;
    f3 wi = point - ref_point;
    if (length_squared(wi) == 0.0f) return ls;
    wi = normalize(wi);
    pdf = pdf * distance_squared(ref_point, point) / fabsf(dot(normal, -wi));
    if (isinf(pdf)) return ls;
    if (pdf == 0.0f) return ls;
    const f3 wl = normalize(point - ref_point);
    const f3 Le = mk(E.radiance[0], E.radiance[1], E.radiance[2]);
    ls.L = dot(normal, -wl) > 0.0f ? Le : splat(0.0f);
    // emitter_sample::transform(to_world) (emitter.cpp:18-26)
    ls.point = ident(point);
    ls.normal = normalize(ident(normal));
    ls.wi = ident(wl);
    ls.pdf = pdf;
    return ls;
}

// surface_light::evaluate through entity::evaluate<emitter> (surface_light.cpp:12-16)
RBW_HD f3 light_evaluate(const dentity& E, f3 normal_w, f3 w_world) {
    const f3 n = normalize(ident(normal_w));
    const f3 w = ident(w_world);
    return dot(n, w) > 0.0f ? mk(E.radiance[0], E.radiance[1], E.radiance[2]) : splat(0.0f);
}

// entity::pdf<emitter> -> shape::pdf (shape.cpp:46-60): a ray cast against the
// light's own triangles, closest hit by linear scan (mesh.cpp:45-58)
// (raw_normal: the reference normal is used as given -- the zero normal of a medium vertex)
RBW_HD float light_pdf(const dscene& S, const dentity& E, f3 ref_point_w, f3 ref_normal_w, f3 wi_w,
                       bool raw_normal = false) {
    const f3 ref_point = ident(ref_point_w);
    const f3 ref_normal = raw_normal ? ref_normal_w : normalize(ident(ref_normal_w));
    const f3 wi = ident(wi_w);
    f3 o, d;
    spawn_ray(ref_point, ref_normal, wi, o, d);
    float len = __builtin_inff();
    uint32_t best = RBW_NONE;
    float bb1 = 0.0f, bb2 = 0.0f;
    for (uint32_t i = 0; i < E.count; i++) {
        const uint32_t tri = E.first + i;
        const f3 p0 = load3(S.pos, S.idx[3 * tri]), p1 = load3(S.pos, S.idx[3 * tri + 1]),
                 p2 = load3(S.pos, S.idx[3 * tri + 2]);
        float t, b1, b2;
        if (mt_test(p0, p1, p2, o, d, len, t, b1, b2)) {
            len = t;
            best = tri;
            bb1 = b1;
            bb2 = b2;
        }
    }
    if (best == RBW_NONE) return 0.0f;
    const surf h = triangle_interaction(S, E, best, d, 1.0f - bb1 - bb2, bb1, bb2);
    const float pdf = distance_squared(ref_point, h.point) / (fabsf(dot(h.normal, -wi)) * E.area);
    if (isinf(pdf)) return 0.0f;
    return pdf;
}

}  // namespace rbw
