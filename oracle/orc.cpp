// TEST INFRASTRUCTURE ONLY -- the CPU oracle.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load liborc.so;
// the product library never links or calls it.
//
// A restatement of rainbow-cpu's path-integrator render call, following the
// reference files line by line (file:line cited at each function) with the
// reference's floating-point operation order.  Scenes arrive as the flat
// rbw_scene_desc of include/rbw.h: mesh entities carry the identity transform
// (world-space triangles), sphere entities a translation.  Also restated: the
// direct integrator (direct_integrator.cpp) and the volume path integrator with
// homogeneous and grid media (volume_path_integrator.cpp, media/*.cpp).
//
// How each part is pinned lives with its tests (tests/test_oracle.py,
// test_direct.py, test_volume.py, ...): golden pcg32 vectors from the reference's
// own random_generator.cpp, closed forms (furnaces, polygon irradiance, Beer-Lambert,
// slab irradiance 2 rho Le E3(tau), HG normalisation and sampling), brute-force
// intersection.  Parity unpinned: the last bits of MSVC's float transcendentals
// (here correctly rounded sin/cos/acos/atan2 and the fixed flog/fexp sequence).
#include "orc.h"
#include "orc_math.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace orc {

constexpr uint32_t NONE = RBW_NONE;

// ---------------------------------------------------------------- geometry types

struct bound3 {
    vec3 min, max;
    bound3() : min(0), max(1) {}  // rainbow-core/math/detail/bound.hpp:7-10
    bound3(const vec3& p0, const vec3& p1) : min(vmin(p0, p1)), max(vmax(p0, p1)) {}
    void union_it(const bound3& b) { min = vmin(min, b.min); max = vmax(max, b.max); }
    void union_it(const vec3& p) { min = vmin(min, p); max = vmax(max, p); }
    vec3 centroid() const { return (min + max) * static_cast<real>(0.5); }
    // shared/accelerators/detail/accelerator.hpp:91-108
    size_t max_dimension() const {
        const vec3 l = max - min;
        if (l.x > l.y && l.x > l.z) return 0;
        if (l.y > l.z) return 1;
        return 2;
    }
};

// A ray with the reference's mutable length (shared/ray.hpp:13-28).
struct ray {
    vec3 direction, origin;
    real length;
};

// shared/ray.cpp:7-10: the constructor normalises the direction
inline ray make_ray(const vec3& d, const vec3& o, real len = std::numeric_limits<real>::infinity()) {
    return ray{normalize(d), o, len};
}

struct coordinate_system {
    vec3 x{1, 0, 0}, y{0, 1, 0}, z{0, 0, 1};
};

// shared/coordinate_system.cpp:8-23
inline coordinate_system coordinate_system_from_normal(const vec3& normal) {
    coordinate_system s;
    s.z = normalize(normal);
    if (std::fabs(s.z.x) > std::fabs(s.z.y))
        s.x = vec3(-s.z.z, 0, s.z.x) / std::sqrt(s.z.x * s.z.x + s.z.z * s.z.z);
    else
        s.x = vec3(0, s.z.z, -s.z.y) / std::sqrt(s.z.y * s.z.y + s.z.z * s.z.z);
    s.y = cross(s.z, s.x);
    return s;
}

// shared/coordinate_system.cpp:64-77
inline vec3 world_to_local(const coordinate_system& s, const vec3& v) { return {dot(v, s.x), dot(v, s.y), dot(v, s.z)}; }
inline vec3 local_to_world(const coordinate_system& s, const vec3& v) {
    return {s.x.x * v.x + s.y.x * v.y + s.z.x * v.z, s.x.y * v.x + s.y.y * v.y + s.z.y * v.z,
            s.x.z * v.x + s.y.z * v.y + s.z.z * v.z};
}

struct interaction {
    vec3 normal{0}, point{0}, wo{0};
};

struct surface_interaction : interaction {
    coordinate_system shading;
    uint32_t entity = NONE;
    uint32_t prim = NONE;
    vec3 dp_du{0}, dp_dv{0};
    vec2 uv;
};

// shared/interactions/interaction.cpp:25-46
inline vec3 offset_ray_origin(const interaction& it, const vec3& direction) {
    const real offset_scale = k_ray_epsilon * (1 + max_component(vabs(it.point)));
    vec3 offset = offset_scale * it.normal;
    if (dot(it.normal, direction) < 0) offset = -offset;
    vec3 o = it.point + offset;
    o.x = std::nextafter(o.x, offset.x > 0 ? std::numeric_limits<real>::max() : std::numeric_limits<real>::min());
    o.y = std::nextafter(o.y, offset.y > 0 ? std::numeric_limits<real>::max() : std::numeric_limits<real>::min());
    o.z = std::nextafter(o.z, offset.z > 0 ? std::numeric_limits<real>::max() : std::numeric_limits<real>::min());
    return o;
}
// interaction.cpp:12-15
inline ray spawn_ray(const interaction& it, const vec3& d) { return make_ray(d, offset_ray_origin(it, d)); }
// interaction.cpp:17-23
inline ray spawn_ray_to(const interaction& it, const vec3& p) {
    const vec3 origin = offset_ray_origin(it, normalize(p - it.point));
    const vec3 direction = p - origin;
    return make_ray(normalize(direction), origin, length(direction) - k_shadow_epsilon);
}

// shared/transform.cpp:94-105: ray transform (re-normalises, scales the length)
inline ray transform_ray(const xform& t, const ray& r) {
    const vec3 d = transform_vector(t, r.direction);
    return make_ray(normalize(d), transform_point(t, r.origin), r.length * length(d));
}

// shared/transform.cpp:64-79 (coordinate_system transform)
inline coordinate_system transform_system(const xform& t, const coordinate_system& s) {
    vec3 x = transform_vector(t, s.x);
    vec3 y = transform_vector(t, s.y);
    const vec3 z = normalize(cross(x, y));
    x = normalize(x);
    y = normalize(cross(z, x));
    return {x, y, z};
}

// shared/transform.cpp:38-50
inline surface_interaction transform_si(const xform& t, const surface_interaction& si) {
    surface_interaction r;
    r.entity = si.entity;
    r.prim = si.prim;
    r.shading = transform_system(t, si.shading);
    r.dp_du = transform_vector(t, si.dp_du);
    r.dp_dv = transform_vector(t, si.dp_dv);
    r.normal = transform_normal(t, si.normal);
    r.point = transform_point(t, si.point);
    r.wo = normalize(transform_vector(t, si.wo));
    r.uv = si.uv;
    return r;
}

// shared/transform.cpp:190-197
inline interaction transform_interaction(const xform& t, const interaction& it) {
    return interaction{transform_normal(t, it.normal), transform_point(t, it.point), transform_vector(t, it.wo)};
}

// surface_interaction.cpp:19-57 (constructor building the shading space from dp_du)
inline void si_from_dpdu(surface_interaction& si) {
    const vec3 sx = normalize(si.dp_du);
    const vec3 sy = cross(si.normal, sx);
    si.shading = {sx, sy, si.normal};
}

// rainbow-core/shading_function.cpp:77-80
inline vec3 face_forward(const vec3& v, const vec3& f) { return dot(v, f) > 0 ? v : -v; }

// ---------------------------------------------------------------- warps (rainbow-core/sample_function.cpp)

// sample_function.cpp:25-44
inline vec2 concentric_sample_disk(const vec2& s) {
    const vec2 r(static_cast<real>(2) * s.x - 1, static_cast<real>(2) * s.y - 1);
    if (r.x == 0 && r.y == 0) return vec2(0, 0);
    real theta, rad;
    if (std::fabs(r.x) > std::fabs(r.y)) {
        rad = r.x;
        theta = k_quarter_pi * (r.y / r.x);
    } else {
        rad = r.y;
        theta = k_half_pi - k_quarter_pi * (r.x / r.y);
    }
    return vec2(rad * rcos(theta), rad * rsin(theta));
}
// sample_function.cpp:15-23
inline vec3 cosine_sample_hemisphere(const vec2& s) {
    const vec2 d = concentric_sample_disk(s);
    const real z = std::sqrt(fmax_g(static_cast<real>(0), 1 - dot(d, d)));
    return vec3(d.x, d.y, z);
}
// sample_function.cpp:57-62
inline vec2 uniform_sample_triangle(const vec2& s) {
    const real u = std::sqrt(s.x);
    return vec2(1 - u, s.y * u);
}

// ---------------------------------------------------------------- shading trig (shading_function.cpp)

inline real cos_theta(const vec3& w) { return w.z; }
inline real cos2_theta(const vec3& w) { return cos_theta(w) * cos_theta(w); }
inline real sin2_theta(const vec3& w) { return fmax_g(static_cast<real>(0), 1 - cos2_theta(w)); }
inline real sin_theta(const vec3& w) { return std::sqrt(sin2_theta(w)); }
inline real tan_theta(const vec3& w) { return sin_theta(w) / cos_theta(w); }
inline real tan2_theta(const vec3& w) { return sin2_theta(w) / cos2_theta(w); }
inline real cos_phi(const vec3& w) {
    const real s = sin_theta(w);
    return s == 0 ? 1 : clamp_g(w.x / s, -1, 1);
}
inline real sin_phi(const vec3& w) {
    const real s = sin_theta(w);
    return s == 0 ? 0 : clamp_g(w.y / s, -1, 1);
}
inline real cos2_phi(const vec3& w) { return cos_phi(w) * cos_phi(w); }
inline real sin2_phi(const vec3& w) { return sin_phi(w) * sin_phi(w); }
inline bool same_hemisphere(const vec3& a, const vec3& b) { return a.z * b.z > 0; }
// shading_function.cpp:103-109
inline vec3 reflect(const vec3& wi, const vec3& n) { return -wi + 2 * dot(wi, n) * n; }

// ---------------------------------------------------------------- spectra

using spectrum = vec3;  // color_spectrum = coefficient_spectrum<3> (element-wise FP32 arithmetic)
inline bool is_black(const spectrum& s) { return s.x == 0 && s.y == 0 && s.z == 0; }
inline spectrum sdiv(const spectrum& a, const spectrum& b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline spectrum ssqrt(const spectrum& a) { return {std::sqrt(a.x), std::sqrt(a.y), std::sqrt(a.z)}; }
// coefficient_specturm.hpp:38-49
inline real smax_component(const spectrum& s) { return fmax_g(fmax_g(s.x, s.y), s.z); }

// ---------------------------------------------------------------- BSDFs

// scatterings/fresnel_effect.cpp:35-62 with eta_i = 1
inline spectrum fresnel_conductor(real cos_i, const spectrum& eta_o, const spectrum& k) {
    cos_i = clamp_g(cos_i, -1, 1);
    const spectrum eta = sdiv(eta_o, spectrum(1));
    const spectrum eta_k = sdiv(k, spectrum(1));
    const real cos2 = cos_i * cos_i;
    const real sin2 = 1 - cos2;
    const spectrum eta2 = eta * eta;
    const spectrum eta_k2 = eta_k * eta_k;
    const spectrum t0 = eta2 - eta_k2 - spectrum(sin2);
    const spectrum a2b2 = ssqrt(t0 * t0 + eta2 * eta_k2 * static_cast<real>(4));
    const spectrum t1 = a2b2 + spectrum(cos2);
    const spectrum a = ssqrt((a2b2 + t0) * static_cast<real>(0.5));
    const spectrum t2 = a * static_cast<real>(2) * cos_i;
    const spectrum t3 = a2b2 * cos2 + spectrum(sin2 * sin2);
    const spectrum t4 = t2 * sin2;
    const spectrum r0 = sdiv(t1 - t2, t1 + t2);
    const spectrum r1 = sdiv(r0 * (t3 - t4), t3 + t4);
    return (r0 + r1) * static_cast<real>(0.5);
}

struct ggx {
    real ax, ay;
    // trowbridge_reitz_distribution.cpp:76-88
    real D(const vec3& wh) const {
        const real tan2 = tan2_theta(wh);
        if (std::isinf(tan2)) return 0;
        const real cos4 = cos2_theta(wh) * cos2_theta(wh);
        const real e = (cos2_phi(wh) / (ax * ax) + sin2_phi(wh) / (ay * ay)) * tan2;
        return 1 / (k_pi * ax * ay * cos4 * (1 + e) * (1 + e));
    }
    // trowbridge_reitz_distribution.cpp:138-150
    real lambda(const vec3& w) const {
        const real abs_tan = std::fabs(tan_theta(w));
        if (std::isinf(abs_tan)) return 0;
        const real alpha = std::sqrt(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
        const real a2 = (alpha * abs_tan) * (alpha * abs_tan);
        return (-1 + std::sqrt(1.f + a2)) / 2;
    }
    // microfacet_distribution.cpp:10-21
    real G1(const vec3& w) const { return 1 / (1 + lambda(w)); }
    real G(const vec3& wo, const vec3& wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    // microfacet_distribution.cpp:23-30 (sample_visible_area = true, metal_material.cpp:30)
    real pdf(const vec3& wo, const vec3& wh) const {
        return D(wh) * G1(wo) * std::fabs(dot(wo, wh)) / std::fabs(cos_theta(wo));
    }
    // trowbridge_reitz_distribution.cpp:90-126 (visible branch)
    vec3 sample(const vec3& wo, const vec2& u) const {
        const bool flip = wo.z < 0;
        const vec3 wh = sample_visible(flip ? -wo : wo, u.x, u.y);
        return flip ? -wh : wh;
    }
    // trowbridge_reitz_distribution.cpp:7-51
    static vec2 sample11(real cos_t, real u1, real u2) {
        if (cos_t > 0.9999) {
            const real r = std::sqrt(u1 / (1 - u1));
            const real phi = k_two_pi * u2;
            return vec2(r * rcos(phi), r * rsin(phi));
        }
        vec2 slope;
        const real sin_t = std::sqrt(fmax_g(static_cast<real>(0), 1 - cos_t * cos_t));
        const real tan_t = sin_t / cos_t;
        const real a = 1 / tan_t;
        const real g1 = 2 / (1 + std::sqrt(1.f + 1.f / (a * a)));
        const real A = 2 * u1 / g1 - 1;
        real tmp = 1.f / (A * A - 1.f);
        if (tmp > 1e10) tmp = 1e10;
        const real B = tan_t;
        const real D = std::sqrt(fmax_g(real(B * B * tmp * tmp - (A * A - B * B) * tmp), real(0)));
        const real sx1 = B * tmp - D;
        const real sx2 = B * tmp + D;
        slope.x = (A < 0 || sx2 > 1.f / tan_t) ? sx1 : sx2;
        real S;
        if (u2 > 0.5f) {
            S = 1.f;
            u2 = 2.f * (u2 - .5f);
        } else {
            S = -1.f;
            u2 = 2.f * (.5f - u2);
        }
        const real z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
                       (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        slope.y = S * z * std::sqrt(1.f + slope.x * slope.x);
        return slope;
    }
    // trowbridge_reitz_distribution.cpp:53-66
    vec3 sample_visible(const vec3& wi, real u1, real u2) const {
        const vec3 ws = normalize(vec3(ax * wi.x, ay * wi.y, wi.z));
        vec2 slope = sample11(cos_theta(ws), u1, u2);
        const real tmp = cos_phi(ws) * slope.x - sin_phi(ws) * slope.y;
        slope.y = sin_phi(ws) * slope.x + cos_phi(ws) * slope.y;
        slope.x = tmp;
        slope.x = ax * slope.x;
        slope.y = ay * slope.y;
        return normalize(vec3(-slope.x, -slope.y, 1.));
    }
};

enum : uint32_t { LOBE_NONE = 0, LOBE_LAMBERT = 1, LOBE_GGX = 2 };

// The scattering_function_collection of one surface: all configurations use
// at most one lobe (matte -> Lambert, metal -> microfacet reflection).
// fresnel_effect.cpp:5-33
inline real fresnel_dielectric(real cos_i, real eta_i, real eta_o) {
    cos_i = std::min(std::max(cos_i, static_cast<real>(-1)), static_cast<real>(1));
    if (cos_i < 0.f) {
        std::swap(eta_i, eta_o);
        cos_i = std::fabs(cos_i);
    }
    const real sin_i = std::sqrt(std::max(static_cast<real>(0), 1 - cos_i * cos_i));
    const real sin_o = sin_i * eta_i / eta_o;
    const real cos_o = std::sqrt(std::max(static_cast<real>(0), 1 - sin_o * sin_o));
    if (sin_o >= 1) return 1;
    const real r0 = (eta_o * cos_i - eta_i * cos_o) / (eta_o * cos_i + eta_i * cos_o);
    const real r1 = (eta_i * cos_i - eta_o * cos_o) / (eta_i * cos_i + eta_o * cos_o);
    return (r0 * r0 + r1 * r1) / 2;
}

// A material's scattering_function_collection: at most a Lambert lobe (added first) and a
// GGX microfacet reflection lobe; matte (Lambert), metal (GGX, conductor Fresnel,
// reflectance 1), plastic (both; GGX with fresnel_effect_dielectric(1, eta) and
// reflectance `specular`, plastic_material.cpp:30-56).
struct bsdf {
    uint32_t kind = LOBE_NONE;  // lobe bits
    spectrum R{0};              // Lambert reflectance
    spectrum Rs{1};             // GGX reflectance
    ggx dist{0.1f, 0.1f};
    spectrum eta{1}, k{0};      // conductor Fresnel
    real ior = 0;               // > 0: dielectric Fresnel (1, ior)
    bool fresnel_one = false;   // fresnel_effect_nop (fresnel_effect.cpp:90-92), used by closed-form tests

    uint32_t count() const { return ((kind & LOBE_LAMBERT) ? 1u : 0u) + ((kind & LOBE_GGX) ? 1u : 0u); }

    // lambertian_reflection.cpp:10-13
    spectrum lambert_eval() const { return R * k_one_over_pi; }
    // microfacet_reflection.cpp:13-40
    spectrum ggx_eval(const vec3& wo, const vec3& wi) const {
        const real cos_o = std::fabs(cos_theta(wo));
        const real cos_i = std::fabs(cos_theta(wi));
        vec3 wh = wi + wo;
        if (cos_o == 0 || cos_i == 0) return spectrum(0);
        if (wh.x == 0 && wh.y == 0 && wh.z == 0) return spectrum(0);
        wh = face_forward(normalize(wh), vec3(0, 0, 1));
        const spectrum F = fresnel_one ? spectrum(1)
                           : ior > 0   ? spectrum(fresnel_dielectric(dot(wi, wh), 1, ior))
                                       : fresnel_conductor(std::fabs(dot(wi, wh)), eta, k);
        const real G = dist.G(wo, wi);
        const real D = dist.D(wh);
        return Rs * (G * D) * F / (4 * cos_i * cos_o);
    }
    // reflection_function.cpp:27-30 / microfacet_reflection.cpp:66-73
    real lambert_pdf(const vec3& wo, const vec3& wi) const {
        return same_hemisphere(wo, wi) ? std::fabs(cos_theta(wi)) * k_one_over_pi : 0;
    }
    real ggx_pdf(const vec3& wo, const vec3& wi) const {
        if (!same_hemisphere(wo, wi)) return 0;
        const vec3 wh = normalize(wo + wi);
        return dist.pdf(wo, wh) / (4 * dot(wo, wh));
    }
    // microfacet_reflection.cpp:42-64; false for the empty scattering_sample {}
    bool ggx_sample(const vec3& wo, const vec2& u, vec3& wi, spectrum& f, real& pdf) const {
        if (wo.z == 0) return false;
        const vec3 wh = dist.sample(wo, u);
        if (dot(wo, wh) < 0) return false;
        wi = reflect(wo, wh);
        if (!same_hemisphere(wo, wi)) return false;
        f = ggx_eval(wo, wi);
        pdf = dist.pdf(wo, wh) / (4 * dot(wo, wh));
        return true;
    }

    // scattering_function_collection.cpp:29-46 (every lobe reflects)
    spectrum evaluate(const vec3& wo, const vec3& wi) const {
        if (wo.z == 0) return spectrum(0);
        spectrum f(0);
        if (count() && same_hemisphere(wo, wi)) {
            if (kind & LOBE_LAMBERT) f = f + lambert_eval();
            if (kind & LOBE_GGX) f = f + ggx_eval(wo, wi);
        }
        return f;
    }
    // scattering_function_collection.cpp:131-148
    real pdf(const vec3& wo, const vec3& wi) const {
        if (count() == 0 || wo.z == 0) return 0;
        real p = 0;
        if (kind & LOBE_LAMBERT) p = p + lambert_pdf(wo, wi);
        if (kind & LOBE_GGX) p = p + ggx_pdf(wo, wi);
        return p / count();
    }
    // scattering_function_collection.cpp:48-103; wi in shading space
    bool sample_local(const vec3& wo, const vec2& u, vec3& wi, spectrum& f, real& pdf) const {
        if (count() == 0) return false;
        const size_t n = count();
        const size_t which = std::min(static_cast<size_t>(std::floor(u.x * n)), n - 1);
        const vec2 remapped(std::min(u.x * n - which, static_cast<real>(1) - std::numeric_limits<real>::epsilon()), u.y);
        if (wo.z == 0) return false;
        if ((kind & LOBE_LAMBERT) && which == 0) {  // reflection_function.cpp:12-25
            wi = cosine_sample_hemisphere(remapped);
            if (wo.z < 0) wi.z = -wi.z;
            pdf = lambert_pdf(wo, wi);
        } else {
            spectrum unused;
            if (!ggx_sample(wo, remapped, wi, unused, pdf)) return false;
        }
        if (pdf == 0) return false;
        // non-specular lobes: value and pdf over every lobe
        const bool reflection = same_hemisphere(wo, wi);
        f = spectrum(0);
        pdf = 0;
        if (kind & LOBE_LAMBERT) {
            if (reflection) f = f + lambert_eval();
            pdf = pdf + lambert_pdf(wo, wi);
        }
        if (kind & LOBE_GGX) {
            if (reflection) f = f + ggx_eval(wo, wi);
            pdf = pdf + ggx_pdf(wo, wi);
        }
        pdf = pdf / n;
        return true;
    }
    bool sample(const surface_interaction& si, const vec2& u, vec3& wi_world, spectrum& f, real& pdf) const {
        if (count() == 0) return false;
        const vec3 wo = world_to_local(si.shading, si.wo);
        vec3 wi;
        if (!sample_local(wo, u, wi, f, pdf)) return false;
        wi_world = local_to_world(si.shading, wi);
        return true;
    }
};

// the lobes of an rbw_material (matte_material.cpp:13-29, metal_material.cpp:17-37,
// plastic_material.cpp:30-56)
inline bsdf material_bsdf(const rbw_material& m) {
    bsdf b;
    if (m.type == RBW_MATERIAL_MATTE) {
        const spectrum d(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
        if (is_black(d)) return b;
        b.kind = LOBE_LAMBERT;
        b.R = d;
    } else if (m.type == RBW_MATERIAL_METAL) {
        b.kind = LOBE_GGX;
        b.dist = ggx{m.alpha_u, m.alpha_v};
        b.eta = spectrum(m.eta[0], m.eta[1], m.eta[2]);
        b.k = spectrum(m.k[0], m.k[1], m.k[2]);
    } else {
        const spectrum d(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
        const spectrum sp(m.specular[0], m.specular[1], m.specular[2]);
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

// ---------------------------------------------------------------- scene

struct entity_ref {
    uint32_t entity;
    uint32_t index;  // global primitive id (triangle or num_triangles+sphere) or NONE = whole shape
    bound3 box;
};

struct bvh_node {
    bound3 box;
    int32_t left = -1, right = -1;
    uint32_t begin = 0, end = 0, axis = 0;
    bool leaf() const { return left < 0 && right < 0; }
};

// shared/accelerators/detail/bounding_volume_hierarchy.hpp:60-233, median split
template <typename Ref>
struct bvh {
    std::vector<Ref> refs;
    std::vector<bvh_node> nodes;
    int32_t root = -1;

    void build() {
        nodes.clear();
        nodes.reserve(refs.empty() ? 1 : 2 * refs.size());
        root = refs.empty() ? -1 : build_rec(0, static_cast<uint32_t>(refs.size()));
    }
    int32_t build_rec(uint32_t begin, uint32_t end) {
        const int32_t id = static_cast<int32_t>(nodes.size());
        nodes.emplace_back();
        bound3 ub = refs[begin].box;
        for (uint32_t i = begin + 1; i < end; i++) ub.union_it(refs[i].box);
        if (end - begin == 1) {
            nodes[id].box = ub; nodes[id].begin = begin; nodes[id].end = end;
            return id;
        }
        bound3 cb(refs[begin].box.centroid(), refs[begin].box.centroid());
        for (uint32_t i = begin + 1; i < end; i++) cb.union_it(refs[i].box.centroid());
        const size_t dim = cb.max_dimension();
        if (cb.max[static_cast<int>(dim)] == cb.min[static_cast<int>(dim)]) {
            nodes[id].box = ub; nodes[id].begin = begin; nodes[id].end = end;
            return id;
        }
        const uint32_t middle = (begin + end) >> 1;  // split_equal_count
        std::nth_element(refs.begin() + begin, refs.begin() + middle, refs.begin() + end,
                         [dim](const Ref& a, const Ref& b) {
                             return a.box.centroid()[static_cast<int>(dim)] < b.box.centroid()[static_cast<int>(dim)];
                         });
        const int32_t l = build_rec(begin, middle);
        const int32_t r = build_rec(middle, end);
        bvh_node& n = nodes[id];
        n.left = l; n.right = r;
        n.box = nodes[l].box;
        n.box.union_it(nodes[r].box);  // bounding_box(box0, box1): min/max of the children
        n.begin = begin; n.end = end; n.axis = static_cast<uint32_t>(dim);
        return id;
    }
};

// shared/accelerators/detail/accelerator.hpp:29-53
inline bool box_intersect(const bound3& box, const ray& r) {
    real t0 = 0, t1 = r.length;
    for (int d = 0; d < 3; d++) {
        const real inv = 1 / r.direction[d];
        real near_t = (box.min[d] - r.origin[d]) * inv;
        real far_t = (box.max[d] - r.origin[d]) * inv;
        if (near_t > far_t) std::swap(near_t, far_t);
        t0 = fmax_g(t0, near_t);
        t1 = fmin_g(t1, far_t);
        if (t0 > t1) return false;
    }
    return true;
}

struct entity_data {
    rbw_entity e;
    xform to_world, to_local;
    real area = 0;          // shape_instance_properties::area (shape.cpp:51-54)
    bool instanced = false; // shape got its own accelerator (scene.cpp:56-75)
    bvh<entity_ref> mesh_bvh;  // refs hold triangle ids
    bound3 bounds;
    bool visible() const { return e.shape != RBW_SHAPE_NONE && e.material != NONE; }  // entity.cpp:80-83
    bool emitter() const { return (e.flags & RBW_ENTITY_EMITTER) != 0; }
};

struct hit_candidate {
    uint32_t entity = NONE;
    uint32_t prim = NONE;
    vec3 o_local, d_local;
    real len_local = 0;  // local length in effect when the accepted test ran
    real t = 0;
    real b1 = 0, b2 = 0;
};

struct counters {
    uint64_t camera = 0, extension = 0, shadow = 0, mis = 0, light = 0;
};

}  // namespace orc

struct orc_scene {
    std::vector<orc::vec3> P, N;
    std::vector<orc::vec2> UV;
    std::vector<uint32_t> idx, tri_entity, tri_material;
    std::vector<float> spheres;  // cx cy cz r
    std::vector<uint32_t> sphere_entity;
    std::vector<orc::entity_data> ents;
    std::vector<rbw_material> mats;
    std::vector<uint32_t> emitters;
    // media (media/medium.hpp:37-75): table, (outside, inside) per entity, grid densities,
    // and 1 / the largest density of each grid medium (heterogeneous_medium.cpp:12-22)
    std::vector<rbw_medium> media;
    std::vector<uint32_t> ent_media;
    std::vector<float> grid;
    std::vector<float> inv_max_density;
    orc::bvh<orc::entity_ref> top;
    uint32_t num_tris = 0, num_spheres = 0;
};

namespace orc {

inline void tri_positions(const orc_scene& s, uint32_t tri, vec3& p0, vec3& p1, vec3& p2) {
    p0 = s.P[s.idx[3 * tri + 0]];
    p1 = s.P[s.idx[3 * tri + 1]];
    p2 = s.P[s.idx[3 * tri + 2]];
}

// shapes/mesh.cpp:328-352: Moller-Trumbore with the reference's tests
inline bool mt_test(const vec3& p0, const vec3& p1, const vec3& p2, const vec3& o, const vec3& d, real len, real& t,
                    real& b0, real& b1, real& b2) {
    const vec3 e1 = p1 - p0;
    const vec3 e2 = p2 - p0;
    const vec3 p_vec = cross(d, e2);
    const vec3 t_vec = o - p0;
    const vec3 q_vec = cross(t_vec, e1);
    const real inv_det = static_cast<real>(1) / dot(e1, p_vec);
    b1 = dot(t_vec, p_vec) * inv_det;
    b2 = dot(d, q_vec) * inv_det;
    b0 = 1 - b1 - b2;
    t = dot(e2, q_vec) * inv_det;
    if (t <= 0 || t >= len) return false;
    if (b0 < 0 || b0 > 1) return false;
    if (b1 < 0 || b1 > 1) return false;
    if (b2 < 0 || b2 > 1) return false;
    return true;
}

// utilities.cpp:7-31 (in double)
inline bool solve_quadratic(real a, real b, real c, real& t0, real& t1) {
    const double discrim = static_cast<double>(b) * static_cast<double>(b) - static_cast<double>(a) * static_cast<double>(c) * 4;
    if (discrim < 0) return false;
    const double root = std::sqrt(discrim);
    double q;
    if (b < 0) q = -0.5 * (b - root);
    else q = -0.5 * (b + root);
    t0 = static_cast<real>(q / a);
    t1 = static_cast<real>(c / q);
    if (t0 > t1) std::swap(t0, t1);
    return true;
}

// sphere.cpp:17-38: returns the hit t (the ray length after the test)
inline bool sphere_test(real radius, const vec3& o, const vec3& d, real len, real& t_hit) {
    const real a = dot(d, d);
    const real b = dot(d, o) * 2;
    const real c = dot(o, o) - radius * radius;
    real t0, t1;
    if (!solve_quadratic(a, b, c, t0, t1)) return false;
    if (t0 > len || t1 <= 0) return false;
    if (t0 <= 0) {
        if (t1 > len) return false;
        t_hit = t1;
    } else
        t_hit = t0;
    return true;
}

// sphere.cpp:40-85: local-space interaction at t_hit
inline surface_interaction sphere_interaction(real radius, bool reverse, const vec3& o, const vec3& d, real t_hit) {
    vec3 p = o + d * t_hit;
    p = p * radius / length(p);
    if (p.x == 0 && p.y == 0) p.x = 1e-5f * radius;
    real phi = ratan2(p.y, p.x);
    if (phi < 0) phi = phi + k_two_pi;
    const real phi_max = k_two_pi;
    const real theta_min = racos(static_cast<real>(-1));
    const real theta_max = racos(static_cast<real>(1));
    const real theta = racos(clamp_g(p.z / radius, -1, 1));
    const real u = phi / phi_max;
    const real v = (theta - theta_min) / (theta_max - theta_min);
    const real rr = std::sqrt(p.x * p.x + p.y * p.y);
    const real inv_r = 1 / rr;
    const real cos_p = p.x * inv_r;
    const real sin_p = p.y * inv_r;
    surface_interaction si;
    si.dp_du = vec3(-phi_max * p.y, phi_max * p.x, 0);
    si.dp_dv = vec3(p.z * cos_p, p.z * sin_p, -radius * rsin(theta)) * (theta_max - theta_min);
    si.normal = reverse ? -normalize(cross(si.dp_du, si.dp_dv)) : normalize(cross(si.dp_du, si.dp_dv));
    si.point = p;
    si.wo = -d;
    si.uv = vec2(u, v);
    si_from_dpdu(si);
    return si;
}

// mesh.cpp:354-441: local-space interaction of an accepted triangle hit
inline surface_interaction triangle_interaction(const orc_scene& s, const entity_data& ent, uint32_t tri, const vec3& d,
                                                real b0, real b1, real b2) {
    vec3 p0, p1, p2;
    tri_positions(s, tri, p0, p1, p2);
    const bool has_uv = (ent.e.flags & RBW_ENTITY_HAS_UVS) != 0;
    const bool has_n = (ent.e.flags & RBW_ENTITY_HAS_NORMALS) != 0;
    const bool reverse = (ent.e.flags & RBW_ENTITY_REVERSE_ORIENTATION) != 0;
    vec3 uv0(0, 0, 0), uv1(1, 0, 0), uv2(1, 1, 0);  // mesh.cpp:201-208
    if (has_uv) {
        const vec2 a = s.UV[s.idx[3 * tri]], b = s.UV[s.idx[3 * tri + 1]], c = s.UV[s.idx[3 * tri + 2]];
        uv0 = vec3(a.x, a.y, 0); uv1 = vec3(b.x, b.y, 0); uv2 = vec3(c.x, c.y, 0);
    }
    const vec3 e1 = p1 - p0;
    const vec3 e2 = p2 - p0;
    const vec3 point = p0 * b0 + p1 * b1 + p2 * b2;
    const vec3 uv = uv0 * b0 + uv1 * b1 + uv2 * b2;
    const vec3 normal = reverse ? -normalize(cross(e1, e2)) : normalize(cross(e1, e2));
    const vec3 duv02 = uv0 - uv2;
    const vec3 duv12 = uv1 - uv2;
    const vec3 dp02 = p0 - p2;
    const vec3 dp12 = p1 - p2;
    const real det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
    const real inv_det = 1 / det;
    vec3 dp_du = (+duv12[1] * dp02 - duv02[1] * dp12) * inv_det;
    vec3 dp_dv = (-duv12[0] * dp02 + duv02[0] * dp12) * inv_det;
    if (std::fabs(det) < 1e-8 || length(cross(dp_du, dp_dv)) == 0) {
        // (the degenerate-normal early return of mesh.cpp:375 cannot be reached after a hit)
        const coordinate_system sys = coordinate_system_from_normal(normal);
        dp_du = sys.x;
        dp_dv = sys.y;
    }
    coordinate_system sh;
    if (has_n) {
        const vec3 n0 = s.N[s.idx[3 * tri]], n1 = s.N[s.idx[3 * tri + 1]], n2 = s.N[s.idx[3 * tri + 2]];
        sh.z = b0 * n0 + b1 * n1 + b2 * n2;
        if (reverse) sh.z = sh.z * -1.f;
        if (length_squared(sh.z) <= 0) sh.z = normal;
    } else
        sh.z = normal;
    sh.x = dp_du;  // no tangents
    sh.x = normalize(sh.x);
    sh.z = normalize(sh.z);
    sh.y = cross(sh.z, sh.x);
    if (length_squared(sh.y) > 0) {
        sh.y = normalize(sh.y);
        sh.x = normalize(cross(sh.y, sh.z));
    } else
        sh = coordinate_system_from_normal(sh.z);
    surface_interaction si;
    si.shading = sh;
    si.dp_du = dp_du;
    si.dp_dv = dp_dv;
    si.normal = face_forward(normal, sh.z);
    si.point = point;
    si.wo = -d;
    si.uv = vec2(uv.x, uv.y);
    return si;
}

// Closest hit among an instanced mesh's triangles (mesh.cpp:45-58 via its own BVH)
inline bool mesh_bvh_intersect(const orc_scene& s, const entity_data& ent, ray& r, hit_candidate& c) {
    bool found = false;
    if (ent.mesh_bvh.root < 0) return false;
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = ent.mesh_bvh.root;
    const bool neg[3] = {1 / r.direction.x < 0, 1 / r.direction.y < 0, 1 / r.direction.z < 0};
    while (sp > 0) {
        const bvh_node& n = ent.mesh_bvh.nodes[stack[--sp]];
        if (!box_intersect(n.box, r)) continue;
        if (n.leaf()) {
            for (uint32_t i = n.begin; i < n.end; i++) {
                const uint32_t tri = ent.mesh_bvh.refs[i].index;
                vec3 p0, p1, p2;
                tri_positions(s, tri, p0, p1, p2);
                real t, b0, b1, b2;
                if (mt_test(p0, p1, p2, r.origin, r.direction, r.length, t, b0, b1, b2)) {
                    r.length = t;
                    c.prim = tri;
                    c.t = t; c.b1 = b1; c.b2 = b2;
                    found = true;
                }
            }
            continue;
        }
        if (neg[n.axis]) { stack[sp++] = n.left; stack[sp++] = n.right; }
        else { stack[sp++] = n.right; stack[sp++] = n.left; }
    }
    return found;
}

// entity::intersect (entity.cpp:18-58) on a world ray whose length shrinks on a hit
inline bool entity_intersect(const orc_scene& s, uint32_t eid, uint32_t index, ray& rw, hit_candidate& best) {
    const entity_data& ent = s.ents[eid];
    if (ent.e.shape == RBW_SHAPE_NONE) return false;
    ray rl = transform_ray(ent.to_local, rw);
    hit_candidate c;
    c.entity = eid;
    c.o_local = rl.origin;
    c.d_local = rl.direction;
    c.len_local = rl.length;
    bool hit = false;
    if (ent.e.shape == RBW_SHAPE_MESH) {
        if (index != NONE) {
            vec3 p0, p1, p2;
            tri_positions(s, index, p0, p1, p2);
            real t, b0, b1, b2;
            if (mt_test(p0, p1, p2, rl.origin, rl.direction, rl.length, t, b0, b1, b2)) {
                rl.length = t;
                c.prim = index; c.t = t; c.b1 = b1; c.b2 = b2;
                hit = true;
            }
        } else {
            hit = mesh_bvh_intersect(s, ent, rl, c);
        }
    } else {
        const uint32_t sph = ent.e.first;
        real t;
        if (sphere_test(s.spheres[4 * sph + 3], rl.origin, rl.direction, rl.length, t)) {
            rl.length = t;
            c.prim = s.num_tris + sph; c.t = t; c.b1 = 0; c.b2 = 0;
            hit = true;
        }
    }
    if (!hit) return false;
    // entity.cpp:384: world length = length of the local ray mapped back to world
    rw.length = transform_ray(ent.to_world, rl).length;
    best = c;
    return true;
}

// full world-space interaction of the final candidate (entity.cpp:381-387)
inline surface_interaction candidate_interaction(const orc_scene& s, const hit_candidate& c) {
    const entity_data& ent = s.ents[c.entity];
    surface_interaction si;
    const bool reverse = (ent.e.flags & RBW_ENTITY_REVERSE_ORIENTATION) != 0;
    if (ent.e.shape == RBW_SHAPE_MESH) {
        const real b1 = c.b1, b2 = c.b2, b0 = 1 - b1 - b2;
        si = triangle_interaction(s, ent, c.prim, c.d_local, b0, b1, b2);
    } else {
        const uint32_t sph = c.prim - s.num_tris;
        si = sphere_interaction(s.spheres[4 * sph + 3], reverse, c.o_local, c.d_local, c.t);
    }
    si.entity = c.entity;
    si.prim = c.prim;
    return transform_si(ent.to_world, si);
}

// scene::intersect / intersect_with_shadow_ray through the top BVH
// (bounding_volume_hierarchy.hpp:60-160)
inline bool scene_intersect(const orc_scene& s, ray& r, bool visible_only, hit_candidate& best) {
    if (s.top.root < 0) return false;
    bool found = false;
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = s.top.root;
    const bool neg[3] = {1 / r.direction.x < 0, 1 / r.direction.y < 0, 1 / r.direction.z < 0};
    while (sp > 0) {
        const bvh_node& n = s.top.nodes[stack[--sp]];
        if (!box_intersect(n.box, r)) continue;
        if (n.leaf()) {
            for (uint32_t i = n.begin; i < n.end; i++) {
                const entity_ref& ref = s.top.refs[i];
                if (visible_only && !s.ents[ref.entity].visible()) continue;
                if (entity_intersect(s, ref.entity, ref.index, r, best)) found = true;
            }
            continue;
        }
        if (neg[n.axis]) { stack[sp++] = n.left; stack[sp++] = n.right; }
        else { stack[sp++] = n.right; stack[sp++] = n.left; }
    }
    return found;
}

// brute force over all entities and primitives (scene.cpp:86-94, mesh.cpp:49-57)
inline bool scene_intersect_linear(const orc_scene& s, ray& r, bool visible_only, hit_candidate& best) {
    bool found = false;
    for (uint32_t e = 0; e < s.ents.size(); e++) {
        const entity_data& ent = s.ents[e];
        if (visible_only && !ent.visible()) continue;
        if (ent.e.shape == RBW_SHAPE_MESH) {
            for (uint32_t i = 0; i < ent.e.count; i++)
                if (entity_intersect(s, e, ent.e.first + i, r, best)) found = true;
        } else if (ent.e.shape == RBW_SHAPE_SPHERE) {
            if (entity_intersect(s, e, NONE, r, best)) found = true;
        }
    }
    return found;
}

// ---------------------------------------------------------------- sampling source

// The reference's stratified samplers (samplers/detail/stratified_sampler.hpp): a pixel's
// reset() fills, per stratified dimension, spp jittered stratum values (1D: n strata;
// 2D: strata_x x strata_y, row by row, x then y jitter) and shuffles them (Fisher-Yates
// with pcg32::uint32(0, n - i - 1)); the 1D sampler's dimensions first, then the 2D
// sampler's (sampler_group::reset, integrator.cpp:29-33), all from the tile's shared
// pcg32 (path_integrator.cpp:50-58).  Sample i's k-th 1D (2D) draw is table[k][i] while
// k < dims, then random (one, or x then y, draws from the shared generator).
struct stratified_tables {
    uint32_t sx = 1, sy = 1, d1 = 0, d2 = 0;
    std::vector<real> t1;   // d1 x n
    std::vector<vec2> t2;   // d2 x n
    uint32_t n() const { return sx * sy; }
    void fill(pcg32& g) {
        const uint32_t N = n();
        const real one = k_one_minus_epsilon;
        t1.resize(size_t(d1) * N);
        t2.resize(size_t(d2) * N);
        for (uint32_t d = 0; d < d1; d++) {  // stratified_sampler_t<1>::stratified_sample, :99-110
            real* v = &t1[size_t(d) * N];
            const real inv = static_cast<real>(1) / static_cast<real>(N);
            for (uint32_t i = 0; i < N; i++) {
                const real delta = g.uniform_real();
                const real x = (static_cast<real>(i) + delta) * inv;
                v[i] = (one < x) ? one : x;
            }
            for (uint32_t i = 0; i < N; i++) std::swap(v[i], v[i + g.range(0, N - i - 1)]);  // :112-119
        }
        for (uint32_t d = 0; d < d2; d++) {  // stratified_sampler_t<2>::stratified_sample, :121-141
            vec2* v = &t2[size_t(d) * N];
            const real dx = static_cast<real>(1) / static_cast<real>(sx);
            const real dy = static_cast<real>(1) / static_cast<real>(sy);
            uint32_t i = 0;
            for (uint32_t y = 0; y < sy; y++)
                for (uint32_t x = 0; x < sx; x++) {
                    const real jx = g.uniform_real();
                    const real jy = g.uniform_real();
                    const real px = (static_cast<real>(x) + jx) * dx, py = (static_cast<real>(y) + jy) * dy;
                    v[i++] = vec2((one < px) ? one : px, (one < py) ? one : py);
                }
            for (uint32_t k = 0; k < N; k++) std::swap(v[k], v[k + g.range(0, N - k - 1)]);
        }
    }
};

struct sample_stream {
    pcg32 gen;
    const float* replay = nullptr;
    uint64_t pos = 0, end = 0;
    std::vector<float>* record = nullptr;
    const stratified_tables* strat = nullptr;
    uint32_t sample = 0, k1 = 0, k2 = 0;  // stratified: sample index in the pixel, dimensions drawn
    real raw() {
        real u;
        if (replay) u = pos < end ? replay[pos++] : 0.0f;
        else u = gen.uniform_real();
        if (record) record->push_back(u);
        return u;
    }
    real next() {  // a sampler1d draw
        if (strat && k1 < strat->d1) {
            const real u = strat->t1[size_t(k1++) * strat->n() + sample];
            if (record) record->push_back(u);
            return u;
        }
        return raw();
    }
    vec2 next2() {  // a sampler2d draw; random_sampler_t<2>::next: x then y
        if (strat && k2 < strat->d2) {
            const vec2 v = strat->t2[size_t(k2++) * strat->n() + sample];
            if (record) {
                record->push_back(v.x);
                record->push_back(v.y);
            }
            return v;
        }
        const real a = raw();
        const real b = raw();
        return vec2(a, b);
    }
};

// ---------------------------------------------------------------- integrator

struct render_ctx {
    const orc_scene* s;
    const rbw_camera_desc* cam;
    const rbw_render_params* p;
    mat4 raster_to_camera, camera_to_world;
    uint32_t* debug = nullptr;
};

inline bsdf build_bsdf(const orc_scene& s, const surface_interaction& si) {
    uint32_t mid = s.ents[si.entity].e.material;
    if (si.prim < s.num_tris && !s.tri_material.empty() && s.tri_material[si.prim] != NONE) mid = s.tri_material[si.prim];
    return material_bsdf(s.mats[mid]);
}

// integrator.cpp:423-426
inline real power_heuristic(real f, real g) { return (f * f) / (f * f + g * g); }

// mesh::sample (mesh.cpp:84-116) + shape::sample (shape.cpp:26-44) + surface_light::sample (surface_light.cpp:39-54),
// through entity::sample<emitter> (detail/entity.hpp:32-41).  Returns pdf 0 for the empty sample.
struct emitter_sample {
    interaction it;
    spectrum intensity{0};
    vec3 wi{0};
    real pdf = 0;
};

inline emitter_sample sample_emitter(const orc_scene& s, uint32_t eid, const interaction& ref_world, const vec2& u) {
    const entity_data& ent = s.ents[eid];
    const interaction ref = transform_interaction(ent.to_local, ref_world);
    emitter_sample es;
    // mesh::sample
    const size_t count = ent.e.count;
    const size_t which = std::min(static_cast<size_t>(std::floor(u.x * count)), count - 1);
    const vec2 rem(std::min(u.x * count - which, static_cast<real>(1) - std::numeric_limits<real>::epsilon()), u.y);
    const uint32_t tri = ent.e.first + static_cast<uint32_t>(which);
    vec3 p0, p1, p2;
    tri_positions(s, tri, p0, p1, p2);
    const vec2 b = uniform_sample_triangle(rem);
    interaction it;
    it.point = p0 * b.x + p1 * b.y + p2 * (1 - b.x - b.y);
    it.normal = normalize(cross(p1 - p0, p2 - p0));
    if (ent.e.flags & RBW_ENTITY_HAS_NORMALS) {
        const vec3 n0 = s.N[s.idx[3 * tri]], n1 = s.N[s.idx[3 * tri + 1]], n2 = s.N[s.idx[3 * tri + 2]];
        it.normal = n0 * b.x + n1 * b.y + n2 * (1 - b.x - b.y);
    }
    if (ent.e.flags & RBW_ENTITY_REVERSE_ORIENTATION) it.normal = it.normal * -1.f;
    real pdf = 1 / ent.area;
    // shape::sample
    vec3 wi = it.point - ref.point;
    if (length_squared(wi) == 0) return es;
    wi = normalize(wi);
    pdf = pdf * distance_squared(ref.point, it.point) / std::fabs(dot(it.normal, -wi));
    if (std::isinf(pdf)) return es;
    // surface_light::sample
    if (pdf == 0) return es;
    const vec3 wl = normalize(it.point - ref.point);
    const spectrum L(ent.e.radiance[0], ent.e.radiance[1], ent.e.radiance[2]);
    const spectrum intensity = dot(it.normal, -wl) > 0 ? L : spectrum(0);
    // emitter_sample::transform (emitter.cpp:18-26)
    es.it = transform_interaction(ent.to_world, it);
    es.intensity = intensity;
    es.wi = transform_vector(ent.to_world, wl);
    es.pdf = pdf;
    return es;
}

// surface_light::evaluate via entity::evaluate<emitter> (detail/entity.hpp:74-79, surface_light.cpp:12-16)
inline spectrum emitter_evaluate(const orc_scene& s, uint32_t eid, const interaction& it_world, const vec3& w_world) {
    const entity_data& ent = s.ents[eid];
    const interaction it = transform_interaction(ent.to_local, it_world);
    const vec3 w = transform_vector(ent.to_local, w_world);
    const spectrum L(ent.e.radiance[0], ent.e.radiance[1], ent.e.radiance[2]);
    return dot(it.normal, w) > 0 ? L : spectrum(0);
}

// entity::pdf<emitter> -> surface_light::pdf -> shape::pdf (shape.cpp:46-60): a cast against the light's own shape
// (raw: the reference interaction is used untransformed, keeping a medium vertex's zero normal)
inline real emitter_pdf(const orc_scene& s, uint32_t eid, const interaction& ref_world, const vec3& wi_world, counters& cnt,
                        bool raw = false) {
    const entity_data& ent = s.ents[eid];
    const interaction ref = raw ? ref_world : transform_interaction(ent.to_local, ref_world);
    const vec3 wi = transform_vector(ent.to_local, wi_world);
    ray r = spawn_ray(ref, wi);
    cnt.light++;
    hit_candidate c;
    bool hit = false;
    if (ent.e.shape == RBW_SHAPE_MESH) {
        if (ent.instanced) {
            hit = mesh_bvh_intersect(s, ent, r, c);
        } else {
            for (uint32_t i = 0; i < ent.e.count; i++) {
                vec3 p0, p1, p2;
                tri_positions(s, ent.e.first + i, p0, p1, p2);
                real t, b0, b1, b2;
                if (mt_test(p0, p1, p2, r.origin, r.direction, r.length, t, b0, b1, b2)) {
                    r.length = t;
                    c.prim = ent.e.first + i; c.t = t; c.b1 = b1; c.b2 = b2;
                    hit = true;
                }
            }
        }
    }
    if (!hit) return 0;
    const real b1 = c.b1, b2 = c.b2, b0 = 1 - b1 - b2;
    const surface_interaction hi = triangle_interaction(s, ent, c.prim, r.direction, b0, b1, b2);
    const real pdf = distance_squared(ref.point, hi.point) / (std::fabs(dot(hi.normal, -wi)) * ent.area);
    if (std::isinf(pdf)) return 0;
    return pdf;
}

// uniform_sample_one_emitter, surface overload (integrator.cpp:101-201), media = false
inline spectrum uniform_sample_one_emitter(const render_ctx& ctx, sample_stream& smp, const surface_interaction& si,
                                           const bsdf& fn, counters& cnt, uint32_t* dbg) {
    const orc_scene& s = *ctx.s;
    spectrum L(0);
    {
        // integrator.cpp:84-99
        const size_t ne = s.emitters.size();
        const real u1 = smp.next();
        const size_t which = std::min(static_cast<size_t>(std::floor(u1 * ne)), ne - 1);
        const real pick_pdf = static_cast<real>(1) / ne;
        const uint32_t eid = s.emitters[which];
        emitter_sample es = sample_emitter(s, eid, si, smp.next2());
        es.pdf = es.pdf * pick_pdf;
        if (!is_black(es.intensity) && es.pdf > 0) {
            const vec3 wi = world_to_local(si.shading, es.wi);
            const vec3 wo = world_to_local(si.shading, si.wo);
            spectrum f = fn.evaluate(wo, wi);
            const real fp = fn.pdf(wo, wi);
            f = f * std::fabs(dot(es.wi, si.shading.z));
            if (!is_black(f) && fp > 0) {
                ray sr = spawn_ray_to(si, es.it.point);
                cnt.shadow++;
                hit_candidate c;
                const bool hit = scene_intersect(s, sr, true, c);
                if (dbg) dbg[1] = hit ? c.prim : RBW_MISS;
                if (!hit || c.entity == eid) {
                    const real w = power_heuristic(es.pdf, fp);
                    L = L + f * es.intensity * w / es.pdf;
                }
            }
        }
    }
    {
        vec3 wi;
        spectrum f;
        real pdf = 0;
        const vec2 u = smp.next2();
        if (fn.sample(si, u, wi, f, pdf)) {
            f = f * std::fabs(dot(wi, si.shading.z));
            if (!is_black(f) && pdf > 0) {
                // find_emitter (integrator.cpp:49-82)
                ray er = spawn_ray(si, wi);
                cnt.mis++;
                hit_candidate c;
                const bool hit = scene_intersect(s, er, true, c);
                if (dbg) dbg[2] = hit ? c.prim : RBW_MISS;
                if (hit && s.ents[c.entity].emitter()) {
                    const real pick_pdf = static_cast<real>(1) / s.emitters.size();
                    const surface_interaction ei = candidate_interaction(s, c);
                    const spectrum Le = emitter_evaluate(s, c.entity, ei, -wi);
                    const real lp = emitter_pdf(s, c.entity, si, wi, cnt) * pick_pdf;
                    if (!is_black(Le) && lp > 0) {
                        const real w = power_heuristic(pdf, lp);
                        L = L + f * Le * w / pdf;
                    }
                }
            }
        }
    }
    return L;
}

struct path_state {
    spectrum value{0}, beta{1};
    bool specular = false;
    real eta = 1;
    ray r;
};

// path_integrator::trace (path_integrator.cpp:15-47) with sample_surface_interaction (integrator.cpp:329-404)
inline spectrum trace_path(const render_ctx& ctx, sample_stream& smp, const ray& first, counters& cnt, uint32_t* dbg) {
    const orc_scene& s = *ctx.s;
    path_state ps;
    ps.r = first;
    const int max_depth = static_cast<int>(ctx.p->max_depth);
    for (int bounces = 0; bounces < max_depth; bounces++) {
        uint32_t* d = dbg ? dbg + 3 * static_cast<size_t>(bounces) : nullptr;
        cnt.extension++;
        hit_candidate c;
        ray r = ps.r;
        const bool hit = scene_intersect(s, r, false, c);
        if (d) d[0] = hit ? c.prim : RBW_MISS;
        // sample_surface_interaction
        if (!hit) break;  // no environment lights
        const surface_interaction si = candidate_interaction(s, c);
        const entity_data& ent = s.ents[si.entity];
        if ((bounces == 0 || ps.specular) && ent.emitter())
            ps.value = ps.value + ps.beta * emitter_evaluate(s, si.entity, si, -ps.r.direction);
        if (ent.e.material == NONE) {
            // invisible entity: continue straight through without counting the bounce
            ps.r = spawn_ray(si, ps.r.direction);
            bounces--;
        } else {
            const bsdf fn = build_bsdf(s, si);
            if (fn.count() != 0)
                ps.value = ps.value + ps.beta * uniform_sample_one_emitter(ctx, smp, si, fn, cnt, d);
            vec3 wi;
            spectrum f;
            real pdf = 0;
            const vec2 u = smp.next2();
            const bool ok = fn.sample(si, u, wi, f, pdf);
            if (!ok || is_black(f) || pdf == 0) break;
            ps.beta = ps.beta * (f * std::fabs(dot(wi, si.shading.z)) / pdf);
            ps.specular = false;  // Lambert and microfacet lobes are not specular
            ps.r = spawn_ray(si, wi);
        }
        // Russian roulette (path_integrator.cpp:35-43)
        const real mc = smax_component(ps.beta * ps.eta);
        if (mc < ctx.p->rr_threshold && bounces > 3) {
            const real q = fmax_g(static_cast<real>(0.05), 1 - mc);
            if (smp.next() < q) break;
            ps.beta = ps.beta / (1 - q);
        }
    }
    return ps.value;
}

// direct_integrator::trace (direct_integrator.cpp:20-128), recursive through surfaces
// without scattering functions as the reference recurses
inline spectrum trace_direct(const render_ctx& ctx, sample_stream& smp, const ray& r, counters& cnt, int guard = 0) {
    const orc_scene& s = *ctx.s;
    if (s.emitters.empty()) return spectrum(0);
    spectrum L(0);
    cnt.extension++;
    hit_candidate c;
    ray rr = r;
    if (!scene_intersect(s, rr, false, c)) return L;  // no environment emitters
    const surface_interaction si = candidate_interaction(s, c);
    const entity_data& ent = s.ents[si.entity];
    if (ent.emitter()) L = L + emitter_evaluate(s, si.entity, si, -r.direction);
    bsdf fn;
    if (ent.e.material != NONE) fn = build_bsdf(s, si);
    if (fn.count() == 0) {
        if (guard > 1024) return L;
        return L + trace_direct(ctx, smp, spawn_ray(si, r.direction), cnt, guard + 1);
    }
    const uint32_t ne = ctx.p->emitter_samples ? ctx.p->emitter_samples : 1u;
    const uint32_t nb = ctx.p->bsdf_samples ? ctx.p->bsdf_samples : 1u;
    const size_t sum = size_t(ne) + nb;
    const real we = static_cast<real>(1) / ne, wb = static_cast<real>(1) / nb;
    const real fe = static_cast<real>(ne) / sum, fb = static_cast<real>(nb) / sum;
    const vec3 wo = world_to_local(si.shading, si.wo);
    for (uint32_t i = 0; i < ne; i++) {
        const size_t n = s.emitters.size();
        const size_t which = std::min(static_cast<size_t>(std::floor(smp.next() * n)), n - 1);
        const real pick_pdf = static_cast<real>(1) / n;
        const uint32_t eid = s.emitters[which];
        emitter_sample es = sample_emitter(s, eid, si, smp.next2());
        es.pdf = es.pdf * pick_pdf;
        if (is_black(es.intensity) || !(es.pdf > 0)) continue;
        const vec3 wi = world_to_local(si.shading, es.wi);
        spectrum f = fn.evaluate(wo, wi);
        const real fp = fn.pdf(wo, wi);
        f = f * std::fabs(dot(es.wi, si.shading.z));
        if (is_black(f) || fp <= 0) continue;
        ray sr = spawn_ray_to(si, es.it.point);
        cnt.shadow++;
        hit_candidate sc;
        const bool hit = scene_intersect(s, sr, true, sc);
        if (!hit || sc.entity == eid) {
            const real w = power_heuristic(es.pdf * fe, fp * fb) * we;
            L = L + f * es.intensity * w / es.pdf;
        }
    }
    for (uint32_t i = 0; i < nb; i++) {
        vec3 wi;
        spectrum f;
        real pdf = 0;
        if (!fn.sample(si, smp.next2(), wi, f, pdf)) continue;
        f = f * std::fabs(dot(wi, si.shading.z));
        if (is_black(f) || !(pdf > 0)) continue;
        ray er = spawn_ray(si, wi);
        cnt.mis++;
        hit_candidate ec;
        if (!scene_intersect(s, er, true, ec) || !s.ents[ec.entity].emitter()) continue;
        const real pick_pdf = static_cast<real>(1) / s.emitters.size();
        const surface_interaction ei = candidate_interaction(s, ec);
        const spectrum Le = emitter_evaluate(s, ec.entity, ei, -wi);
        const real lp = emitter_pdf(s, ec.entity, si, wi, cnt) * pick_pdf;
        if (is_black(Le) || lp <= 0) continue;
        const real w = power_heuristic(pdf * fb, lp * fe) * wb;
        L = L + f * Le * w / pdf;
    }
    return L;
}

// ---------------------------------------------------------------- media

constexpr real k_one_over_four_pi = static_cast<real>(0.0795774715459476678844418816862571810);

// henyey_greenstein_phase_function (utilities.cpp:46-51)
inline real hg_phase(real cos_theta, real g) {
    const real den = 1 + g * g + 2 * g * cos_theta;
    return k_one_over_four_pi * (1 - g * g) / (den * std::sqrt(den));
}

// henyey_greenstein::sample (henyey_greenstein.cpp:17-48); returns wi, pdf = the phase value
inline vec3 hg_sample(const vec3& wo, real g, const vec2& u, real& pdf) {
    real cos_t;
    if (std::fabs(g) >= 1e-3) {
        const real sq = (1 - g * g) / (1 - g + 2 * g * u.x);
        cos_t = (1 + g * g - sq * sq) / (2 * g);
    } else {
        cos_t = 1 - 2 * u.x;
    }
    const real sin_t = std::sqrt(fmax_g(static_cast<real>(0), 1 - cos_t * cos_t));
    const real phi = k_two_pi * u.y;
    const vec3 wl(sin_t * rcos(phi), sin_t * rsin(phi), cos_t);
    coordinate_system sys = coordinate_system_from_normal(wo);
    sys.z = -wo;
    pdf = hg_phase(-cos_t, g);
    return local_to_world(sys, wl);
}

// image_texture_t<3, real>::sample (textures/detail/image_texture.hpp:119-162)
inline real grid_sample(const orc_scene& s, const rbw_medium& m, const vec3& p) {
    const int n[3] = {static_cast<int>(m.grid_size[0]), static_cast<int>(m.grid_size[1]), static_cast<int>(m.grid_size[2])};
    const real pc[3] = {p.x, p.y, p.z};
    real uv[3];
    int a[3], b[3];
    real f[3];
    for (int k = 0; k < 3; k++) {
        const real md = pc[k] - 1 * std::floor(pc[k] / 1);  // glm::mod
        uv[k] = static_cast<real>(md * static_cast<real>(n[k]) - 0.5);
        a[k] = std::min(std::max(static_cast<int>(std::floor(uv[k])), 0), n[k] - 1);
        b[k] = std::min(std::max(a[k] + 1, 0), n[k] - 1);
        f[k] = std::min(std::max(uv[k] - static_cast<real>(a[k]), static_cast<real>(0)), static_cast<real>(1));
    }
    auto at = [&](int x, int y, int z) {
        return s.grid[m.grid_offset + static_cast<size_t>(z) * n[0] * n[1] + static_cast<size_t>(y) * n[0] + x];
    };
    auto lerp = [](real v0, real v1, real x) { return v0 * (1 - x) + v1 * x; };
    const real c00 = lerp(at(a[0], a[1], a[2]), at(b[0], a[1], a[2]), f[0]);
    const real c01 = lerp(at(a[0], a[1], b[2]), at(b[0], a[1], b[2]), f[0]);
    const real c10 = lerp(at(a[0], b[1], a[2]), at(b[0], b[1], a[2]), f[0]);
    const real c11 = lerp(at(a[0], b[1], b[2]), at(b[0], b[1], b[2]), f[0]);
    return lerp(lerp(c00, c10, f[1]), lerp(c01, c11, f[1]), f[2]);
}

// the grid's local ray: world_to_local applied without re-normalising (heterogeneous_medium.cpp:28-33)
inline void grid_local_ray(const rbw_medium& m, const ray& r, vec3& lo, vec3& ld) {
    const float* w = m.world_to_local;
    lo = vec3(w[0] * r.origin.x + w[1] * r.origin.y + w[2] * r.origin.z + w[3],
              w[4] * r.origin.x + w[5] * r.origin.y + w[6] * r.origin.z + w[7],
              w[8] * r.origin.x + w[9] * r.origin.y + w[10] * r.origin.z + w[11]);
    ld = vec3(w[0] * r.direction.x + w[1] * r.direction.y + w[2] * r.direction.z,
              w[4] * r.direction.x + w[5] * r.direction.y + w[6] * r.direction.z,
              w[8] * r.direction.x + w[9] * r.direction.y + w[10] * r.direction.z);
}

// heterogeneous_medium::intersect (heterogeneous_medium.cpp:122-140) with the unit box
inline bool unit_box(const vec3& o, const vec3& d, real len, real& t0, real& t1) {
    t0 = 0;
    t1 = len;
    for (int k = 0; k < 3; k++) {
        const real inv = 1 / d[k];
        real near_t = (0 - o[k]) * inv, far_t = (1 - o[k]) * inv;
        if (near_t > far_t) std::swap(near_t, far_t);
        t0 = fmax_g(t0, near_t);
        t1 = fmin_g(t1, far_t);
        if (t0 > t1) return false;
    }
    return true;
}

struct medium_sample_result {
    bool scattered = false;
    vec3 point{0};
    spectrum value{1};
    real g = 0;
};

// medium_info::sample (medium.cpp:45-53) -> homogeneous_medium::sample (homogeneous_medium.cpp:20-58)
// or heterogeneous_medium::sample (heterogeneous_medium.cpp:70-120).  A grid collision's world
// point is the world ray at the collision parameter.
inline medium_sample_result medium_sample(const orc_scene& s, uint32_t mid, sample_stream& smp, const ray& r) {
    medium_sample_result out;
    const rbw_medium& m = s.media[mid];
    out.g = m.g;
    if (m.type == RBW_MEDIUM_HOMOGENEOUS) {
        const spectrum st = spectrum(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]) + spectrum(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
        const size_t channel = std::min(static_cast<size_t>(smp.next() * 3), static_cast<size_t>(2));  // sample_function.cpp:97-102
        const real t = fmin_g(-flog(1 - smp.next()) / st[static_cast<int>(channel)], r.length);
        const bool sampled = t < r.length;
        const real tc = fmin_g(t, std::numeric_limits<real>::max());
        const spectrum value(fexp(-st.x * tc), fexp(-st.y * tc), fexp(-st.z * tc));
        const spectrum density = sampled ? st * value : value;
        real pdf = 0;
        pdf = pdf + density.x;
        pdf = pdf + density.y;
        pdf = pdf + density.z;
        pdf = pdf / 3;
        if (pdf == 0) pdf = 1;
        if (!sampled) {
            out.value = value / pdf;
            return out;
        }
        out.scattered = true;
        out.point = r.origin + r.direction * t;
        out.value = value * spectrum(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]) / pdf;
        return out;
    }
    vec3 lo, ld;
    grid_local_ray(m, r, lo, ld);
    real t0, t1;
    if (!unit_box(lo, ld, r.length, t0, t1)) return out;
    const real inv = s.inv_max_density[mid];
    real t = t0;
    while (true) {
        t = t - flog(1 - smp.next()) * inv;
        if (t >= t1) break;
        const real sigma_t = grid_sample(s, m, lo + ld * t);
        if (sigma_t * inv > smp.next()) {
            out.scattered = true;
            out.point = r.origin + r.direction * t;
            out.value = spectrum(m.albedo[0], m.albedo[1], m.albedo[2]);
            return out;
        }
    }
    return out;
}

// medium_info::evaluate (medium.cpp:35-43): vacuum 1; homogeneous_medium::evaluate
// (homogeneous_medium.cpp:12-18); heterogeneous_medium::evaluate, ratio tracking (:24-68)
inline spectrum medium_evaluate(const orc_scene& s, uint32_t mid, sample_stream& smp, const ray& r) {
    if (mid == NONE) return spectrum(1);
    const rbw_medium& m = s.media[mid];
    if (m.type == RBW_MEDIUM_HOMOGENEOUS) {
        const spectrum st = spectrum(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]) + spectrum(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
        const real tc = fmin_g(r.length, std::numeric_limits<real>::max());
        return spectrum(fexp(-st.x * tc), fexp(-st.y * tc), fexp(-st.z * tc));
    }
    vec3 lo, ld;
    grid_local_ray(m, r, lo, ld);
    real t0, t1;
    if (!unit_box(lo, ld, r.length, t0, t1)) return spectrum(1);
    const real inv = s.inv_max_density[mid];
    real value = 1, t = t0;
    while (true) {
        t = t - flog(1 - smp.next()) * inv;
        if (t >= t1) break;
        const real sigma_t = grid_sample(s, m, lo + ld * t);
        value = value * (1 - fmax_g(static_cast<real>(0), sigma_t * inv));
    }
    return spectrum(value);
}

// medium_info(entity, normal, wi) (medium.cpp:24-33)
inline uint32_t medium_after(const orc_scene& s, uint32_t entity, const vec3& normal, const vec3& wi) {
    return dot(normal, wi) > 0 ? s.ent_media[2 * entity] : s.ent_media[2 * entity + 1];
}
inline bool has_media(const orc_scene& s, uint32_t entity) { return (s.ents[entity].e.flags & RBW_ENTITY_HAS_MEDIA) != 0; }

// scene::evaluate_media_beam (scene.cpp:114-148)
inline spectrum evaluate_media_beam(const orc_scene& s, sample_stream& smp, uint32_t medium, const interaction& from,
                                    const vec3& to, counters& cnt) {
    spectrum L(1);
    ray beam = spawn_ray_to(from, to);
    hit_candidate c;
    cnt.shadow++;
    bool hit = scene_intersect(s, beam, false, c);
    while (hit) {
        L = L * medium_evaluate(s, medium, smp, beam);
        if (s.ents[c.entity].visible()) return spectrum(0);
        const surface_interaction si = candidate_interaction(s, c);
        beam = spawn_ray_to(si, to);
        if (has_media(s, c.entity)) medium = medium_after(s, c.entity, si.normal, beam.direction);
        c = hit_candidate();
        cnt.shadow++;
        hit = scene_intersect(s, beam, false, c);
    }
    L = L * medium_evaluate(s, medium, smp, beam);
    return L;
}

// uniform_sample_one_emitter, surface overload with media = true (integrator.cpp:101-201)
inline spectrum sample_one_emitter_media(const render_ctx& ctx, sample_stream& smp, const surface_interaction& si,
                                         const bsdf& fn, uint32_t medium, counters& cnt) {
    const orc_scene& s = *ctx.s;
    spectrum L(0);
    {
        const size_t ne = s.emitters.size();
        const size_t which = std::min(static_cast<size_t>(std::floor(smp.next() * ne)), ne - 1);
        const real pick_pdf = static_cast<real>(1) / ne;
        const uint32_t eid = s.emitters[which];
        emitter_sample es = sample_emitter(s, eid, si, smp.next2());
        es.pdf = es.pdf * pick_pdf;
        if (!is_black(es.intensity) && es.pdf > 0) {
            const vec3 wi = world_to_local(si.shading, es.wi);
            const vec3 wo = world_to_local(si.shading, si.wo);
            spectrum f = fn.evaluate(wo, wi);
            const real fp = fn.pdf(wo, wi);
            f = f * std::fabs(dot(es.wi, si.shading.z));
            if (!is_black(f) && fp > 0) {
                ray sr = spawn_ray_to(si, es.it.point);
                cnt.shadow++;
                hit_candidate c;
                const bool hit = scene_intersect(s, sr, true, c);
                if (!hit || c.entity == eid) {
                    f = f * evaluate_media_beam(s, smp, medium, si, es.it.point, cnt);
                    const real w = power_heuristic(es.pdf, fp);
                    L = L + f * es.intensity * w / es.pdf;
                }
            }
        }
    }
    {
        vec3 wi;
        spectrum f;
        real pdf = 0;
        const vec2 u = smp.next2();
        if (fn.sample(si, u, wi, f, pdf)) {
            f = f * std::fabs(dot(wi, si.shading.z));
            if (!is_black(f) && pdf > 0) {
                ray er = spawn_ray(si, wi);
                cnt.mis++;
                hit_candidate c;
                const bool hit = scene_intersect(s, er, true, c);
                if (hit && s.ents[c.entity].emitter()) {
                    const real pick_pdf = static_cast<real>(1) / s.emitters.size();
                    const surface_interaction ei = candidate_interaction(s, c);
                    const spectrum Le = emitter_evaluate(s, c.entity, ei, -wi);
                    const real lp = emitter_pdf(s, c.entity, si, wi, cnt) * pick_pdf;
                    if (!is_black(Le) && lp > 0) {
                        f = f * evaluate_media_beam(s, smp, medium, si, ei.point, cnt);
                        const real w = power_heuristic(pdf, lp);
                        L = L + f * Le * w / pdf;
                    }
                }
            }
        }
    }
    return L;
}

// uniform_sample_one_emitter, medium overload (integrator.cpp:203-291).  The emitter pdf
// of a phase sample is entity::pdf<emitter> of the medium interaction, whose zero normal
// transform_normal normalises to NaN (detail/entity.hpp:66-72): the test `emitter_pdf > 0`
// fails and the phase-sampled emitter contributes nothing.  medium_mis = 1 evaluates the
// pdf from the untransformed interaction instead.
inline spectrum sample_one_emitter_medium(const render_ctx& ctx, sample_stream& smp, const interaction& mi, real g,
                                          uint32_t medium, counters& cnt) {
    const orc_scene& s = *ctx.s;
    spectrum L(0);
    {
        const size_t ne = s.emitters.size();
        const size_t which = std::min(static_cast<size_t>(std::floor(smp.next() * ne)), ne - 1);
        const real pick_pdf = static_cast<real>(1) / ne;
        const uint32_t eid = s.emitters[which];
        emitter_sample es = sample_emitter(s, eid, mi, smp.next2());
        es.pdf = es.pdf * pick_pdf;
        if (!is_black(es.intensity) && es.pdf > 0) {
            const real phase = hg_phase(dot(mi.wo, es.wi), g);
            const spectrum f(phase);
            const real fp = phase;
            if (!is_black(f) && fp > 0) {
                ray sr = spawn_ray_to(mi, es.it.point);
                cnt.shadow++;
                hit_candidate c;
                const bool hit = scene_intersect(s, sr, true, c);
                if (!hit || c.entity == eid) {
                    const spectrum beam = evaluate_media_beam(s, smp, medium, mi, es.it.point, cnt);
                    const real w = power_heuristic(es.pdf, fp);
                    L = L + f * es.intensity * beam * w / es.pdf;
                }
            }
        }
    }
    {
        real fp = 0;
        const vec3 wi = hg_sample(mi.wo, g, smp.next2(), fp);
        const spectrum f(fp);
        if (!is_black(f) && fp > 0) {
            ray er = spawn_ray(mi, wi);
            cnt.mis++;
            hit_candidate c;
            const bool hit = scene_intersect(s, er, true, c);
            if (hit && s.ents[c.entity].emitter()) {
                const real pick_pdf = static_cast<real>(1) / s.emitters.size();
                const surface_interaction ei = candidate_interaction(s, c);
                const spectrum Le = emitter_evaluate(s, c.entity, ei, -wi);
                const real lp = emitter_pdf(s, c.entity, mi, wi, cnt, ctx.p->medium_mis != 0) * pick_pdf;
                if (!is_black(Le) && lp > 0) {
                    const spectrum beam = evaluate_media_beam(s, smp, medium, mi, ei.point, cnt);
                    const real w = power_heuristic(fp, lp);
                    L = L + f * Le * beam * w / fp;
                }
            }
        }
    }
    return L;
}

// volume_path_integrator::trace (volume_path_integrator.cpp:13-79); camera rays start in vacuum
inline spectrum trace_volume(const render_ctx& ctx, sample_stream& smp, const ray& first, counters& cnt) {
    const orc_scene& s = *ctx.s;
    path_state ps;
    ps.r = first;
    uint32_t medium = NONE;
    const int max_depth = static_cast<int>(ctx.p->max_depth);
    for (int bounces = 0; bounces < max_depth; bounces++) {
        cnt.extension++;
        hit_candidate c;
        const bool hit = scene_intersect(s, ps.r, false, c);  // the ray's length becomes the hit distance
        bool surface = true;
        if (medium != NONE) {
            const medium_sample_result ms = medium_sample(s, medium, smp, ps.r);
            ps.beta = ps.beta * ms.value;
            if (is_black(ps.beta)) break;
            if (ms.scattered) {
                // sample_medium_interaction (integrator.cpp:406-421)
                interaction mi;
                mi.point = ms.point;
                mi.normal = vec3(0);
                mi.wo = -ps.r.direction;
                ps.value = ps.value + ps.beta * sample_one_emitter_medium(ctx, smp, mi, ms.g, medium, cnt);
                real pp;
                const vec3 wi = hg_sample(mi.wo, ms.g, smp.next2(), pp);
                ps.specular = false;
                ps.r = spawn_ray(mi, wi);
                surface = false;
            }
        }
        if (surface) {
            // sample_surface_interaction (integrator.cpp:329-404), media = true
            if (!hit) break;  // no environment lights
            const surface_interaction si = candidate_interaction(s, c);
            const entity_data& ent = s.ents[si.entity];
            if ((bounces == 0 || ps.specular) && ent.emitter())
                ps.value = ps.value + ps.beta * emitter_evaluate(s, si.entity, si, -ps.r.direction);
            if (ent.e.material == NONE) {
                ps.r = spawn_ray(si, ps.r.direction);
                bounces--;
            } else {
                const bsdf fn = build_bsdf(s, si);
                if (fn.count() != 0)
                    ps.value = ps.value + ps.beta * sample_one_emitter_media(ctx, smp, si, fn, medium, cnt);
                vec3 wi;
                spectrum f;
                real pdf = 0;
                const vec2 u = smp.next2();
                const bool ok = fn.sample(si, u, wi, f, pdf);
                if (!ok || is_black(f) || pdf == 0) break;
                ps.beta = ps.beta * (f * std::fabs(dot(wi, si.shading.z)) / pdf);
                ps.specular = false;
                ps.r = spawn_ray(si, wi);
            }
            if (has_media(s, si.entity)) medium = medium_after(s, si.entity, si.normal, ps.r.direction);
        }
        const real mc = smax_component(ps.beta * ps.eta);
        if (mc < ctx.p->rr_threshold && bounces > 3) {
            const real q = fmax_g(static_cast<real>(0.05), 1 - mc);
            if (smp.next() < q) break;
            ps.beta = ps.beta / (1 - q);
        }
    }
    return ps.value;
}

// perspective_camera::sample(position, sample) (perspective_camera.cpp:78-99)
inline ray camera_ray(const render_ctx& ctx, const vec2& pos, const vec2& lens_u) {
    const vec3 target = ctx.raster_to_camera.mul(vec3(pos.x, pos.y, 0), 1);
    ray cr = make_ray(normalize(target), vec3(0));
    if (ctx.cam->lens_radius > 0) {
        const vec3 focus = cr.origin + cr.direction * (ctx.cam->focus_distance / cr.direction.z);
        const vec2 disk = concentric_sample_disk(lens_u);
        const vec2 lens(disk.x * ctx.cam->lens_radius, disk.y * ctx.cam->lens_radius);
        cr.origin = vec3(lens.x, lens.y, 0);
        cr.direction = normalize(focus - cr.origin);
    }
    const xform c2w{ctx.camera_to_world, ctx.camera_to_world};
    return transform_ray(c2w, cr);
}

struct pixel_acc {
    spectrum sum{0};
    real weight = 0;
};

struct film_tile {
    int32_t fmin[2], fmax[2];  // filter region
    std::vector<pixel_acc> px;
};

}  // namespace orc

using namespace orc;

extern "C" {

int orc_scene_create(const rbw_scene_desc* d, orc_scene** out) {
    if (!d || !out) return RBW_ERR_INVALID_ARGUMENT;
    auto s = std::make_unique<orc_scene>();
    s->num_tris = d->num_triangles;
    s->num_spheres = d->num_spheres;
    s->P.resize(d->num_vertices);
    for (uint32_t i = 0; i < d->num_vertices; i++) s->P[i] = vec3(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2]);
    if (d->normals) {
        s->N.resize(d->num_vertices);
        for (uint32_t i = 0; i < d->num_vertices; i++) s->N[i] = vec3(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
    }
    if (d->uvs) {
        s->UV.resize(d->num_vertices);
        for (uint32_t i = 0; i < d->num_vertices; i++) s->UV[i] = vec2(d->uvs[2 * i], d->uvs[2 * i + 1]);
    }
    s->idx.assign(d->indices, d->indices + 3 * static_cast<size_t>(d->num_triangles));
    s->tri_entity.assign(d->triangle_entity, d->triangle_entity + d->num_triangles);
    if (d->triangle_material) s->tri_material.assign(d->triangle_material, d->triangle_material + d->num_triangles);
    if (d->num_spheres) {
        s->spheres.assign(d->spheres, d->spheres + 4 * static_cast<size_t>(d->num_spheres));
        s->sphere_entity.assign(d->sphere_entity, d->sphere_entity + d->num_spheres);
    }
    s->mats.assign(d->materials, d->materials + d->num_materials);
    if (d->entity_media) {
        s->media.assign(d->media, d->media + d->num_media);
        s->ent_media.assign(d->entity_media, d->entity_media + 2 * static_cast<size_t>(d->num_entities));
        s->grid.assign(d->medium_grid, d->medium_grid + d->medium_grid_count);
        for (const rbw_medium& m : s->media) {
            float mx = 0;
            if (m.type == RBW_MEDIUM_GRID)
                for (size_t k = 0; k < size_t(m.grid_size[0]) * m.grid_size[1] * m.grid_size[2]; k++)
                    mx = std::max(mx, s->grid[m.grid_offset + k]);
            s->inv_max_density.push_back(mx > 0 ? 1 / mx : 0);
        }
    }
    s->ents.resize(d->num_entities);
    // reference_count of every shape is 1 (each entity owns its shape)
    constexpr size_t sub_shape_threshold = 1 << 17;  // scene.cpp:43
    for (uint32_t e = 0; e < d->num_entities; e++) {
        entity_data& ent = s->ents[e];
        ent.e = d->entities[e];
        if (ent.e.shape == RBW_SHAPE_SPHERE) {
            const float* sp = &s->spheres[4 * ent.e.first];
            ent.to_world = xform::translate(vec3(sp[0], sp[1], sp[2]));
        } else
            ent.to_world = xform::identity();
        ent.to_local = ent.to_world.inverse();
        if (ent.e.flags & RBW_ENTITY_EMITTER) s->emitters.push_back(e);
        if (ent.e.shape == RBW_SHAPE_MESH) {
            // mesh::area(transform) (mesh.cpp:133-152)
            real area = 0;
            for (uint32_t i = 0; i < ent.e.count; i++) {
                vec3 p0, p1, p2;
                tri_positions(*s, ent.e.first + i, p0, p1, p2);
                p0 = transform_point(ent.to_world, p0);
                p1 = transform_point(ent.to_world, p1);
                p2 = transform_point(ent.to_world, p2);
                area = area + static_cast<real>(0.5) * length(cross(p2 - p0, p2 - p1));
            }
            ent.area = area;
            ent.instanced = ent.e.count > sub_shape_threshold;
        }
    }
    // top-level references (scene.cpp:47-80)
    for (uint32_t e = 0; e < d->num_entities; e++) {
        entity_data& ent = s->ents[e];
        if (ent.e.shape == RBW_SHAPE_MESH) {
            auto tri_box = [&](uint32_t tri) {
                vec3 p0, p1, p2;
                tri_positions(*s, tri, p0, p1, p2);
                p0 = transform_point(ent.to_world, p0);
                p1 = transform_point(ent.to_world, p1);
                p2 = transform_point(ent.to_world, p2);
                bound3 b(p0, p1);
                b.union_it(p2);
                return b;
            };
            if (!ent.instanced) {
                for (uint32_t i = 0; i < ent.e.count; i++) s->top.refs.push_back({e, ent.e.first + i, tri_box(ent.e.first + i)});
            } else {
                // mesh::build_accelerator (mesh.cpp:159-172): boxes in the mesh's own space
                bound3 mb;
                for (uint32_t i = 0; i < ent.e.count; i++) {
                    vec3 p0, p1, p2;
                    tri_positions(*s, ent.e.first + i, p0, p1, p2);
                    bound3 b(p0, p1);
                    b.union_it(p2);
                    ent.mesh_bvh.refs.push_back({e, ent.e.first + i, b});
                    if (i == 0) mb = b; else mb.union_it(b);
                }
                ent.mesh_bvh.build();
                // transform(mBoundingBox): 8 corners (transform.cpp:80-93); identity for meshes
                bound3 wb(transform_point(ent.to_world, mb.min), transform_point(ent.to_world, mb.max));
                s->top.refs.push_back({e, NONE, wb});
            }
        } else if (ent.e.shape == RBW_SHAPE_SPHERE) {
            // sphere::bounding_box (sphere.cpp:93-100)
            const real r = s->spheres[4 * ent.e.first + 3];
            const vec3 c = transform_point(ent.to_world, vec3(0));
            const vec3 p = transform_point(ent.to_world, vec3(0, 0, r));
            const real rad = length(p - c);
            s->top.refs.push_back({e, s->num_tris + ent.e.first, bound3(c - vec3(rad), c + vec3(rad))});
        }
    }
    s->top.build();
    *out = s.release();
    return RBW_OK;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

uint32_t orc_scene_num_bvh_nodes(const orc_scene* s) {
    size_t n = s->top.nodes.size();
    for (const auto& e : s->ents) n += e.mesh_bvh.nodes.size();
    return static_cast<uint32_t>(n);
}

static int trace_rays_impl(orc_scene* s, uint32_t n, const float* o, const float* dir, const float* tmax,
                           uint32_t visible_only, rbw_hit* out, bool linear) {
    for (uint32_t i = 0; i < n; i++) {
        ray r{vec3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), tmax[i]};
        hit_candidate c;
        const bool hit = linear ? scene_intersect_linear(*s, r, visible_only != 0, c)
                                : scene_intersect(*s, r, visible_only != 0, c);
        out[i].prim = hit ? c.prim : RBW_MISS;
        out[i].entity = hit ? c.entity : RBW_NONE;
        out[i].t = hit ? c.t : 0;
        out[i].b1 = hit ? c.b1 : 0;
        out[i].b2 = hit ? c.b2 : 0;
    }
    return RBW_OK;
}

void orc_pcg32_range(uint64_t seed, uint32_t n, const uint32_t* lo, const uint32_t* hi, uint32_t* out) {
    pcg32 g(seed);
    for (uint32_t i = 0; i < n; i++) out[i] = g.range(lo[i], hi[i]);
}

uint64_t orc_stratified_tables(uint64_t seed, uint64_t skip, uint32_t strata_x, uint32_t strata_y, uint32_t dims_1d,
                               uint32_t dims_2d, float* out_1d, float* out_2d) {
    pcg32 g(seed);
    for (uint64_t i = 0; i < skip; i++) g.next_u32();
    stratified_tables t;
    t.sx = strata_x;
    t.sy = strata_y;
    t.d1 = dims_1d;
    t.d2 = dims_2d;
    const pcg32 before = g;
    t.fill(g);
    // count the draws by replaying the generator up to its new state
    pcg32 c = before;
    uint64_t draws = 0;
    while (c.state != g.state) {
        c.next_u32();
        draws++;
    }
    for (size_t i = 0; i < t.t1.size(); i++) out_1d[i] = t.t1[i];
    for (size_t i = 0; i < t.t2.size(); i++) {
        out_2d[2 * i] = t.t2[i].x;
        out_2d[2 * i + 1] = t.t2[i].y;
    }
    return draws;
}

int orc_trace_rays(orc_scene* s, uint32_t n, const float* o, const float* d, const float* tmax, uint32_t vis, rbw_hit* out) {
    return trace_rays_impl(s, n, o, d, tmax, vis, out, false);
}

int orc_trace_rays_linear(orc_scene* s, uint32_t n, const float* o, const float* d, const float* tmax, uint32_t vis,
                          rbw_hit* out) {
    return trace_rays_impl(s, n, o, d, tmax, vis, out, true);
}

int orc_render(orc_scene* s, const rbw_camera_desc* cam, const rbw_render_params* p, orc_render_opts* opts,
               float* out_rgb, orc_stats* stats) {
    if (!s || !cam || !p || !opts) return RBW_ERR_INVALID_ARGUMENT;
    if (s->emitters.empty()) return RBW_ERR_INVALID_ARGUMENT;  // integrator.cpp:89,121 needs an emitter
    if (opts->sampler == ORC_SAMPLER_TILE_STRATIFIED &&
        (opts->strata_x == 0 || opts->strata_y == 0 || opts->strata_x * opts->strata_y != p->spp))
        return RBW_ERR_INVALID_ARGUMENT;  // the stratified sampler's spp is strata_x * strata_y
    render_ctx ctx;
    ctx.s = s;
    ctx.cam = cam;
    ctx.p = p;
    std::memcpy(ctx.raster_to_camera.m, cam->raster_to_camera, sizeof(float) * 16);
    std::memcpy(ctx.camera_to_world.m, cam->camera_to_world, sizeof(float) * 16);
    const int32_t W = static_cast<int32_t>(cam->width);
    const int32_t bx0 = cam->bound_min[0], by0 = cam->bound_min[1], bx1 = cam->bound_max[0], by1 = cam->bound_max[1];
    const uint32_t spp = p->spp;
    const int32_t tile = 16;  // sampler_integrator.cpp:30
    struct tile_in {
        uint32_t index;
        int32_t x0, y0, x1, y1;
    };
    std::vector<tile_in> tiles;
    for (int32_t y = by0; y < by1; y += tile)
        for (int32_t x = bx0; x < bx1; x += tile)
            tiles.push_back({static_cast<uint32_t>(tiles.size()), x, y, std::min(x + tile, bx1), std::min(y + tile, by1)});
    const uint32_t stride = opts->tile_stride > 1 ? opts->tile_stride : 1;
    std::vector<film_tile> ftiles(tiles.size());
    const bool record = opts->record_offsets != nullptr;
    std::vector<std::vector<float>> rec_vals(record ? tiles.size() : 0);
    std::vector<std::vector<uint32_t>> rec_counts(record ? tiles.size() : 0);
    std::vector<counters> tile_cnt(tiles.size());
    std::atomic<uint32_t> next_tile{0};
    std::atomic<uint64_t> samples{0};

    auto work = [&]() {
        for (;;) {
            const uint32_t ti = next_tile.fetch_add(1);
            if (ti >= tiles.size()) break;
            const tile_in& tin = tiles[ti];
            if (tin.index % stride != 0) continue;
            film_tile& ft = ftiles[ti];
            // film_tile ctor (film.cpp:33-55), box filter radius 0.5
            const double rad = 0.5;
            ft.fmin[0] = std::max(static_cast<int32_t>(std::ceil(tin.x0 - 0.5 - rad)), bx0);
            ft.fmin[1] = std::max(static_cast<int32_t>(std::ceil(tin.y0 - 0.5 - rad)), by0);
            ft.fmax[0] = std::min(static_cast<int32_t>(std::floor(tin.x1 - 0.5 + rad) + 1), bx1);
            ft.fmax[1] = std::min(static_cast<int32_t>(std::floor(tin.y1 - 0.5 + rad) + 1), by1);
            const int32_t fw = ft.fmax[0] - ft.fmin[0];
            ft.px.assign(static_cast<size_t>(fw) * (ft.fmax[1] - ft.fmin[1]), pixel_acc());
            sample_stream smp;
            const bool stratified = opts->sampler == ORC_SAMPLER_TILE_STRATIFIED;
            if (opts->sampler == ORC_SAMPLER_TILE || stratified) smp.gen = pcg32(static_cast<uint64_t>(tin.index));
            stratified_tables tables;
            if (stratified) {
                tables.sx = opts->strata_x;
                tables.sy = opts->strata_y;
                tables.d1 = opts->dims_1d;
                tables.d2 = opts->dims_2d;
                smp.strat = &tables;
            }
            if (record) smp.record = &rec_vals[ti];
            counters& cnt = tile_cnt[ti];
            for (int32_t y = tin.y0; y < tin.y1; y++) {
                for (int32_t x = tin.x0; x < tin.x1; x++) {
                    const uint64_t pixel = static_cast<uint64_t>(y) * W + x;
                    if (stratified) tables.fill(smp.gen);  // trace_samplers.reset()
                    for (uint32_t si = 0; si < spp; si++) {
                        smp.sample = si;  // next_sample(): dimension counters back to 0
                        smp.k1 = smp.k2 = 0;
                        const uint64_t sid = pixel * spp + si;
                        if (opts->sampler == ORC_SAMPLER_PER_SAMPLE)
                            smp.gen = pcg32(static_cast<uint64_t>(p->seed) * 0x9E3779B97F4A7C15ull + sid);
                        else if (opts->sampler == ORC_SAMPLER_REPLAY) {
                            smp.replay = p->replay_values;
                            smp.pos = p->replay_offsets[sid];
                            smp.end = p->replay_offsets[sid + 1];
                        }
                        const size_t before = record ? rec_vals[ti].size() : 0;
                        const vec2 j = smp.next2();
                        const vec2 pos(static_cast<real>(x) + j.x, static_cast<real>(y) + j.y);
                        const vec2 lens = smp.next2();
                        const ray r = camera_ray(ctx, pos, lens);
                        cnt.camera++;
                        uint32_t* dbg = p->debug_hits ? p->debug_hits + sid * 3 * p->max_depth : nullptr;
                        if (dbg)
                            for (uint32_t k = 0; k < 3 * p->max_depth; k++) dbg[k] = RBW_NONE;
                        const spectrum L = p->integrator == RBW_INTEGRATOR_DIRECT ? trace_direct(ctx, smp, r, cnt)
                                           : p->integrator == RBW_INTEGRATOR_VOLUME_PATH ? trace_volume(ctx, smp, r, cnt)
                                                                                         : trace_path(ctx, smp, r, cnt, dbg);
                        // film_tile::add_sample (film.cpp:57-92)
                        const vec2 dp(pos.x - 0.5f, pos.y - 0.5f);
                        int32_t sx0 = static_cast<int32_t>(std::ceil(dp.x - 0.5f));
                        int32_t sy0 = static_cast<int32_t>(std::ceil(dp.y - 0.5f));
                        int32_t sx1 = static_cast<int32_t>(std::floor(dp.x + 0.5f)) + 1;
                        int32_t sy1 = static_cast<int32_t>(std::floor(dp.y + 0.5f)) + 1;
                        sx0 = std::max(sx0, ft.fmin[0]); sy0 = std::max(sy0, ft.fmin[1]);
                        sx1 = std::min(sx1, ft.fmax[0]); sy1 = std::min(sy1, ft.fmax[1]);
                        for (int32_t yy = sy0; yy < sy1; yy++)
                            for (int32_t xx = sx0; xx < sx1; xx++) {
                                pixel_acc& pa = ft.px[static_cast<size_t>(yy - ft.fmin[1]) * fw + (xx - ft.fmin[0])];
                                pa.sum = pa.sum + L * 1.0f;
                                pa.weight = pa.weight + 1.0f;
                            }
                        if (record) rec_counts[ti].push_back(static_cast<uint32_t>(rec_vals[ti].size() - before));
                    }
                }
            }
            samples.fetch_add(static_cast<uint64_t>(tin.x1 - tin.x0) * (tin.y1 - tin.y0) * spp);
        }
    };

    uint32_t nthreads = opts->threads ? opts->threads : std::max(1u, std::thread::hardware_concurrency());
    const auto t0 = std::chrono::high_resolution_clock::now();
    std::vector<std::thread> pool;
    for (uint32_t i = 1; i < nthreads; i++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    // film::add_tile in tile order (sampler_integrator.cpp:137-138, film.cpp:207-223)
    const int32_t H = static_cast<int32_t>(cam->height);
    std::vector<pixel_acc> film(static_cast<size_t>(W) * H);
    for (size_t ti = 0; ti < tiles.size(); ti++) {
        const film_tile& ft = ftiles[ti];
        if (ft.px.empty()) continue;
        const int32_t fw = ft.fmax[0] - ft.fmin[0];
        for (int32_t y = ft.fmin[1]; y < ft.fmax[1]; y++)
            for (int32_t x = ft.fmin[0]; x < ft.fmax[0]; x++) {
                const pixel_acc& tp = ft.px[static_cast<size_t>(y - ft.fmin[1]) * fw + (x - ft.fmin[0])];
                pixel_acc& fp = film[static_cast<size_t>(y) * W + x];
                fp.sum = fp.sum + tp.sum;
                fp.weight = fp.weight + tp.weight;
            }
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (out_rgb) {
        for (int32_t y = 0; y < H; y++)
            for (int32_t x = 0; x < W; x++) {
                const pixel_acc& fp = film[static_cast<size_t>(y) * W + x];
                const size_t o = (static_cast<size_t>(y) * W + x) * 3;
                const bool in = x >= bx0 && x < bx1 && y >= by0 && y < by1 && fp.weight != 0;
                out_rgb[o + 0] = in ? fp.sum.x / fp.weight : 0.0f;
                out_rgb[o + 1] = in ? fp.sum.y / fp.weight : 0.0f;
                out_rgb[o + 2] = in ? fp.sum.z / fp.weight : 0.0f;
            }
    }
    if (record) {
        // offsets in (pixel*spp + s) order
        uint64_t total = 0;
        for (auto& v : rec_vals) total += v.size();
        opts->record_count = total;
        if (total > opts->record_capacity || !opts->record_values) return RBW_ERR_OUT_OF_MEMORY;
        const size_t nsamples = static_cast<size_t>(W) * H * spp;
        std::vector<uint32_t> cnt_of(nsamples, 0);
        std::vector<uint64_t> src_of(nsamples, 0);
        std::vector<uint32_t> tile_of(nsamples, 0);
        for (size_t ti = 0; ti < tiles.size(); ti++) {
            const tile_in& tin = tiles[ti];
            if (rec_counts[ti].empty()) continue;
            size_t k = 0;
            uint64_t src = 0;
            for (int32_t y = tin.y0; y < tin.y1; y++)
                for (int32_t x = tin.x0; x < tin.x1; x++)
                    for (uint32_t si = 0; si < spp; si++) {
                        const size_t sid = (static_cast<size_t>(y) * W + x) * spp + si;
                        cnt_of[sid] = rec_counts[ti][k++];
                        src_of[sid] = src;
                        tile_of[sid] = static_cast<uint32_t>(ti);
                        src += cnt_of[sid];
                    }
        }
        uint64_t off = 0;
        for (size_t sid = 0; sid < nsamples; sid++) {
            opts->record_offsets[sid] = off;
            if (cnt_of[sid])
                std::memcpy(opts->record_values + off, rec_vals[tile_of[sid]].data() + src_of[sid], sizeof(float) * cnt_of[sid]);
            off += cnt_of[sid];
        }
        opts->record_offsets[nsamples] = off;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->samples = samples.load();
        stats->threads = nthreads;
        for (auto& c : tile_cnt) {
            stats->camera_rays += c.camera;
            stats->extension_rays += c.extension;
            stats->shadow_rays += c.shadow;
            stats->mis_rays += c.mis;
            stats->light_casts += c.light;
        }
    }
    return RBW_OK;
}

// perspective_camera::sample(position, sample) alone: n raster positions (x, y) and lens
// samples (u, v) -> world rays (origin, normalised direction)
void orc_camera_rays(const rbw_camera_desc* cam, uint32_t n, const float* raster_xy, const float* lens_uv,
                     float* out_o, float* out_d) {
    render_ctx ctx{};
    ctx.cam = cam;
    std::memcpy(ctx.raster_to_camera.m, cam->raster_to_camera, sizeof(float) * 16);
    std::memcpy(ctx.camera_to_world.m, cam->camera_to_world, sizeof(float) * 16);
    for (uint32_t i = 0; i < n; i++) {
        const ray r = camera_ray(ctx, vec2(raster_xy[2 * i], raster_xy[2 * i + 1]),
                                 vec2(lens_uv[2 * i], lens_uv[2 * i + 1]));
        for (int c = 0; c < 3; c++) {
            out_o[3 * i + c] = r.origin[c];
            out_d[3 * i + c] = r.direction[c];
        }
    }
}

void orc_pcg32_reals(uint64_t seed, uint32_t n, float* out) {
    pcg32 g(seed);
    for (uint32_t i = 0; i < n; i++) out[i] = g.uniform_real();
}

void orc_pcg32_u32(uint64_t seed, uint32_t n, uint32_t* out) {
    pcg32 g(seed);
    for (uint32_t i = 0; i < n; i++) out[i] = g.next_u32();
}

void orc_concentric_sample_disk(float u0, float u1, float* o) {
    const vec2 r = concentric_sample_disk(vec2(u0, u1));
    o[0] = r.x; o[1] = r.y;
}

void orc_cosine_sample_hemisphere(float u0, float u1, float* o) {
    const vec3 r = cosine_sample_hemisphere(vec2(u0, u1));
    o[0] = r.x; o[1] = r.y; o[2] = r.z;
}

void orc_uniform_sample_triangle(float u0, float u1, float* o) {
    const vec2 r = uniform_sample_triangle(vec2(u0, u1));
    o[0] = r.x; o[1] = r.y;
}

static vec3 v3(const float* p) { return vec3(p[0], p[1], p[2]); }

float orc_ggx_D(float ax, float ay, const float* wh) { return ggx{ax, ay}.D(v3(wh)); }
float orc_ggx_lambda(float ax, float ay, const float* w) { return ggx{ax, ay}.lambda(v3(w)); }

void orc_ggx_sample_wh(float ax, float ay, const float* wo, float u0, float u1, float* wh) {
    const vec3 r = ggx{ax, ay}.sample(v3(wo), vec2(u0, u1));
    wh[0] = r.x; wh[1] = r.y; wh[2] = r.z;
}

static bsdf make_ggx(float ax, float ay, const float* eta, const float* k) {
    bsdf b;
    b.kind = LOBE_GGX;
    b.dist = ggx{ax, ay};
    // no conductor given: Fresnel term 1
    b.fresnel_one = (eta == nullptr);
    b.eta = eta ? v3(eta) : spectrum(1);
    b.k = k ? v3(k) : spectrum(0);
    return b;
}

void orc_ggx_sample_f(float ax, float ay, const float* eta, const float* k, const float* wo, float u0, float u1,
                      float* wi, float* f, float* pdf) {
    const bsdf b = make_ggx(ax, ay, eta, k);
    vec3 w;
    spectrum ff(0);
    real p = 0;
    if (!b.ggx_sample(v3(wo), vec2(u0, u1), w, ff, p)) { w = vec3(0); ff = spectrum(0); p = 0; }
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    f[0] = ff.x; f[1] = ff.y; f[2] = ff.z;
    *pdf = p;
}

float orc_ggx_pdf(float ax, float ay, const float* wo, const float* wi) {
    const bsdf b = make_ggx(ax, ay, nullptr, nullptr);
    return b.ggx_pdf(v3(wo), v3(wi));
}

void orc_ggx_eval(float ax, float ay, const float* eta, const float* k, const float* wo, const float* wi, float* f) {
    const bsdf b = make_ggx(ax, ay, eta, k);
    const spectrum r = b.ggx_eval(v3(wo), v3(wi));
    f[0] = r.x; f[1] = r.y; f[2] = r.z;
}

float orc_fresnel_dielectric(float cos_theta_i, float eta_i, float eta_o) {
    return fresnel_dielectric(cos_theta_i, eta_i, eta_o);
}

float orc_material_pdf(const rbw_material* m, const float* wo, const float* wi) {
    return material_bsdf(*m).pdf(v3(wo), v3(wi));
}

void orc_material_eval(const rbw_material* m, const float* wo, const float* wi, float* f) {
    const spectrum r = material_bsdf(*m).evaluate(v3(wo), v3(wi));
    f[0] = r.x; f[1] = r.y; f[2] = r.z;
}

int orc_material_sample(const rbw_material* m, const float* wo, float u0, float u1, float* wi, float* f, float* pdf) {
    vec3 w(0);
    spectrum ff(0);
    real p = 0;
    const bool ok = material_bsdf(*m).sample_local(v3(wo), vec2(u0, u1), w, ff, p);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    f[0] = ff.x; f[1] = ff.y; f[2] = ff.z;
    *pdf = ok ? p : 0;
    return ok ? 1 : 0;
}

float orc_hg_phase(float cos_theta, float g) { return hg_phase(cos_theta, g); }
float orc_flog(float x) { return flog(x); }
float orc_fexp(float x) { return fexp(x); }

float orc_hg_sample(const float* wo, float g, float u0, float u1, float* wi) {
    real pdf = 0;
    const vec3 w = hg_sample(v3(wo), g, vec2(u0, u1), pdf);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    return pdf;
}

// the emitter pdf of direction wi from a medium vertex at p (zero normal), as the reference
// computes it (raw = 0) or from the untransformed interaction (raw = 1)
float orc_medium_emitter_pdf(orc_scene* s, uint32_t entity, const float* p, const float* wi, int raw) {
    interaction it;
    it.point = v3(p);
    it.normal = vec3(0);
    it.wo = vec3(0, 0, 1);
    counters cnt;
    return emitter_pdf(*s, entity, it, v3(wi), cnt, raw != 0);
}

void orc_fresnel_conductor(float cos_theta_i, const float* eta, const float* k, float* out3) {
    const spectrum r = fresnel_conductor(std::fabs(cos_theta_i), v3(eta), v3(k));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

}  // extern "C"
