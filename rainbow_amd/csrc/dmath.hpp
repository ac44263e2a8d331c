// Float arithmetic of the render path, usable on host and device.
//
// Every helper fixes the evaluation order that rainbow-core gets from glm
// (rainbow-core/math/detail/vector.hpp:9-105):
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)   = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v) = v * (1 / sqrt(dot(v,v)))
// All HIP sources are compiled with -ffp-contract=off and IEEE division and
// square root, so these expressions round exactly as written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RBW_HD __host__ __device__ __forceinline__

namespace rbw {

constexpr float k_pi = 3.14159265358979323846f;
constexpr float k_two_pi = 6.28318530717958647692f;
constexpr float k_half_pi = 1.57079632679489661923f;
constexpr float k_quarter_pi = 0.785398163397448309616f;
constexpr float k_one_over_pi = 0.318309886183790671538f;
constexpr float k_flt_epsilon = 1.19209290e-07f;
// shared/ray.hpp:10-11
constexpr float k_ray_epsilon = k_flt_epsilon / 2 * 1500;
constexpr float k_shadow_epsilon = k_ray_epsilon * 10;
constexpr float k_one_minus_epsilon = 1.0f - k_flt_epsilon;
constexpr float k_flt_max = 3.402823466e+38f;
constexpr float k_flt_min = 1.175494351e-38f;

struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};

RBW_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RBW_HD f3 splat(float s) { return f3{s, s, s}; }
RBW_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RBW_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RBW_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
RBW_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RBW_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
RBW_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RBW_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
RBW_HD f3 div3(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }

RBW_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RBW_HD float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
RBW_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
RBW_HD float length(f3 v) { return sqrtf(dot(v, v)); }
RBW_HD f3 normalize(f3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
RBW_HD float length_squared(f3 v) {
    const f3 d = v - splat(0.0f);
    return dot(d, d);
}
RBW_HD float distance_squared(f3 a, f3 b) {
    const f3 d = a - b;
    return dot(d, d);
}
RBW_HD f3 abs3(f3 v) { return f3{fabsf(v.x), fabsf(v.y), fabsf(v.z)}; }
// glm::min / glm::max on scalars
RBW_HD float gmin(float x, float y) { return (y < x) ? y : x; }
RBW_HD float gmax(float x, float y) { return (x < y) ? y : x; }
RBW_HD float gclamp(float v, float lo, float hi) { return gmin(gmax(v, lo), hi); }
// rainbow-core/math/detail/vector.hpp:85-93
RBW_HD float max_component(f3 v) {
    float m = v.x;
    if (m < v.y) m = v.y;
    if (m < v.z) m = v.z;
    return m;
}
// coefficient_specturm.hpp:38-49 (math::max fold)
RBW_HD float spectrum_max(f3 s) { return gmax(gmax(s.x, s.y), s.z); }
RBW_HD bool is_black(f3 s) { return s.x == 0.0f && s.y == 0.0f && s.z == 0.0f; }
// transform_vector / transform_point by the identity matrix: glm's
// (m0*x + m1*y) + (m2*z + m3*w) reduces to x + 0 (signed zero becomes +0)
RBW_HD f3 ident(f3 v) { return f3{v.x + 0.0f, v.y + 0.0f, v.z + 0.0f}; }
// transform_point by a translation t: (x + 0*y) + (0*z + t) = x + t for t != 0
RBW_HD f3 translate(f3 p, f3 t) { return f3{p.x + t.x, p.y + t.y, p.z + t.z}; }

// Transcendentals: correctly rounded float results via the double routine.
RBW_HD float rsin(float x) { return (float)sin((double)x); }
RBW_HD float rcos(float x) { return (float)cos((double)x); }
RBW_HD float racos(float x) { return (float)acos((double)x); }
RBW_HD float ratan2(float y, float x) { return (float)atan2((double)y, (double)x); }
RBW_HD float rlog(float x) { return (float)log((double)x); }
RBW_HD float rexp(float x) { return (float)exp((double)x); }

// std::nextafter for float (IEEE 754 nextUp/nextDown on the bit pattern)
RBW_HD float next_after(float from, float to) {
    if (from != from || to != to) return from + to;
    if (from == to) return to;
    if (from == 0.0f) {
        const uint32_t tiny = 1u;
        float r;
        uint32_t b = tiny | (to < 0.0f ? 0x80000000u : 0u);
        __builtin_memcpy(&r, &b, 4);
        return r;
    }
    uint32_t b;
    __builtin_memcpy(&b, &from, 4);
    if ((from < to) == (from > 0.0f)) b += 1u;
    else b -= 1u;
    float r;
    __builtin_memcpy(&r, &b, 4);
    return r;
}

// PCG32 XSH-RR 64/32 (shared/random_generator.cpp:91-120)
struct pcg32 {
    uint64_t state, inc;
    RBW_HD void seed(uint64_t s) {
        state = 0u;
        inc = (s << 1u) | 1u;
        next_u32();
        state += 0x853c49e6748fea9bULL;
        next_u32();
    }
    RBW_HD uint32_t next_u32() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    RBW_HD float uniform() {
        const float x = (float)next_u32() * 2.3283064365386963e-10f;
        return (x < k_one_minus_epsilon) ? x : k_one_minus_epsilon;
    }
};

// stream of the camera sample `sid` (= pixel * spp + s) in RBW_SAMPLER_PER_SAMPLE_PCG mode
RBW_HD uint64_t sample_stream(uint32_t seed, uint64_t sid) { return (uint64_t)seed * 0x9E3779B97F4A7C15ull + sid; }

}  // namespace rbw
