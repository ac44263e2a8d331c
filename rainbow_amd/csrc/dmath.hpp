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
// Float log and exp with one fixed sequence of IEEE operations (no contraction), so that
// the oracle, which carries the same definition, agrees with the kernels bit for bit at
// float cost.  Published methods: log by reduction to m in [sqrt(1/2), sqrt(2)) and the
// series log(m) = 2 atanh(s), s = (m - 1) / (m + 1), through s^9; exp by Cody-Waite
// reduction x = k ln2 + r and the Taylor polynomial of degree 7, scaled by 2^k.  Both
// are within 2 ulp of the correctly rounded value on the ranges used (tested).
RBW_HD float flog(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_huge_valf() : __builtin_nanf("");
    if (x == __builtin_huge_valf()) return x;
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    int e = 0;
    if (b < 0x00800000u) {  // subnormal: scale into the normal range
        const float y = x * 8388608.0f;
        __builtin_memcpy(&b, &y, 4);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    b = (b & 0x007FFFFFu) | 0x3F800000u;
    float m;
    __builtin_memcpy(&m, &b, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float p = s2 * (0.666666667f + s2 * (0.4f + s2 * (0.285714286f + s2 * 0.222222222f)));
    const float lm = (s + s) + s * p;
    const float fe = (float)e;
    return fe * 0.693145752f + (lm + fe * 1.42860677e-6f);
}

RBW_HD float fexp(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return __builtin_huge_valf();
    if (x < -103.972084f) return 0.0f;
    const float kf = __builtin_rintf(x * 1.44269504f);
    const float r = (x - kf * 0.693145752f) - kf * 1.42860677e-6f;
    const float p =
        1.0f + r * (1.0f + r * (0.5f + r * (0.166666667f + r * (0.0416666667f + r * (0.00833333333f +
                                                                                    r * (0.00138888889f +
                                                                                         r * 0.000198412698f))))));
    int k = (int)kf;
    float scale = 1.0f;
    if (k < -125) {  // subnormal result: scale in two steps
        scale = 5.42101086e-20f;  // 2^-64
        k += 64;
    } else if (k > 127) {  // x in (88.38, 88.72]: 2^128 is not a float, so 2^127 * 2
        scale = 2.0f;
        k -= 1;
    }
    const uint32_t eb = (uint32_t)(k + 127) << 23;
    float two_k;
    __builtin_memcpy(&two_k, &eb, 4);
    return p * two_k * scale;
}

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
