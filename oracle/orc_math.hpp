// TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Scalar/vector arithmetic of rainbow-core restated with glm's operation
// order (rainbow-core/math/detail/vector.hpp:9-105 forwards to glm):
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)    = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v)  = v * (1 / sqrt(dot(v,v)))
//   mat4 * vec4   = (m0*v.x + m1*v.y) + (m2*v.z + m3*v.w)   (columns m0..m3)
// The file must be compiled with -ffp-contract=off: the reference is built
// by MSVC /fp:precise, which does not contract a*b+c into an FMA.
#pragma once

#include <cmath>
#include <cstdint>
#include <cfloat>
#include <limits>

namespace orc {

using real = float;

// glm constants rounded to float (glm/gtc/constants.inl)
constexpr real k_pi = static_cast<real>(3.14159265358979323846264338327950288);
constexpr real k_two_pi = static_cast<real>(6.28318530717958647692528676655900576);
constexpr real k_half_pi = static_cast<real>(1.57079632679489661923132169163975144);
constexpr real k_quarter_pi = static_cast<real>(0.785398163397448309615660845819875721);
constexpr real k_one_over_pi = static_cast<real>(0.318309886183790671537767526745028724);
// shared/ray.hpp:10-11
constexpr real k_ray_epsilon = std::numeric_limits<real>::epsilon() / 2 * 1500;
constexpr real k_shadow_epsilon = k_ray_epsilon * 10;
constexpr real k_one_minus_epsilon = static_cast<real>(1) - std::numeric_limits<real>::epsilon();

// Transcendentals: correctly rounded float results, obtained from the double
// routine.  The reference calls MSVC's float sin/cos/atan2/acos, whose exact
// bits are not reproducible here; the HIP kernels use the same double-rounded
// definition so both sides agree bit for bit.
inline real rsin(real x) { return static_cast<real>(std::sin(static_cast<double>(x))); }
inline real rcos(real x) { return static_cast<real>(std::cos(static_cast<double>(x))); }
inline real racos(real x) { return static_cast<real>(std::acos(static_cast<double>(x))); }
inline real ratan2(real y, real x) { return static_cast<real>(std::atan2(static_cast<double>(y), static_cast<double>(x))); }
// Float log and exp with one fixed sequence of IEEE operations (no contraction), so that
// the oracle, which carries the same definition, agrees with the kernels bit for bit at
// real cost.  Published methods: log by reduction to m in [sqrt(1/2), sqrt(2)) and the
// series log(m) = 2 atanh(s), s = (m - 1) / (m + 1), through s^9; exp by Cody-Waite
// reduction x = k ln2 + r and the Taylor polynomial of degree 7, scaled by 2^k.  Both
// are within 2 ulp of the correctly rounded value on the ranges used (tested).
inline real flog(real x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_huge_valf() : __builtin_nanf("");
    if (x == __builtin_huge_valf()) return x;
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    int e = 0;
    if (b < 0x00800000u) {  // subnormal: scale into the normal range
        const real y = x * 8388608.0f;
        __builtin_memcpy(&b, &y, 4);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    b = (b & 0x007FFFFFu) | 0x3F800000u;
    real m;
    __builtin_memcpy(&m, &b, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const real s = (m - 1.0f) / (m + 1.0f);
    const real s2 = s * s;
    const real p = s2 * (0.666666667f + s2 * (0.4f + s2 * (0.285714286f + s2 * 0.222222222f)));
    const real lm = (s + s) + s * p;
    const real fe = static_cast<real>(e);
    return fe * 0.693145752f + (lm + fe * 1.42860677e-6f);
}

inline real fexp(real x) {
    if (x != x) return x;
    if (x > 88.7228394f) return __builtin_huge_valf();
    if (x < -103.972084f) return 0.0f;
    const real kf = __builtin_rintf(x * 1.44269504f);
    const real r = (x - kf * 0.693145752f) - kf * 1.42860677e-6f;
    const real p =
        1.0f + r * (1.0f + r * (0.5f + r * (0.166666667f + r * (0.0416666667f + r * (0.00833333333f +
                                                                                    r * (0.00138888889f +
                                                                                         r * 0.000198412698f))))));
    int k = (int)kf;
    real scale = 1.0f;
    if (k < -125) {  // subnormal result: scale in two steps
        scale = 5.42101086e-20f;  // 2^-64
        k += 64;
    } else if (k > 127) {  // x in (88.38, 88.72]: 2^128 is not a float, so 2^127 * 2
        scale = 2.0f;
        k -= 1;
    }
    const uint32_t eb = (uint32_t)(k + 127) << 23;
    real two_k;
    __builtin_memcpy(&two_k, &eb, 4);
    return p * two_k * scale;
}

// glm::min/max on scalars: (y < x) ? y : x  and  (x < y) ? y : x
inline real fmin_g(real x, real y) { return (y < x) ? y : x; }
inline real fmax_g(real x, real y) { return (x < y) ? y : x; }
inline real clamp_g(real v, real lo, real hi) { return fmin_g(fmax_g(v, lo), hi); }

struct vec2 {
    real x = 0, y = 0;
    vec2() = default;
    vec2(real a, real b) : x(a), y(b) {}
};

struct vec3 {
    real x = 0, y = 0, z = 0;
    vec3() = default;
    explicit vec3(real s) : x(s), y(s), z(s) {}
    vec3(real a, real b, real c) : x(a), y(b), z(c) {}
    real operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    real& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline vec3 operator+(const vec3& a, const vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(const vec3& a, const vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(const vec3& a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(const vec3& a, real s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(real s, const vec3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator*(const vec3& a, const vec3& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator/(const vec3& a, real s) { return {a.x / s, a.y / s, a.z / s}; }

inline real dot(const vec3& a, const vec3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline real dot(const vec2& a, const vec2& b) { return a.x * b.x + a.y * b.y; }
inline vec3 cross(const vec3& a, const vec3& b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline real length(const vec3& v) { return std::sqrt(dot(v, v)); }
inline vec3 normalize(const vec3& v) { return v * (static_cast<real>(1) / std::sqrt(dot(v, v))); }
inline real length_squared(const vec3& v) { vec3 d = v - vec3(0); return dot(d, d); }
inline real distance_squared(const vec3& a, const vec3& b) { vec3 d = a - b; return dot(d, d); }
inline vec3 vabs(const vec3& v) { return {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)}; }
inline vec3 vmin(const vec3& a, const vec3& b) { return {fmin_g(a.x, b.x), fmin_g(a.y, b.y), fmin_g(a.z, b.z)}; }
inline vec3 vmax(const vec3& a, const vec3& b) { return {fmax_g(a.x, b.x), fmax_g(a.y, b.y), fmax_g(a.z, b.z)}; }
// rainbow-core/math/detail/vector.hpp:85-93
inline real max_component(const vec3& v) {
    real m = v.x;
    if (m < v.y) m = v.y;
    if (m < v.z) m = v.z;
    return m;
}

// Column-major 4x4 matrix with glm's mat*vec evaluation order.
struct mat4 {
    real m[16];  // m[col*4 + row]
    static mat4 identity() {
        mat4 r{};
        for (int i = 0; i < 16; i++) r.m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        return r;
    }
    static mat4 translation(const vec3& t) {  // glm::translate(identity, t)
        mat4 r = identity();
        r.m[12] = t.x; r.m[13] = t.y; r.m[14] = t.z;
        return r;
    }
    mat4 transpose() const {
        mat4 r{};
        for (int c = 0; c < 4; c++)
            for (int w = 0; w < 4; w++) r.m[c * 4 + w] = m[w * 4 + c];
        return r;
    }
    // returns xyz of M * (x, y, z, w)
    vec3 mul(const vec3& v, real w) const {
        vec3 r;
        for (int row = 0; row < 3; row++) {
            real a0 = m[0 * 4 + row] * v.x;
            real a1 = m[1 * 4 + row] * v.y;
            real a2 = m[2 * 4 + row] * v.z;
            real a3 = m[3 * 4 + row] * w;
            r[row] = (a0 + a1) + (a2 + a3);
        }
        return r;
    }
};

// An affine rainbow transform: matrix plus its inverse (shared/transform.hpp:16-48)
struct xform {
    mat4 M, Minv;
    static xform identity() { return {mat4::identity(), mat4::identity()}; }
    static xform translate(const vec3& t) { return {mat4::translation(t), mat4::translation(-t)}; }
    xform inverse() const { return {Minv, M}; }
};

// shared/transform.cpp:172-197
inline vec3 transform_point(const xform& t, const vec3& p) { return t.M.mul(p, 1); }
inline vec3 transform_vector(const xform& t, const vec3& v) { return t.M.mul(v, 0); }
inline vec3 transform_normal(const xform& t, const vec3& n) { return normalize(t.Minv.transpose().mul(n, 0)); }

// shared/random_generator.cpp:91-120 -- PCG32 XSH-RR 64/32 with the
// reference's seeding and float conversion.
struct pcg32 {
    static constexpr uint64_t k_default_state = 0x853c49e6748fea9bULL;
    static constexpr uint64_t k_mult = 0x5851f42d4c957f2dULL;
    uint64_t state = 0, inc = 1;
    pcg32() = default;
    explicit pcg32(uint64_t seed) {
        state = 0u;
        inc = (seed << 1u) | 1u;
        next_u32();
        state += k_default_state;
        next_u32();
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * k_mult + inc;
        uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    real uniform_real() {
        // uint32 * float: the integer is rounded to float first; std::min(a, b) = (b < a) ? b : a
        const real x = static_cast<real>(next_u32()) * 2.3283064365386963e-10f;
        return (x < k_one_minus_epsilon) ? x : k_one_minus_epsilon;
    }
    // pcg32::uint32(value), random_generator.cpp:78-85: rejection below (2^32 - value) % value
    uint32_t bounded(uint32_t value) {
        const uint32_t threshold = (~value + 1u) % value;
        for (;;) {
            const uint32_t r = next_u32();
            if (r >= threshold) return r % value;
        }
    }
    // pcg32::uint32(min, max), random_generator.cpp:73-76
    uint32_t range(uint32_t lo, uint32_t hi) { return bounded(hi - lo + 1u) + lo; }
};

}  // namespace orc
