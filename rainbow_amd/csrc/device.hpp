// Device-side scene, path state and the per-vertex math of the wavefront
// tracer.  Each function cites the reference routine whose arithmetic it
// reproduces; the operation order matches so that CPU and GPU agree bit for
// bit wherever the inputs agree.
#pragma once

#include "dmath.hpp"
#include "../../include/rbw.h"

namespace rbw {

// primitive id word of a BVH leaf slot (prim_a[slot].w)
constexpr uint32_t k_prim_sphere = 0x80000000u;     // sphere (else triangle)
constexpr uint32_t k_prim_invisible = 0x40000000u;  // entity without material (skipped by shadow rays)
constexpr uint32_t k_prim_instanced = 0x20000000u;  // triangle of an entity with more than 2^17 triangles
constexpr uint32_t k_prim_id_mask = 0x1FFFFFFFu;

struct dentity {
    uint32_t shape, first, count, material, flags;
    float radiance[3];
    float area;       // shape_instance_properties::area (shapes/shape.cpp:51-54)
    float center[3];  // sphere translation
    float radius;
    uint32_t pad;
};

struct dmaterial {
    uint32_t type;
    float diffuse[3];
    float eta[3];
    float k[3];
    float alpha_u, alpha_v;
    float specular[3];
    float ior;
};

// A participating medium (media/homogeneous_medium.cpp, media/heterogeneous_medium.cpp)
struct dmedium {
    uint32_t type;               // RBW_MEDIUM_*
    float sigma_t[3];            // homogeneous: sigma_a + sigma_s
    float sigma_s[3];
    float g;
    float albedo[3];             // grid
    float inv_max_density;       // grid: 1 / the largest density
    uint32_t grid_offset, nx, ny, nz;
    float w2l[12];               // grid: world -> unit cube, 3x4 row-major
};

struct dscene {
    const float4* nodes;     // bvh2 nodes, 4 float4 each
    const float4* prim_a;    // leaf slots: (p0 | centre, id word)
    const float4* prim_b;    // (p1 | radius)
    const float4* prim_c;    // p2
    const float* pos;        // 3 per vertex
    const float* nrm;        // 3 per vertex (valid where the entity has normals)
    const float* uv;         // 2 per vertex or null
    const uint32_t* idx;     // 3 per triangle
    const uint32_t* tri_entity;
    const uint32_t* tri_material;  // null = entity material
    const float4* spheres;   // centre, radius
    const uint32_t* sphere_entity;
    const dentity* ents;
    const dmaterial* mats;
    const uint32_t* emitters;
    uint32_t num_emitters;
    uint32_t num_tris;
    uint32_t num_spheres;
    uint32_t root;           // BVH root node (a part's root inside a merged scene)
    // four-wide BVH over the same primitives and leaf slots (8 float4 per node)
    const float4* nodes4;
    uint32_t root4;
    // eight-wide BVH over the same primitives, with its own leaf slots
    const float4* nodes8;    // 5 float4 per node
    const float4* prim8_a;
    const float4* prim8_b;
    const float4* prim8_c;
    uint32_t root8;
    uint32_t width;          // 2, 4 or 8: the tree traversals use
    uint32_t light_flags;    // k_lights_* : what the connect stage may shortcut (set per scene)
    // media (volume path integrator only)
    const dmedium* media;
    const uint32_t* ent_media;  // 2 per entity: outside, inside (RBW_NONE = vacuum); null = no media
    const float* grid;          // grid media densities
    uint32_t num_media;
};

// light_flags: the emitters' triangles are few enough to test an MIS ray against them
// first; every emitter mesh is planar (a shadow ray to a point on it cannot reach it
// earlier, so "closest visible hit is the light or none" is "no other visible hit")
constexpr uint32_t k_lights_few = 1u, k_lights_planar = 2u;
constexpr uint32_t k_light_first_max_tris = 64;

struct hitrec {
    uint32_t prim;  // global primitive id or RBW_MISS
    float t;        // entity-local parameter
    float b1, b2;
};

// ------------------------------------------------------------------ traversal

// Moller-Trumbore with the reference's acceptance tests (shapes/mesh.cpp:333-352)
RBW_HD bool mt_test(f3 v0, f3 v1, f3 v2, f3 o, f3 d, float len, float& t, float& b1, float& b2) {
    const f3 e1 = v1 - v0;
    const f3 e2 = v2 - v0;
    const f3 p = cross(d, e2);
    const f3 tv = o - v0;
    const f3 q = cross(tv, e1);
    const float inv_det = 1.0f / dot(e1, p);
    b1 = dot(tv, p) * inv_det;
    b2 = dot(d, q) * inv_det;
    const float b0 = 1.0f - b1 - b2;
    t = dot(e2, q) * inv_det;
    if (t <= 0.0f || t >= len) return false;
    if (b0 < 0.0f || b0 > 1.0f) return false;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    if (b2 < 0.0f || b2 > 1.0f) return false;
    return true;
}

// Watertight ray/triangle test (Woop, Benthin and Wald, JCGT 2013): shear the
// triangle into ray space; edge functions decide the hit with no gaps between
// triangles sharing an edge.  Returns t in the same parameterisation as MT.
RBW_HD bool wt_test(f3 p0, f3 p1, f3 p2, f3 o, f3 d, float len, int kz, float Sx, float Sy, float Sz, float& t,
                    float& b1, float& b2) {
    const int kx = kz == 2 ? 0 : kz + 1;
    const int ky = kx == 2 ? 0 : kx + 1;
    const f3 A = p0 - o, B = p1 - o, Cc = p2 - o;
    auto comp = [](f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); };
    const float Az = comp(A, kz), Bz = comp(B, kz), Cz = comp(Cc, kz);
    const float Ax = comp(A, kx) - Sx * Az, Ay = comp(A, ky) - Sy * Az;
    const float Bx = comp(B, kx) - Sx * Bz, By = comp(B, ky) - Sy * Bz;
    const float Cx = comp(Cc, kx) - Sx * Cz, Cy = comp(Cc, ky) - Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        const double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        const double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (Sz * Az) + V * (Sz * Bz) + W * (Sz * Cz);
    const float inv = 1.0f / det;
    t = T * inv;
    if (!(t > 0.0f) || !(t < len)) return false;
    b1 = V * inv;
    b2 = W * inv;
    return true;
}

// rainbow-core/utilities.cpp:7-31 in double, then shapes/sphere.cpp:22-38
RBW_HD bool sphere_test(float radius, f3 o, f3 d, float len, float& t_hit) {
    const float a = dot(d, d);
    const float b = dot(d, o) * 2.0f;
    const float c = dot(o, o) - radius * radius;
    const double disc = (double)b * (double)b - (double)a * (double)c * 4.0;
    if (disc < 0.0) return false;
    const double root = sqrt(disc);
    const double q = (b < 0.0f) ? -0.5 * ((double)b - root) : -0.5 * ((double)b + root);
    float t0 = (float)(q / (double)a);
    float t1 = (float)((double)c / q);
    if (t0 > t1) {
        const float s = t0;
        t0 = t1;
        t1 = s;
    }
    if (t0 > len || t1 <= 0.0f) return false;
    if (t0 <= 0.0f) {
        if (t1 > len) return false;
        t_hit = t1;
    } else
        t_hit = t0;
    return true;
}

struct trace_counters {
    uint32_t nodes, prims;
};

// The ray as every entity's local space sees it (entity.cpp:18-58 with the
// identity or translation transforms of flattened scenes):
//   d_l = normalize(normalize(d_w + 0)),  local length = world length * |d_w + 0|,
//   and an accepted hit at local t sets the world length to t * |d_l + 0|.
struct local_ray {
    f3 o;        // world origin (+0)
    f3 d;        // local direction
    f3 inv;      // 1 / d
    float L0;    // |transform_vector(to_local, d_w)|
    float L1;    // |transform_vector(to_world, d_l)|
};

RBW_HD local_ray make_local_ray(f3 o_w, f3 d_w) {
    local_ray r;
    r.o = ident(o_w);
    const f3 dv = ident(d_w);
    r.d = normalize(normalize(dv));
    r.L0 = length(dv);
    r.L1 = length(ident(r.d));
    r.inv = f3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    return r;
}

struct wt_setup {
    int kz;
    float Sx, Sy, Sz;
};

RBW_HD wt_setup make_wt(f3 d) {
    wt_setup w;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    w.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    const int kx = w.kz == 2 ? 0 : w.kz + 1;
    const int ky = kx == 2 ? 0 : kx + 1;
    const float dz = w.kz == 0 ? d.x : (w.kz == 1 ? d.y : d.z);
    const float dx = kx == 0 ? d.x : (kx == 1 ? d.y : d.z);
    const float dy = ky == 0 ? d.x : (ky == 1 ? d.y : d.z);
    w.Sx = dx / dz;
    w.Sy = dy / dz;
    w.Sz = 1.0f / dz;
    return w;
}

}  // namespace rbw
