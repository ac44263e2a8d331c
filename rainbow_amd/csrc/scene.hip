// Scene flattening for the tracer: validation of an rbw_scene_desc, the SAH
// BVH over every primitive, the leaf-slot arrays the traversal reads, and the
// merge of several flattened scenes into one (the batched scheduler's scene
// pool).  The flattening follows scene::add / scene::build
// (scenes/scene.cpp:15-80) with the BVH replaced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "bvh.hpp"
#include "host.hpp"
#include "shading.hpp"

namespace rbw {
namespace host {

namespace {

inline float expand_lo(float v) { return v - (fabsf(v) * 4.8e-7f + 1e-30f); }
inline float expand_hi(float v) { return v + (fabsf(v) * 4.8e-7f + 1e-30f); }

inline float4 with_word(float x, float y, float z, uint32_t word) {
    float wf;
    std::memcpy(&wf, &word, 4);
    return make_float4(x, y, z, wf);
}

inline uint32_t word_of(const float4& v) {
    uint32_t w;
    std::memcpy(&w, &v.w, 4);
    return w;
}

// What the connect stage may shortcut for this scene's emitters (k_lights_*): few visible
// emitter triangles (an MIS ray is tested against them before any walk), and planar
// emitter meshes (every vertex within 1e-6 of the first triangle's plane, relative to the
// mesh's extent).
uint32_t light_flags_of(const rbw_scene_desc* d, const host_scene& h) {
    uint64_t tris = 0;
    bool planar = true;
    for (uint32_t e : h.emitters) {
        const rbw_entity& in = d->entities[e];
        if (in.material != RBW_NONE) tris += in.count;
        const uint32_t* I = d->indices + size_t(in.first) * 3;
        auto P = [&](uint32_t v) { return &d->positions[size_t(v) * 3]; };
        const float* a = P(I[0]);
        const float* b = P(I[1]);
        const float* c = P(I[2]);
        const double u[3] = {double(b[0]) - a[0], double(b[1]) - a[1], double(b[2]) - a[2]};
        const double v[3] = {double(c[0]) - a[0], double(c[1]) - a[1], double(c[2]) - a[2]};
        double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (!(nl > 0)) {
            planar = false;
            continue;
        }
        double ext = 0, dev = 0;
        for (uint32_t k = 0; k < 3 * in.count; k++) {
            const float* p = P(I[k]);
            const double w[3] = {double(p[0]) - a[0], double(p[1]) - a[1], double(p[2]) - a[2]};
            ext = std::max(ext, std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));
            dev = std::max(dev, std::fabs(w[0] * n[0] + w[1] * n[1] + w[2] * n[2]) / nl);
        }
        if (dev > 1e-6 * ext) planar = false;
    }
    return (tris <= k_light_first_max_tris ? k_lights_few : 0u) | (planar ? k_lights_planar : 0u);
}

// The media table (media/medium.hpp:37-75): homogeneous media keep sigma_t = sigma_a + sigma_s
// (homogeneous_medium.cpp:7-10); a grid keeps 1 / its largest density
// (heterogeneous_medium.cpp:12-22).  Entity media are (outside, inside) pairs.
void build_media(const rbw_scene_desc* d, host_scene& h) {
    h.media.clear();
    h.ent_media.clear();
    h.grid.clear();
    if (!d->entity_media) {
        require(d->num_media == 0 || d->media, RBW_ERR_INVALID_ARGUMENT, "media table missing");
        for (uint32_t e = 0; e < d->num_entities; e++)
            require(!(d->entities[e].flags & RBW_ENTITY_HAS_MEDIA), RBW_ERR_INVALID_ARGUMENT,
                    "an entity has media but the scene has no entity_media");
        return;
    }
    require(d->num_media > 0 && d->media, RBW_ERR_INVALID_ARGUMENT, "entity media given without a media table");
    require(d->medium_grid_count == 0 || d->medium_grid, RBW_ERR_INVALID_ARGUMENT, "medium grid missing");
    h.grid.assign(d->medium_grid, d->medium_grid + d->medium_grid_count);
    for (float v : h.grid)
        require(std::isfinite(v) && v >= 0.0f, RBW_ERR_INVALID_ARGUMENT, "grid densities must be finite and >= 0");
    h.media.resize(d->num_media);
    for (uint32_t i = 0; i < d->num_media; i++) {
        const rbw_medium& m = d->media[i];
        dmedium& o = h.media[i];
        std::memset(&o, 0, sizeof(o));
        o.type = m.type;
        o.g = m.g;
        require(std::fabs(m.g) < 1.0f, RBW_ERR_INVALID_ARGUMENT, "Henyey-Greenstein g must lie in (-1, 1)");
        if (m.type == RBW_MEDIUM_HOMOGENEOUS) {
            for (int c = 0; c < 3; c++) {
                require(m.sigma_a[c] >= 0.0f && m.sigma_s[c] >= 0.0f && std::isfinite(m.sigma_a[c] + m.sigma_s[c]),
                        RBW_ERR_INVALID_ARGUMENT, "medium coefficients must be finite and >= 0");
                o.sigma_t[c] = m.sigma_a[c] + m.sigma_s[c];
                o.sigma_s[c] = m.sigma_s[c];
            }
        } else if (m.type == RBW_MEDIUM_GRID) {
            const uint64_t n = uint64_t(m.grid_size[0]) * m.grid_size[1] * m.grid_size[2];
            require(n > 0 && uint64_t(m.grid_offset) + n <= d->medium_grid_count, RBW_ERR_INVALID_ARGUMENT,
                    "medium grid out of range");
            float mx = 0.0f;
            for (uint64_t k = 0; k < n; k++) mx = std::max(mx, d->medium_grid[m.grid_offset + k]);
            require(mx > 0.0f, RBW_ERR_INVALID_ARGUMENT, "medium grid has no density");
            o.inv_max_density = 1.0f / mx;
            o.grid_offset = m.grid_offset;
            o.nx = m.grid_size[0];
            o.ny = m.grid_size[1];
            o.nz = m.grid_size[2];
            for (int c = 0; c < 3; c++) o.albedo[c] = m.albedo[c];
            for (int k = 0; k < 12; k++) o.w2l[k] = m.world_to_local[k];
        } else {
            require(false, RBW_ERR_INVALID_ARGUMENT, "unknown medium type");
        }
    }
    h.ent_media.assign(d->entity_media, d->entity_media + size_t(d->num_entities) * 2);
    for (uint32_t v : h.ent_media)
        require(v == RBW_NONE || v < d->num_media, RBW_ERR_INVALID_ARGUMENT, "entity medium index out of range");
    // only entities with the media component switch the path's medium
    for (uint32_t e = 0; e < d->num_entities; e++)
        if (!(d->entities[e].flags & RBW_ENTITY_HAS_MEDIA)) h.ent_media[2 * e] = h.ent_media[2 * e + 1] = RBW_NONE;
}

// leaf slots in BVH order: (p0 | sphere centre, id word), (p1 | radius), p2
void fill_slots(const rbw_scene_desc* d, const host_scene& h, const std::vector<uint32_t>& order,
                std::vector<float4>& pa, std::vector<float4>& pb, std::vector<float4>& pc) {
    const uint32_t nt = d->num_triangles;
    const size_t nslots = order.size();
    pa.resize(nslots);
    pb.resize(nslots);
    pc.resize(nslots);
    for (size_t sl = 0; sl < nslots; sl++) {
        const uint32_t prim = order[sl];
        uint32_t word = prim;
        if (prim < nt) {
            const dentity& E = h.ents[d->triangle_entity[prim]];
            if (E.material == RBW_NONE) word |= k_prim_invisible;
            if (E.count > k_instanced_threshold) word |= k_prim_instanced;
            const float* p0 = d->positions + 3 * size_t(d->indices[3 * prim]);
            const float* p1 = d->positions + 3 * size_t(d->indices[3 * prim + 1]);
            const float* p2 = d->positions + 3 * size_t(d->indices[3 * prim + 2]);
            pa[sl] = with_word(p0[0], p0[1], p0[2], word);
            pb[sl] = make_float4(p1[0], p1[1], p1[2], 0.0f);
            pc[sl] = make_float4(p2[0], p2[1], p2[2], 0.0f);
        } else {
            const uint32_t s = prim - nt;
            const dentity& E = h.ents[d->sphere_entity[s]];
            word |= k_prim_sphere;
            if (E.material == RBW_NONE) word |= k_prim_invisible;
            pa[sl] = with_word(d->spheres[4 * s], d->spheres[4 * s + 1], d->spheres[4 * s + 2], word);
            pb[sl] = make_float4(d->spheres[4 * s + 3], 0.0f, 0.0f, 0.0f);
            pc[sl] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

}  // namespace

uint32_t resolve_width(uint32_t requested) {
    if (requested == 0) return k_default_width;
    require(requested == 2 || requested == 4 || requested == 8, RBW_ERR_INVALID_ARGUMENT,
            "bvh_width must be 0, 2, 4 or 8");
    return requested;
}

size_t host_scene::bytes() const {
    return (nodes.size() + nodes4.size() + nodes8.size()) * sizeof(float4) + 3 * (pa.size() + pa8.size()) * sizeof(float4) +
           (pos.size() + nrm.size() + uv.size()) * 4 +
           (idx.size() + tri_entity.size() + tri_material.size() + sphere_entity.size() + emitters.size()) * 4 +
           spheres.size() * sizeof(float4) + ents.size() * sizeof(dentity) + mats.size() * sizeof(dmaterial);
}

void build_host_scene(const rbw_scene_desc* d, host_scene& h) {
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
        require(m.type == RBW_MATERIAL_MATTE || m.type == RBW_MATERIAL_METAL || m.type == RBW_MATERIAL_PLASTIC,
                RBW_ERR_INVALID_ARGUMENT, "unknown material type");
        require(m.type != RBW_MATERIAL_PLASTIC || (m.alpha_u > 0.0f && m.alpha_v > 0.0f && m.ior > 0.0f),
                RBW_ERR_INVALID_ARGUMENT, "plastic alpha and ior must be positive");
        require(m.type != RBW_MATERIAL_MATTE || m.sigma == 0.0f, RBW_ERR_UNSUPPORTED,
                "matte sigma != 0 (Oren-Nayar) is not supported");
        require(m.type != RBW_MATERIAL_METAL || (m.alpha_u > 0.0f && m.alpha_v > 0.0f), RBW_ERR_INVALID_ARGUMENT,
                "metal alpha must be positive");
    }
    const uint32_t nt = d->num_triangles, ns = d->num_spheres, nv = d->num_vertices;
    for (uint64_t i = 0; i < 3ull * nt; i++)
        require(d->indices[i] < nv, RBW_ERR_INVALID_ARGUMENT, "triangle index out of range");
    h = host_scene();
    h.ents.resize(d->num_entities);
    for (uint32_t e = 0; e < d->num_entities; e++) {
        const rbw_entity& in = d->entities[e];
        dentity& o = h.ents[e];
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
            h.emitters.push_back(e);
        }
    }
    // integrator.cpp:89,121 dereference the sampled emitter: a scene needs one
    require(!h.emitters.empty(), RBW_ERR_INVALID_ARGUMENT, "scene has no emitter");
    h.light_flags = light_flags_of(d, h);
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
    h.smin = smin;
    h.smax = smax;
    bvh_build_result bvh;
    build_bvh(boxes, bvh, 4);
    require(bvh.depth < k_stack_depth, RBW_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    bvh4_result bvh4;
    build_bvh4(bvh, bvh4);
    require(bvh4.stack < k_stack_depth, RBW_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    bvh8_result bvh8;
    build_bvh8(bvh, bvh8);
    require(bvh8.depth < k_stack_depth, RBW_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    fill_slots(d, h, bvh.order, h.pa, h.pb, h.pc);
    fill_slots(d, h, bvh8.order, h.pa8, h.pb8, h.pc8);
    h.nodes.resize(bvh.nodes2.size() * 4);
    std::memcpy(h.nodes.data(), bvh.nodes2.data(), bvh.nodes2.size() * sizeof(bvh2_node));
    h.nodes4.resize(bvh4.nodes.size() * 8);
    std::memcpy(h.nodes4.data(), bvh4.nodes.data(), bvh4.nodes.size() * sizeof(bvh4_node));
    h.stack4 = bvh4.stack;
    h.nodes8.resize(bvh8.nodes.size() * 5);
    std::memcpy(h.nodes8.data(), bvh8.nodes.data(), bvh8.nodes.size() * sizeof(bvh8_node));
    h.depth8 = bvh8.depth;
    h.mats.resize(d->num_materials);
    for (uint32_t i = 0; i < d->num_materials; i++) {
        const rbw_material& m = d->materials[i];
        dmaterial& o = h.mats[i];
        o.type = m.type;
        for (int c = 0; c < 3; c++) {
            o.diffuse[c] = m.diffuse[c];
            o.eta[c] = m.eta[c];
            o.k[c] = m.k[c];
            o.specular[c] = m.specular[c];
        }
        o.alpha_u = m.alpha_u;
        o.alpha_v = m.alpha_v;
        o.ior = m.ior;
    }
    build_media(d, h);
    h.nt = nt;
    h.ns = ns;
    h.nv = nv;
    h.pos.assign(d->positions, d->positions + size_t(nv) * 3);
    h.has_nrm = d->normals != nullptr;
    if (h.has_nrm) h.nrm.assign(d->normals, d->normals + size_t(nv) * 3);
    h.has_uv = d->uvs != nullptr;
    if (h.has_uv) h.uv.assign(d->uvs, d->uvs + size_t(nv) * 2);
    h.idx.assign(d->indices, d->indices + size_t(nt) * 3);
    h.tri_entity.assign(d->triangle_entity, d->triangle_entity + nt);
    h.has_tri_material = d->triangle_material != nullptr;
    if (h.has_tri_material) h.tri_material.assign(d->triangle_material, d->triangle_material + nt);
    h.spheres.resize(ns);
    for (uint32_t i = 0; i < ns; i++)
        h.spheres[i] = make_float4(d->spheres[4 * i], d->spheres[4 * i + 1], d->spheres[4 * i + 2], d->spheres[4 * i + 3]);
    if (ns) h.sphere_entity.assign(d->sphere_entity, d->sphere_entity + ns);
    h.depth = bvh.depth;
}

void merge_scenes(const std::vector<const host_scene*>& parts, host_scene& u, std::vector<uint32_t>& roots,
                  std::vector<uint32_t>& roots4, std::vector<uint32_t>& roots8, std::vector<uint32_t>& em_first,
                  std::vector<uint32_t>& em_count) {
    u = host_scene();
    u.light_flags = k_lights_few | k_lights_planar;
    for (const host_scene* p : parts) u.light_flags &= p->light_flags;
    uint64_t T = 0, Sp = 0, V = 0, N = 0, L = 0, E = 0, M = 0, N4 = 0, N8 = 0, L8 = 0;
    bool any_nrm = false, any_uv = false, any_tm = false;
    for (const host_scene* p : parts) {
        T += p->nt;
        Sp += p->ns;
        V += p->nv;
        N += p->nodes.size() / 4;
        L += p->pa.size();
        N4 += p->nodes4.size() / 8;
        N8 += p->nodes8.size() / 5;
        L8 += p->pa8.size();
        E += p->ents.size();
        M += p->mats.size();
        any_nrm |= p->has_nrm;
        any_uv |= p->has_uv;
        any_tm |= p->has_tri_material;
    }
    require(T + Sp < (1ull << 29), RBW_ERR_UNSUPPORTED, "merged scene has too many primitives");
    require(L < (1ull << 27), RBW_ERR_UNSUPPORTED, "merged scene has too many leaf slots");
    require(N < (1ull << 31) && N4 < (1ull << 31) && N8 < (1ull << 32) && L8 < (1ull << 32) && V < (1ull << 32) && E < (1ull << 31) &&
                M < (1ull << 31),
            RBW_ERR_UNSUPPORTED, "merged scene too large");
    u.nt = uint32_t(T);
    u.ns = uint32_t(Sp);
    u.nv = uint32_t(V);
    u.has_nrm = any_nrm;
    u.has_uv = any_uv;
    u.has_tri_material = any_tm;
    u.nodes.reserve(N * 4);
    u.pa.reserve(L);
    u.pb.reserve(L);
    u.pc.reserve(L);
    u.pos.reserve(V * 3);
    if (any_nrm) u.nrm.reserve(V * 3);
    if (any_uv) u.uv.reserve(V * 2);
    u.idx.reserve(T * 3);
    u.tri_entity.reserve(T);
    if (any_tm) u.tri_material.reserve(T);
    u.spheres.reserve(Sp);
    u.sphere_entity.reserve(Sp);
    u.ents.reserve(E);
    u.mats.reserve(M);
    u.nodes4.reserve(N4 * 8);
    u.nodes8.reserve(N8 * 5);
    u.pa8.reserve(L8);
    u.pb8.reserve(L8);
    u.pc8.reserve(L8);
    roots.clear();
    roots4.clear();
    roots8.clear();
    em_first.clear();
    em_count.clear();
    uint32_t tb = 0, sb = 0, vb = 0, nb = 0, lb = 0, eb = 0, mb = 0, n4b = 0, n8b = 0, l8b = 0;
    u.smin = f3{3.4e38f, 3.4e38f, 3.4e38f};
    u.smax = f3{-3.4e38f, -3.4e38f, -3.4e38f};
    for (const host_scene* p : parts) {
        roots.push_back(nb);
        roots4.push_back(n4b);
        roots8.push_back(n8b);
        em_first.push_back(uint32_t(u.emitters.size()));
        em_count.push_back(uint32_t(p->emitters.size()));
        // nodes: inner children + nb, leaves' first slot + lb
        for (size_t i = 0; i < p->nodes.size(); i += 4) {
            bvh2_node nd;
            std::memcpy(&nd, &p->nodes[i], sizeof(nd));
            for (int c = 0; c < 2; c++) {
                const int32_t ch = nd.child[c];
                if (ch == k_empty_child) continue;
                if (ch >= 0) {
                    nd.child[c] = ch + int32_t(nb);
                } else {
                    const uint32_t w = uint32_t(ch);
                    const uint32_t first = (w & 0x07FFFFFFu) + lb;
                    nd.child[c] = int32_t((w & 0xF8000000u) | first);
                }
            }
            float4 q[4];
            std::memcpy(q, &nd, sizeof(nd));
            for (int k = 0; k < 4; k++) u.nodes.push_back(q[k]);
        }
        // four-wide nodes: inner children + n4b, leaves' first slot + lb (the binary tree's slots)
        for (size_t i = 0; i < p->nodes4.size(); i += 8) {
            bvh4_node nd;
            std::memcpy(&nd, &p->nodes4[i], sizeof(nd));
            for (int c = 0; c < 4; c++) {
                const int32_t ch = nd.child[c];
                if (ch == k_empty_child) continue;
                if (ch >= 0) {
                    nd.child[c] = ch + int32_t(n4b);
                } else {
                    const uint32_t w = uint32_t(ch);
                    nd.child[c] = int32_t((w & 0xF8000000u) | ((w & 0x07FFFFFFu) + lb));
                }
            }
            float4 q[8];
            std::memcpy(q, &nd, sizeof(nd));
            for (int k = 0; k < 8; k++) u.nodes4.push_back(q[k]);
        }
        // wide nodes: first inner child + n8b, first leaf slot + l8b
        for (size_t i = 0; i < p->nodes8.size(); i += 5) {
            bvh8_node nd;
            std::memcpy(&nd, &p->nodes8[i], sizeof(nd));
            nd.child_base += n8b;
            nd.prim_base += l8b;
            float4 q[5];
            std::memcpy(q, &nd, sizeof(nd));
            for (int k = 0; k < 5; k++) u.nodes8.push_back(q[k]);
        }
        // leaf slots: primitive ids into the merged id space (triangles first, then spheres)
        auto rebase_slots = [&](const std::vector<float4>& pa, const std::vector<float4>& pb,
                                const std::vector<float4>& pc, std::vector<float4>& oa, std::vector<float4>& ob,
                                std::vector<float4>& oc) {
            for (size_t sl = 0; sl < pa.size(); sl++) {
                const uint32_t w = word_of(pa[sl]);
                const uint32_t flags = w & ~k_prim_id_mask;
                const uint32_t id = w & k_prim_id_mask;
                const uint32_t nid = (w & k_prim_sphere) ? uint32_t(T) + sb + (id - p->nt) : tb + id;
                oa.push_back(with_word(pa[sl].x, pa[sl].y, pa[sl].z, flags | nid));
                ob.push_back(pb[sl]);
                oc.push_back(pc[sl]);
            }
        };
        rebase_slots(p->pa, p->pb, p->pc, u.pa, u.pb, u.pc);
        rebase_slots(p->pa8, p->pb8, p->pc8, u.pa8, u.pb8, u.pc8);
        u.pos.insert(u.pos.end(), p->pos.begin(), p->pos.end());
        if (any_nrm) {
            if (p->has_nrm) u.nrm.insert(u.nrm.end(), p->nrm.begin(), p->nrm.end());
            else u.nrm.resize(u.nrm.size() + size_t(p->nv) * 3, 0.0f);
        }
        if (any_uv) {
            if (p->has_uv) u.uv.insert(u.uv.end(), p->uv.begin(), p->uv.end());
            else u.uv.resize(u.uv.size() + size_t(p->nv) * 2, 0.0f);
        }
        for (uint32_t v : p->idx) u.idx.push_back(v + vb);
        for (uint32_t e : p->tri_entity) u.tri_entity.push_back(e + eb);
        if (any_tm) {
            if (p->has_tri_material)
                for (uint32_t m : p->tri_material) u.tri_material.push_back(m == RBW_NONE ? RBW_NONE : m + mb);
            else
                u.tri_material.resize(u.tri_material.size() + p->nt, RBW_NONE);
        }
        u.spheres.insert(u.spheres.end(), p->spheres.begin(), p->spheres.end());
        for (uint32_t e : p->sphere_entity) u.sphere_entity.push_back(e + eb);
        for (dentity e : p->ents) {
            if (e.shape == RBW_SHAPE_MESH) e.first += tb;
            else if (e.shape == RBW_SHAPE_SPHERE) e.first += sb;
            if (e.material != RBW_NONE) e.material += mb;
            u.ents.push_back(e);
        }
        u.mats.insert(u.mats.end(), p->mats.begin(), p->mats.end());
        for (uint32_t e : p->emitters) u.emitters.push_back(e + eb);
        u.depth = std::max(u.depth, p->depth);
        u.depth8 = std::max(u.depth8, p->depth8);
        u.stack4 = std::max(u.stack4, p->stack4);
        u.smin = f3{std::min(u.smin.x, p->smin.x), std::min(u.smin.y, p->smin.y), std::min(u.smin.z, p->smin.z)};
        u.smax = f3{std::max(u.smax.x, p->smax.x), std::max(u.smax.y, p->smax.y), std::max(u.smax.z, p->smax.z)};
        tb += p->nt;
        sb += p->ns;
        vb += p->nv;
        nb += uint32_t(p->nodes.size() / 4);
        lb += uint32_t(p->pa.size());
        n4b += uint32_t(p->nodes4.size() / 8);
        n8b += uint32_t(p->nodes8.size() / 5);
        l8b += uint32_t(p->pa8.size());
        eb += uint32_t(p->ents.size());
        mb += uint32_t(p->mats.size());
    }
}

void dev_scene::upload(const host_scene& h, hipStream_t s) {
    valid = false;
    nodes.upload(h.nodes.data(), h.nodes.size(), s);
    prim_a.upload(h.pa.data(), h.pa.size(), s);
    prim_b.upload(h.pb.data(), h.pb.size(), s);
    prim_c.upload(h.pc.data(), h.pc.size(), s);
    nodes4.upload(h.nodes4.data(), h.nodes4.size(), s);
    nodes8.upload(h.nodes8.data(), h.nodes8.size(), s);
    prim8_a.upload(h.pa8.data(), h.pa8.size(), s);
    prim8_b.upload(h.pb8.data(), h.pb8.size(), s);
    prim8_c.upload(h.pc8.data(), h.pc8.size(), s);
    pos.upload(h.pos.data(), h.pos.size(), s);
    if (h.has_nrm) nrm.upload(h.nrm.data(), h.nrm.size(), s);
    else nrm.ensure(4);
    has_uv = h.has_uv;
    if (h.has_uv) uv.upload(h.uv.data(), h.uv.size(), s);
    idx.upload(h.idx.data(), h.idx.size(), s);
    tri_entity.upload(h.tri_entity.data(), h.tri_entity.size(), s);
    has_tri_material = h.has_tri_material;
    if (h.has_tri_material) tri_material.upload(h.tri_material.data(), h.tri_material.size(), s);
    spheres.upload(h.spheres.data(), h.spheres.size(), s);
    sphere_entity.upload(h.sphere_entity.data(), h.sphere_entity.size(), s);
    ents.upload(h.ents.data(), h.ents.size(), s);
    mats.upload(h.mats.data(), h.mats.size(), s);
    emitters.upload(h.emitters.data(), h.emitters.size(), s);
    num_media = uint32_t(h.media.size());
    if (num_media) {
        media.upload(h.media.data(), h.media.size(), s);
        ent_media.upload(h.ent_media.data(), h.ent_media.size(), s);
        if (!h.grid.empty()) grid.upload(h.grid.data(), h.grid.size(), s);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    nt = h.nt;
    ns = h.ns;
    num_emitters = uint32_t(h.emitters.size());
    light_flags = h.light_flags;
    valid = true;
}

dscene dev_scene::view(uint32_t width) const {
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
    S.ents = ents.as<dentity>();
    S.mats = mats.as<dmaterial>();
    S.emitters = emitters.as<uint32_t>();
    S.num_emitters = num_emitters;
    S.num_tris = nt;
    S.num_spheres = ns;
    S.root = 0;
    S.nodes4 = nodes4.as<float4>();
    S.root4 = 0;
    S.nodes8 = nodes8.as<float4>();
    S.prim8_a = prim8_a.as<float4>();
    S.prim8_b = prim8_b.as<float4>();
    S.prim8_c = prim8_c.as<float4>();
    S.root8 = 0;
    S.width = width;
    S.light_flags = light_flags;
    S.num_media = num_media;
    S.media = num_media ? media.as<dmedium>() : nullptr;
    S.ent_media = num_media ? ent_media.as<uint32_t>() : nullptr;
    S.grid = num_media && grid.p ? grid.as<float>() : nullptr;
    return S;
}

}  // namespace host
}  // namespace rbw
