/*
 * TEST INFRASTRUCTURE ONLY: C API of the CPU oracle (liborc.so).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product (librbw.so) never links or calls it.
 *
 * The oracle restates rainbow-cpu's path-integrator render call on the CPU
 * with the reference's arithmetic order, BVH (median split, one primitive per
 * leaf), per-tile pcg32 stream and tile-parallel film.  It consumes the same
 * flat descriptors as the C ABI (include/rbw.h).
 */
#ifndef ORC_H
#define ORC_H

#include "../include/rbw.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_SAMPLER_TILE = 0,        /* the reference: one pcg32 per 16x16 tile, seeded with the tile index */
    ORC_SAMPLER_PER_SAMPLE = 1,  /* RBW_SAMPLER_PER_SAMPLE_PCG seeding */
    ORC_SAMPLER_REPLAY = 2,      /* consume rbw_render_params replay buffers */
    ORC_SAMPLER_TILE_STRATIFIED = 3  /* the reference's stratified samplers (1D and 2D) sharing the
                                        per-tile pcg32: strata_x * strata_y = spp, dims_1d / dims_2d
                                        stratified dimensions, random beyond */
};

typedef struct orc_render_opts {
    uint32_t sampler;            /* ORC_SAMPLER_* */
    uint32_t threads;            /* 0 = std::thread::hardware_concurrency() */
    uint32_t tile_stride;        /* render only tiles with index % tile_stride == 0 (0 or 1 = all) */
    uint64_t* record_offsets;    /* NULL or width*height*spp+1: replay offsets of the draws made */
    float* record_values;        /* NULL or record_capacity floats */
    uint64_t record_capacity;
    uint64_t record_count;       /* out: number of recorded draws */
    uint32_t strata_x, strata_y; /* ORC_SAMPLER_TILE_STRATIFIED */
    uint32_t dims_1d, dims_2d;
} orc_render_opts;

typedef struct orc_stats {
    double seconds;              /* tile loop + film merge (sampler_integrator.cpp:90-143) */
    uint64_t samples;
    uint64_t camera_rays, extension_rays, shadow_rays, mis_rays, light_casts;
    uint32_t threads;
} orc_stats;

typedef struct orc_scene orc_scene;

/* builds entities, emitter list and the two-level reference BVH (scene.cpp:30-80) */
int orc_scene_create(const rbw_scene_desc* desc, orc_scene** out);
void orc_scene_destroy(orc_scene* scene);
uint32_t orc_scene_num_bvh_nodes(const orc_scene* scene);

/* sampler_integrator::render + path_integrator::trace; out_rgb = film pixels (w*h*3) */
int orc_render(orc_scene* scene, const rbw_camera_desc* camera, const rbw_render_params* params,
               orc_render_opts* opts, float* out_rgb, orc_stats* stats);

/* scene::intersect (visible_only=0) / intersect_with_shadow_ray (visible_only=1) */
int orc_trace_rays(orc_scene* scene, uint32_t n, const float* origins, const float* directions,
                   const float* tmax, uint32_t visible_only, rbw_hit* out);

/* brute force: linear scan over all entities and primitives (scene.cpp:86-94, mesh.cpp:49-57) */
int orc_trace_rays_linear(orc_scene* scene, uint32_t n, const float* origins, const float* directions,
                          const float* tmax, uint32_t visible_only, rbw_hit* out);

/* perspective_camera::sample(position, sample) (perspective_camera.cpp:78-99) for n raster
 * positions raster_xy (2n) and lens samples lens_uv (2n): world origins and directions (3n each) */
void orc_camera_rays(const rbw_camera_desc* cam, uint32_t n, const float* raster_xy, const float* lens_uv,
                     float* out_o, float* out_d);

/* pcg32 of the reference: n uniform_real() draws of pcg32(seed) */
void orc_pcg32_reals(uint64_t seed, uint32_t n, float* out);
void orc_pcg32_u32(uint64_t seed, uint32_t n, uint32_t* out);
/* pcg32(seed).uint32(lo[i], hi[i]) for i < n (bounded draws with rejection) */
void orc_pcg32_range(uint64_t seed, uint32_t n, const uint32_t* lo, const uint32_t* hi, uint32_t* out);
/* the tables one pixel's reset() fills (stratified_sampler.hpp:82-127) from pcg32(seed)
 * advanced by `skip` draws: dims_1d x spp floats, then dims_2d x spp x 2 floats; returns
 * the number of draws made */
uint64_t orc_stratified_tables(uint64_t seed, uint64_t skip, uint32_t strata_x, uint32_t strata_y,
                               uint32_t dims_1d, uint32_t dims_2d, float* out_1d, float* out_2d);

/* warps and BSDF pieces, for closed-form tests */
void orc_concentric_sample_disk(float u0, float u1, float* out2);
void orc_cosine_sample_hemisphere(float u0, float u1, float* out3);
void orc_uniform_sample_triangle(float u0, float u1, float* out2);
/* GGX (Trowbridge-Reitz, visible normals): D(wh), pdf(wo,wi), eval f(wo,wi) with F=1 if eta==NULL */
float orc_ggx_D(float ax, float ay, const float* wh);
float orc_ggx_lambda(float ax, float ay, const float* w);
void orc_ggx_sample_wh(float ax, float ay, const float* wo, float u0, float u1, float* wh);
void orc_ggx_sample_f(float ax, float ay, const float* eta, const float* k, const float* wo,
                      float u0, float u1, float* wi, float* f, float* pdf);
float orc_ggx_pdf(float ax, float ay, const float* wo, const float* wi);
void orc_ggx_eval(float ax, float ay, const float* eta, const float* k, const float* wo,
                  const float* wi, float* f);
void orc_fresnel_conductor(float cos_theta_i, const float* eta, const float* k, float* out3);
/* fresnel_reflect_dielectric (fresnel_effect.cpp:5-33) */
float orc_fresnel_dielectric(float cos_theta_i, float eta_i, float eta_o);
/* a material's lobe collection in shading space (scattering_function_collection.cpp) */
float orc_material_pdf(const rbw_material* m, const float* wo, const float* wi);
void orc_material_eval(const rbw_material* m, const float* wo, const float* wi, float* f);
int orc_material_sample(const rbw_material* m, const float* wo, float u0, float u1, float* wi, float* f, float* pdf);
/* Henyey-Greenstein (utilities.cpp:46-51, henyey_greenstein.cpp:17-48): phase value; sample -> pdf */
float orc_hg_phase(float cos_theta, float g);
/* the float log and exp the media code uses on both sides (orc_math.hpp flog / fexp) */
float orc_flog(float x);
float orc_fexp(float x);
float orc_hg_sample(const float* wo, float g, float u0, float u1, float* wi);
/* emitter pdf from a medium vertex (zero normal): raw = 0 as the reference (entity::pdf<emitter>
 * transforms the interaction, normalising the zero normal), raw = 1 untransformed */
float orc_medium_emitter_pdf(orc_scene* s, uint32_t entity, const float* p, const float* wi, int raw);

#ifdef __cplusplus
}
#endif

#endif
