/*
 * rbw.h -- C ABI of the MI355X wavefront path tracer (librbw.so).
 *
 * The library replaces the render call of rainbow-cpu's path integrator:
 *
 *   integrator::render(camera, scene)        rainbow-cpu/integrators/integrator.hpp:64-66
 *   sampler_integrator::render               rainbow-cpu/integrators/sampler_integrator.cpp:19-144
 *   path_integrator::trace                   rainbow-cpu/integrators/path_integrator.cpp:15-47
 *
 * The caller (the C++ wavefront_path_integrator, or any FFI) flattens the
 * scene and camera into the plain descriptors below.  Every pointer is a
 * host pointer owned by the caller; the library copies what it needs.
 * Errors are reported as rbw_status codes, with a message available from
 * rbw_last_error() on the calling thread; no C++ exception crosses the ABI.
 * One context owns one HIP stream; calls on one context are not thread-safe.
 */
#ifndef RBW_H
#define RBW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rbw_status;

enum {
    RBW_OK = 0,
    RBW_ERR_INVALID_ARGUMENT = 1,
    RBW_ERR_OUT_OF_MEMORY = 2,
    RBW_ERR_DEVICE = 3,
    RBW_ERR_NO_SCENE = 4,
    RBW_ERR_UNSUPPORTED = 5
};

#define RBW_NONE 0xFFFFFFFFu

/* ------------------------------------------------------------------ scene */

/* shape kinds: mesh (rainbow-cpu/shapes/mesh.hpp:11) and sphere (shapes/sphere.hpp:7) */
enum { RBW_SHAPE_NONE = 0, RBW_SHAPE_MESH = 1, RBW_SHAPE_SPHERE = 2 };

/* entity flags */
enum {
    RBW_ENTITY_EMITTER = 1u,             /* surface_light component (emitters/surface_light.hpp:7) */
    RBW_ENTITY_REVERSE_ORIENTATION = 2u, /* shape::reverse_orientation() (shapes/shape.hpp:81) */
    RBW_ENTITY_HAS_NORMALS = 4u,         /* mesh::has_normal() (shapes/mesh.cpp:215) */
    RBW_ENTITY_HAS_UVS = 8u,             /* mesh::has_uv() (shapes/mesh.cpp:220) */
    RBW_ENTITY_HAS_MEDIA = 16u           /* media component (media/medium.hpp:60-75): the entity's pair
                                            in rbw_scene_desc::entity_media */
};

/* materials: matte (materials/matte_material.cpp:13-29, Lambert only: sigma must be 0)
 *            metal (materials/metal_material.cpp:17-37, GGX conductor, visible-normal sampling) */
enum { RBW_MATERIAL_MATTE = 1, RBW_MATERIAL_METAL = 2, RBW_MATERIAL_PLASTIC = 3 };

typedef struct rbw_material {
    uint32_t type;
    float diffuse[3];   /* matte reflectance */
    float sigma;        /* matte Oren-Nayar sigma (only 0 is supported) */
    float eta[3];       /* metal: conductor eta */
    float k[3];         /* metal: conductor absorption */
    float alpha_u;      /* metal, plastic: Trowbridge-Reitz alpha_x (after any roughness map) */
    float alpha_v;      /* metal, plastic: Trowbridge-Reitz alpha_y (plastic: = alpha_u) */
    /* plastic (plastic_material.cpp:30-56): a Lambert lobe of `diffuse` and a GGX lobe of
     * reflectance `specular` with the dielectric Fresnel of eta_i = 1, eta_o = ior */
    float specular[3];
    float ior;
} rbw_material;

/* Participating media (media/medium.hpp:37-75) for the volume path integrator:
 *   homogeneous (media/homogeneous_medium.cpp:8-58): sigma_a, sigma_s, g;
 *   grid (media/heterogeneous_medium.cpp:5-120): sigma_t an image_texture3d over the unit
 *     cube of the medium's local space, sampled trilinearly as
 *     textures/detail/image_texture.hpp:119-162 does (x fastest, then y, then z), with
 *     constant albedo and g (constant_texture3d); world_to_local maps world points and
 *     vectors into that space and is not re-normalised (heterogeneous_medium.cpp:28-33). */
enum { RBW_MEDIUM_HOMOGENEOUS = 0, RBW_MEDIUM_GRID = 1 };

typedef struct rbw_medium {
    uint32_t type;             /* RBW_MEDIUM_* */
    float sigma_a[3];          /* homogeneous: absorption per unit length */
    float sigma_s[3];          /* homogeneous: scattering per unit length */
    float g;                   /* Henyey-Greenstein asymmetry, |g| < 1 (shared/phases/henyey_greenstein.cpp) */
    float albedo[3];           /* grid: sigma_s / sigma_t */
    uint32_t grid_offset;      /* grid: index of its first density in rbw_scene_desc::medium_grid */
    uint32_t grid_size[3];     /* grid: nx, ny, nz */
    float world_to_local[12];  /* grid: 3x4 row-major affine map (rows of [R | t]) */
} rbw_medium;

/* One rainbow entity (scenes/entity.hpp:45-70): shape + material + emitter.
 * Mesh entities hold world-space triangles [first, first+count) and carry the
 * identity transform; sphere entities hold sphere index `first` and carry the
 * translation to the sphere centre (the reference's sphere lives at the
 * origin of its local space, shapes/sphere.cpp:17-86). */
typedef struct rbw_entity {
    uint32_t shape;     /* RBW_SHAPE_* */
    uint32_t first;
    uint32_t count;
    uint32_t material;  /* index into materials, or RBW_NONE (invisible pass-through) */
    uint32_t flags;     /* RBW_ENTITY_* */
    float radiance[3];  /* surface_light radiance when RBW_ENTITY_EMITTER */
} rbw_entity;

typedef struct rbw_scene_desc {
    uint32_t num_vertices;
    const float* positions;          /* 3 * num_vertices, world space */
    const float* normals;            /* 3 * num_vertices or NULL */
    const float* uvs;                /* 2 * num_vertices or NULL */
    uint32_t num_triangles;
    const uint32_t* indices;         /* 3 * num_triangles */
    const uint32_t* triangle_entity; /* entity of each triangle */
    const uint32_t* triangle_material; /* per-triangle material or NULL (entity material) */
    uint32_t num_spheres;
    const float* spheres;            /* 4 * num_spheres: centre xyz, radius */
    const uint32_t* sphere_entity;
    uint32_t num_entities;
    const rbw_entity* entities;      /* in scene::add_entity order (scenes/scene.cpp:15-28) */
    uint32_t num_materials;
    const rbw_material* materials;
    /* media component of each entity (media::media, media/medium.hpp:60-75): two medium
     * indices per entity, outside then inside, RBW_NONE for vacuum; NULL: no entity has
     * media.  Only the volume path integrator reads them. */
    uint32_t num_media;
    const rbw_medium* media;
    const uint32_t* entity_media;
    uint32_t medium_grid_count;
    const float* medium_grid;        /* the grids' densities (sigma_t per unit length) */
} rbw_scene_desc;

/* ----------------------------------------------------------------- camera */

/* perspective_camera (cameras/perspective_camera.cpp:78-99) flattened:
 * both matrices are 4x4 column-major float32 as built by the host. */
typedef struct rbw_camera_desc {
    float raster_to_camera[16];
    float camera_to_world[16];
    uint32_t width, height;          /* film resolution */
    int32_t bound_min[2];            /* film pixels_bound (cameras/film.cpp:186-204) */
    int32_t bound_max[2];
    float lens_radius;               /* 0 = pinhole */
    float focus_distance;
} rbw_camera_desc;

/* ----------------------------------------------------------------- render */

enum {
    /* one pcg32 stream per camera sample, seeded from (seed, pixel, sample);
     * draws in the reference's order (sampler_integrator.cpp:104-126,
     * integrator.cpp:118-165, 380, path_integrator.cpp:43) */
    RBW_SAMPLER_PER_SAMPLE_PCG = 0,
    /* replay of recorded per-sample draws (the reference's per-tile stream) */
    RBW_SAMPLER_REPLAY = 1,
    /* stratified (the contract of stratified_sampler.hpp:48-127: per pixel, each of the
     * first dims_1d 1D and dims_2d 2D draws of the spp = strata_x*strata_y samples falls in
     * a distinct stratum, strata shuffled independently per dimension; later draws from
     * the per-sample pcg32 stream).  A design of this build: strata assigned by a
     * hash-based permutation and jitter per (seed, pixel, dimension), no tables. */
    RBW_SAMPLER_STRATIFIED = 2
};

enum { RBW_TRI_MOLLER_TRUMBORE = 0, RBW_TRI_WATERTIGHT = 1 };

enum {
    /* path_integrator (path_integrator.cpp:15-47), the default */
    RBW_INTEGRATOR_PATH = 0,
    /* direct_integrator (direct_integrator.cpp:20-140): emission plus emitter_samples
     * light samples and bsdf_samples BSDF samples at the first surface with lobes (the
     * ray passes through surfaces without lobes), MIS by the power heuristic over the
     * sample fractions; max_depth and rr_threshold are not used */
    RBW_INTEGRATOR_DIRECT = 1,
    /* volume_path_integrator (volume_path_integrator.cpp:13-79): the path integrator
     * through the scene's media (rbw_scene_desc::media): distance sampling in the
     * current medium, Henyey-Greenstein vertices with next-event estimation, beam
     * transmittance on light and MIS rays through invisible medium boundaries
     * (scene.cpp:114-148).  Camera rays start in vacuum. */
    RBW_INTEGRATOR_VOLUME_PATH = 2
};

enum {
    RBW_SHADOW_CLOSEST_VISIBLE = 0,  /* scene.cpp:97-112 semantics */
    RBW_SHADOW_ANY_HIT = 1           /* occlusion test with early exit */
};

/* debug_hits records, per camera sample and bounce, three u32 primitive ids:
 * [extension hit, shadow-ray hit, MIS-ray hit].  Triangles are numbered
 * 0..num_triangles-1, spheres num_triangles+i.  RBW_NONE = ray not traced,
 * RBW_MISS = traced and missed. */
#define RBW_MISS 0xFFFFFFFEu

/* Learned path guiding (no counterpart in the reference): the continuation
 * sample of every vertex with a BSDF is drawn from the MLP's vMF mixture with
 * probability mis_fraction, else from the BSDF, and weighted by the one-sample
 * MIS density p_G = mis_fraction * q + (1 - mis_fraction) * p_bsdf (unbiased).
 * The continuation doubles as the second technique of next-event estimation:
 * the light sample is weighted by the power heuristic against p_G and emission
 * the continuation reaches by the power heuristic against the light pdf (the
 * reference's separate MIS ray is not traced).  With train set, the context's
 * network is trained online between passes on the radiance the render's own
 * paths carried back along their continuation directions. */
typedef struct rbw_guiding_opts {
    uint32_t enable;
    float mis_fraction;              /* probability of sampling the guide, e.g. 0.5 */
    uint32_t train;                  /* train between passes */
    float learning_rate;             /* Adam learning rate, e.g. 1e-2 */
    uint32_t batch;                  /* training samples per Adam step (<= 2^20), e.g. 2^18 */
    uint32_t steps_per_pass;         /* Adam steps after each pass */
    float scene_min[3], scene_max[3];/* position normalisation; all zero = scene bounds */
    uint32_t light_records;          /* also train on unoccluded light samples (their MIS-weighted
                                        radiance), so the guide learns all incident light */
    uint32_t guide_depth;            /* vertices 0 .. guide_depth-1 of a path are guided; later ones
                                        use the plain estimator (light + BSDF samples); 0 = all */
    uint32_t product_targets;        /* 1: a training record's weight is f(wo, wi) |cos| L_i / pdf (the
                                        product the continuation's contribution follows), so the guide
                                        learns BSDF x cosine x incident radiance; 0: L_i / pdf */
} rbw_guiding_opts;

typedef struct rbw_render_params {
    uint32_t spp;
    uint32_t max_depth;
    float rr_threshold;              /* path_integrator threshold (path_integrator.hpp:9-12) */
    uint32_t seed;
    uint32_t sampler;                /* RBW_SAMPLER_* */
    uint32_t triangle_test;          /* RBW_TRI_* */
    uint32_t shadow_mode;            /* RBW_SHADOW_* */
    uint32_t samples_per_pass;       /* camera samples per pixel in flight per pass; 0 = auto */
    /* RBW_SAMPLER_REPLAY: draws of sample (pixel p = y*width + x, index s) are
     * replay_values[replay_offsets[p*spp+s] .. replay_offsets[p*spp+s+1]) */
    const uint64_t* replay_offsets;  /* width*height*spp + 1 entries */
    const float* replay_values;
    uint64_t replay_count;
    uint32_t* debug_hits;            /* NULL or width*height*spp*max_depth*3 entries */
    uint32_t count_visits;           /* nonzero: count BVH node visits and primitive tests */
    const rbw_guiding_opts* guiding; /* NULL = plain BSDF continuation (the reference's) */
    uint32_t time_kernels;           /* nonzero: per-kernel HIP-event times in rbw_stats::kernel_ms */
    uint32_t bvh_width;              /* 0 (default: 4), 2 (binary BVH), 4 (four-wide BVH) or 8 (eight-wide BVH) */
    /* RBW_SAMPLER_STRATIFIED: strata_x * strata_y must equal spp; stratified 1D and 2D
     * dimensions (at most 65535 each) */
    uint32_t strata_x, strata_y, dims_1d, dims_2d;
    uint32_t integrator;             /* RBW_INTEGRATOR_* */
    uint32_t emitter_samples, bsdf_samples;  /* RBW_INTEGRATOR_DIRECT (0 = 1) */
    /* RBW_INTEGRATOR_VOLUME_PATH: 0 = as the reference, where an emitter that a phase-function
     * sample reaches from a medium vertex contributes nothing (entity::pdf<emitter>,
     * detail/entity.hpp:66-72, normalises the medium interaction's zero normal, so the light
     * pdf is not a number and fails the test at integrator.cpp:262); 1 = it contributes with
     * the power-heuristic weight, which makes the estimator unbiased */
    uint32_t medium_mis;
    uint32_t schedule;               /* RBW_SCHEDULE_*: which wavefront schedule runs the plain estimator */
} rbw_render_params;

/* rbw_render_params.schedule: 0 picks per render; persistent path slots take the plain
 * estimator with the per-sample or stratified sampler only (no replay, debug hits, visit
 * counts, guiding, direct or volume integrator); passes of samples take every render */
enum {
    RBW_SCHEDULE_AUTO = 0,
    RBW_SCHEDULE_PERSISTENT = 1,
    RBW_SCHEDULE_PASSES = 2
};
/* Schedules: the plain estimator with the per-sample or stratified sampler (no replay, no debug
 * hits, no visit counts, no guiding) runs on persistent path slots regenerated in
 * place (the batched-jobs wavefront with one job, its bounce loop a hipGraph;
 * time_kernels runs it without the graph).  Every other mode runs in passes of
 * samples with compacted queues.  Both give the same image bit for bit. */

enum {
    RBW_KERNEL_RAYGEN = 0,
    RBW_KERNEL_EXTEND = 1,
    RBW_KERNEL_SHADE = 2,
    RBW_KERNEL_CONNECT = 3,
    RBW_KERNEL_FILM = 4,
    RBW_KERNEL_GUIDE_INFER = 5,      /* MLP inference over the guided vertices */
    RBW_KERNEL_SHADE_GUIDED = 6,     /* guided vertex: light, MIS and continuation samples */
    RBW_KERNEL_GUIDE_TRAIN = 7,      /* training-sample collection and Adam steps */
    RBW_KERNEL_COUNT = 8
};

typedef struct rbw_stats {
    double render_ms;                /* HIP-event time of the whole render on the context stream */
    double msamples_per_s;
    double mrays_per_s;
    uint64_t samples;
    uint64_t camera_rays;
    uint64_t extension_rays;         /* includes camera rays */
    uint64_t shadow_rays;
    uint64_t mis_rays;
    uint64_t light_casts;            /* light-local pdf casts (shapes/shape.cpp:46-60) */
    uint64_t node_visits;            /* when count_visits */
    uint64_t prim_tests;             /* when count_visits */
    uint64_t extend_node_visits;     /* node visits of extension rays only */
    uint64_t extend_prim_tests;
    double kernel_ms[RBW_KERNEL_COUNT];
    uint64_t kernel_launches[RBW_KERNEL_COUNT];
    uint64_t guide_samples;          /* guided + train: training samples collected */
    uint64_t guide_steps;            /* guided + train: Adam steps taken */
    double guide_loss;               /* guided + train: loss of the last step, before its update */
    uint32_t bvh_width;              /* the BVH width the render walked (2, 4 or 8) */
    uint32_t pad;
} rbw_stats;

typedef struct rbw_device_opts {
    int32_t device;                  /* HIP device ordinal */
} rbw_device_opts;

typedef struct rbw_ctx rbw_ctx;

typedef struct rbw_hit {
    uint32_t prim;                   /* primitive id or RBW_MISS */
    uint32_t entity;                 /* entity id or RBW_NONE */
    float t;                         /* entity-local ray parameter of the hit */
    float b1, b2;                    /* triangle barycentrics (0 for spheres) */
} rbw_hit;

/* Context with its own HIP stream on `opts->device` (NULL = device 0). */
rbw_status rbw_create(const rbw_device_opts* opts, rbw_ctx** out);
void rbw_destroy(rbw_ctx* ctx);

/* Thread-local message of the last failing call on this thread. */
const char* rbw_last_error(void);

/* Flattens, builds the BVH and uploads the scene (scene::build_accelerator,
 * scenes/scene.cpp:30-80).  Replaces any previous scene of the context. */
rbw_status rbw_upload_scene(rbw_ctx* ctx, const rbw_scene_desc* scene);

/* The render call: writes the film's pixel values, width*height*3 floats in
 * row-major RGB, into out_rgb (host memory; NULL keeps the film on the device).
 * Pixels outside the camera's bound are written as 0. */
rbw_status rbw_render(rbw_ctx* ctx, const rbw_camera_desc* camera,
                      const rbw_render_params* params, float* out_rgb,
                      rbw_stats* stats);

/* Closest-hit (visible_only = 0, scenes/scene.cpp:82-95) or closest-visible
 * (visible_only = 1, scenes/scene.cpp:97-112) queries for n world-space rays:
 * origins/directions 3*n floats (directions normalised by the caller as
 * shared::ray does, shared/ray.cpp:7-10), tmax n floats (world length). */
rbw_status rbw_trace_rays(rbw_ctx* ctx, uint32_t n, const float* origins,
                          const float* directions, const float* tmax,
                          uint32_t visible_only, uint32_t triangle_test,
                          uint32_t bvh_width /* 0 (default), 2, 4 or 8 */, rbw_hit* out);

/* rbw_trace_rays without any device allocation: the same queries, with the rays and hit
 * records staged in a fixed device buffer that the library holds per device from its load
 * until the process ends (12 MiB), instead of in memory allocated per call.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code running
 * under HIP graph capture, serving loops that forbid allocation).  n must be at most
 * RBW_STATIC_MAX_RAYS (RBW_ERR_INVALID_ARGUMENT beyond).  The buffer is shared by every
 * context on a device: calls of one process's threads on one device take turns; calls
 * that overlap on one device from several processes each use their own copy. */
#define RBW_STATIC_MAX_RAYS 262144u
rbw_status rbw_trace_rays_static(rbw_ctx* ctx, uint32_t n, const float* origins, const float* directions,
                                 const float* tmax, uint32_t visible_only, uint32_t triangle_test,
                                 uint32_t bvh_width /* 0 (default), 2, 4 or 8 */, rbw_hit* out);

/* Diagnostic: the draws the sampler of `params` (per-sample pcg32 or stratified)
 * gives the spp samples of pixel p (y*width + x): per sample, n1 1D draws then n2 2D
 * draws (x, y), out = spp * (n1 + 2*n2) floats.  Replaces sampler::next
 * (samplers/samplers.hpp:31-33) called outside a render. */
rbw_status rbw_sampler_draws(rbw_ctx* ctx, const rbw_render_params* params, uint32_t pixel, uint32_t n1,
                             uint32_t n2, float* out);

/* ---------------------------------------------------------------- guiding
 * The guiding network of a context (no counterpart in the reference).
 * Parameters use a padded layout of RBW_GUIDE_PARAMS floats: for each layer
 * l = 0..5, W_l[out][in] row-major then b_l[out], with in = 224, 64, 64, 64,
 * 64, 64 and out = 64, 64, 64, 64, 64, 48 (216 live inputs, 40 live outputs). */
#define RBW_GUIDE_PARAMS 34160

/* He-uniform re-initialisation from `seed`, Adam state cleared */
rbw_status rbw_guide_reset(rbw_ctx* ctx, uint32_t seed);
rbw_status rbw_guide_get_params(rbw_ctx* ctx, float* params /* RBW_GUIDE_PARAMS */);
rbw_status rbw_guide_set_params(rbw_ctx* ctx, const float* params /* RBW_GUIDE_PARAMS */);
/* batched inference: inputs n x 9 in [0,1] (position, (normal+1)/2, (wo+1)/2),
 * out n x 48 raw head outputs (lobe j: mean xyz, kappa, weight logit at 5j..5j+4) */
rbw_status rbw_guide_forward(rbw_ctx* ctx, uint32_t n, const float* inputs, float* out);
/* one training step on n samples: loss = -sum c_i log q(d_i) / sum c_i.  grad
 * (nullable, RBW_GUIDE_PARAMS) receives d loss / d params; apply = 1 also takes
 * an Adam step with learning_rate.  loss (nullable) receives the loss. */
rbw_status rbw_guide_train(rbw_ctx* ctx, uint32_t n, const float* inputs, const float* dirs, const float* c,
                           float learning_rate, uint32_t apply, float* loss, float* grad);
/* HIP-event time, on the context stream, of the network kernels of the last
 * rbw_guide_forward or rbw_guide_train call: host transfers of the inputs and
 * outputs excluded, the read-back of grad / loss included when requested;
 * 0 before the first such call */
rbw_status rbw_guide_last_ms(rbw_ctx* ctx, double* ms);

/* ---------------------------------------------------------------- batches
 * Many render jobs of uneven size in one wavefront (no counterpart in the
 * reference, whose integrator::render, integrators/integrator.hpp:64-66,
 * renders one camera of one scene per call).  Scenes are added to the
 * context's scene pool once; a job names its scene, camera and sampling
 * parameters.  All jobs of a submission share one pool of path slots: a path
 * that ends is replaced by the next camera sample of the job queue, and the
 * bounce loop (regenerate, extend, shade, connect, retire) runs as a captured
 * hipGraph.  Each job's image equals the image rbw_render gives for the same
 * scene, camera and parameters (per-sample pcg32 stream, same seed). */

/* adds a scene to the pool (validated and BVH-built; the caller keeps the arrays) */
rbw_status rbw_add_scene(rbw_ctx* ctx, const rbw_scene_desc* scene, uint32_t* scene_id);
/* frees a pooled scene's host copy; its id stays invalid afterwards */
rbw_status rbw_remove_scene(rbw_ctx* ctx, uint32_t scene_id);

typedef struct rbw_job {
    uint32_t scene;                  /* id from rbw_add_scene */
    uint32_t spp;
    uint32_t max_depth;
    float rr_threshold;              /* path_integrator.cpp:35-43 */
    uint32_t seed;                   /* per-sample pcg32 stream seed (RBW_SAMPLER_PER_SAMPLE_PCG) */
    rbw_camera_desc camera;
    float* out_rgb;                  /* width*height*3 host floats, or NULL */
} rbw_job;

typedef struct rbw_batch_opts {
    uint32_t block_size;             /* threads per block of the wavefront kernels: 64, 128, 256 or 512; 0 = 256 */
    uint32_t pool_size;              /* path slots of the shared launch; 0 = default 2^21.  Jobs with at
                                        least this many samples run alone, one after another, with at
                                        least a single render's 2^24 slots (fewer for smaller jobs); the
                                        rest share the launch */
    uint32_t bvh_width;              /* 0 (default: 4), 2, 4 or 8; the autotuner tries all three */
    uint32_t use_graph;              /* 1: replay the bounce loop as a hipGraph; 0: plain launches */
    uint32_t autotune;               /* 1: first probe the jobs themselves (their real spp, each probe
                                        cut after 12 graph launches) over BVH widths 2/8, block sizes
                                        64..512 and pools of 2^20..2^22 slots, and run the batch with
                                        the configuration of the least estimated batch time */
    uint32_t triangle_test;          /* RBW_TRI_* */
    uint32_t shadow_mode;            /* RBW_SHADOW_* */
    /* NULL: the plain estimator.  Otherwise every job is rendered with learned guiding (the
     * estimator of rbw_guiding_opts) by its own network, job i by network i of the context's
     * job networks (rbw_guide_jobs_reset): all jobs share one persistent launch, their guided
     * vertices are inferred by one launch over all networks, and with train set every network
     * takes steps_per_pass Adam steps of up to `batch` of its job's training rows after every
     * 8 iterations of the bounce loop, all networks in one batched training kernel.
     * light_records must be 0 and scene_min/scene_max are ignored (each job's scene box). */
    const rbw_guiding_opts* guiding;
} rbw_batch_opts;

typedef struct rbw_job_stats {
    uint64_t samples;
    double active_ms;                /* from the job's first camera sample to its last retired path */
    double msamples_per_s;           /* samples / active_ms */
    uint32_t first_iteration, last_iteration;
} rbw_job_stats;

typedef struct rbw_batch_stats {
    double render_ms;                /* HIP-event time of the whole batch (films resolved) */
    double msamples_per_s;           /* all jobs' samples / render_ms */
    double mrays_per_s;
    double autotune_ms;              /* probe time when autotune = 1 */
    uint64_t samples, extension_rays, shadow_rays, mis_rays;
    uint32_t iterations;             /* bounce-loop iterations */
    uint32_t graph_launches;
    uint32_t block_size, pool_size, bvh_width;  /* the configuration that ran */
    uint32_t scene_triangles;        /* triangles of the merged scene */
    uint64_t guide_samples;          /* guided batches: training rows the jobs offered */
    uint32_t guide_steps;            /* guided batches: batched training steps taken */
    uint32_t pad;
} rbw_batch_stats;

/* renders every job; job_stats (nullable) has n entries */
rbw_status rbw_submit_jobs(rbw_ctx* ctx, const rbw_job* jobs, uint32_t n, const rbw_batch_opts* opts,
                           rbw_job_stats* job_stats, rbw_batch_stats* stats);
/* The job networks of guided batches (no counterpart in the reference): n_networks networks
 * of RBW_GUIDE_PARAMS parameters each (at most 32), He-uniform from seed + network index, Adam
 * state cleared.  Job i of a guided rbw_submit_jobs call uses network i. */
rbw_status rbw_guide_jobs_reset(rbw_ctx* ctx, uint32_t n_networks, uint32_t seed);
rbw_status rbw_guide_jobs_get_params(rbw_ctx* ctx, uint32_t network, float* params /* RBW_GUIDE_PARAMS */);
rbw_status rbw_guide_jobs_set_params(rbw_ctx* ctx, uint32_t network, const float* params);
/* one training step of the first n_networks job networks in one batched launch: network k
 * trains on rows[k] records (inputs x 9, dirs x 3, c), the records of the networks one
 * after another; loss (nullable) receives n_networks losses and grad (nullable) the
 * n_networks gradients (RBW_GUIDE_PARAMS each), as rbw_guide_train does for one network */
rbw_status rbw_guide_jobs_train(rbw_ctx* ctx, uint32_t n_networks, const uint32_t* rows, const float* inputs,
                                const float* dirs, const float* c, float learning_rate, uint32_t apply, float* loss,
                                float* grad);

/* the autotuner alone: fills block_size/pool_size/bvh_width of *opts with the
 * fastest candidate on a 1-spp probe of the jobs */
rbw_status rbw_autotune_jobs(rbw_ctx* ctx, const rbw_job* jobs, uint32_t n, rbw_batch_opts* opts,
                             double* probe_ms);

/* ---- film output (host code; no context needed) ----
 * The film's 8-bit image as film::write makes it (cameras/film.cpp:114-153):
 * every byte 255, then for each pixel inside [bound_min, bound_max) (NULL: the whole
 * film) and channel c: clamp(255 * gamma_correct(rgb * scale) + 0.5, 0, 255)
 * truncated to a byte (gamma_correct: utilities.cpp:33-38), alpha 255.
 * rgb: width*height*3 floats as rbw_render writes them; out_rgba: width*height*4. */
rbw_status rbw_film_to_rgba8(const float* rgb, uint32_t width, uint32_t height, const int32_t* bound_min,
                             const int32_t* bound_max, float scale, uint8_t* out_rgba);
/* rbw_film_to_rgba8 written as an 8-bit RGBA PNG to `path` (film::write +
 * file_system::write_image, rainbow-core/file_system.cpp:17-24, which appends
 * ".png" to the name; this call takes the whole path) */
rbw_status rbw_write_png(const char* path, const float* rgb, uint32_t width, uint32_t height,
                         const int32_t* bound_min, const int32_t* bound_max, float scale);
/* the film's float pixel values as a colour PFM (little-endian, rows bottom-up):
 * the raw FP32 dump for error measurements; no counterpart in the reference */
rbw_status rbw_write_pfm(const char* path, const float* rgb, uint32_t width, uint32_t height);

#ifdef __cplusplus
}
#endif

#endif /* RBW_H */
