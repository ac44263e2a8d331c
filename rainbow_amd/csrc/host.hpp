// Host-side state behind the C ABI: error handling, device buffers, the
// flattened scene (host copy and device copy), and the context (one HIP stream,
// the single-render buffers, the guiding model and the batched scheduler).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rbw.h"
#include "device.hpp"
#include "guiding.hpp"

namespace rbw {
namespace host {

struct rbw_error : std::runtime_error {
    rbw_status code;
    rbw_error(rbw_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                                      \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess)                                                                                \
            throw ::rbw::host::rbw_error(_e == hipErrorOutOfMemory ? RBW_ERR_OUT_OF_MEMORY : RBW_ERR_DEVICE, \
                                         std::string(#expr) + ": " + hipGetErrorString(_e));                 \
    } while (0)

inline void require(bool cond, rbw_status code, const std::string& msg) {
    if (!cond) throw rbw_error(code, msg);
}

inline uint32_t blocks(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
    dbuf() = default;
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    ~dbuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        HIP_CHECK(hipMalloc(&p, n ? n : 4));
        bytes = n;
    }
    template <typename T>
    void upload(const T* src, size_t count, hipStream_t s) {
        ensure(count * sizeof(T));
        if (count) HIP_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

constexpr uint32_t k_instanced_threshold = 1u << 17;  // scene.cpp:43
constexpr uint32_t k_max_light_triangles = 4096;
constexpr uint32_t k_train_pool = 8;  // training samples kept per pass, in batches
constexpr uint32_t k_stack_depth = 64; // traversal stack entries (path.hpp RBW_STACK)

struct kernel_timer {
    struct rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<rec> recs;
    size_t used = 0;
    hipEvent_t get() {
        if (used < pool.size()) return pool[used++];
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        pool.push_back(e);
        used++;
        return e;
    }
    void reset() {
        recs.clear();
        used = 0;
    }
    void begin(int kind, hipStream_t s) {
        rec r{kind, get(), nullptr};
        HIP_CHECK(hipEventRecord(r.a, s));
        recs.push_back(r);
    }
    void end(hipStream_t s) {
        recs.back().b = get();
        HIP_CHECK(hipEventRecord(recs.back().b, s));
    }
    // milliseconds of the last begin/end pair (its end event must have completed)
    double elapsed_last() const {
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, recs.back().a, recs.back().b));
        return ms;
    }
    ~kernel_timer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// A validated, flattened scene with its BVH, in host memory: exactly the arrays
// the device copy holds.  Built once per rbw_upload_scene / rbw_add_scene.
struct host_scene {
    std::vector<float4> nodes;      // 4 per bvh2 node
    std::vector<float4> pa, pb, pc; // leaf slots (of the bvh2 and the bvh4)
    std::vector<float4> nodes4;     // 8 per bvh4 node
    std::vector<float4> nodes8;     // 5 per bvh8 node
    std::vector<float4> pa8, pb8, pc8;  // leaf slots of the bvh8
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx, tri_entity, tri_material, sphere_entity, emitters;
    std::vector<float4> spheres;
    std::vector<dentity> ents;
    std::vector<dmaterial> mats;
    std::vector<dmedium> media;     // empty: no media
    std::vector<uint32_t> ent_media;  // 2 per entity when media is not empty
    std::vector<float> grid;
    uint32_t nt = 0, ns = 0, nv = 0;
    bool has_nrm = false, has_uv = false, has_tri_material = false;
    uint32_t depth = 0, depth8 = 0, stack4 = 0;
    uint32_t light_flags = 0;  // k_lights_*
    f3 smin{0, 0, 0}, smax{0, 0, 0};
    size_t bytes() const;
};

// validation + SAH BVH (throws rbw_error)
void build_host_scene(const rbw_scene_desc* d, host_scene& out);

// One scene holding several: every index rebased (nodes, slots, primitives,
// vertices, entities, materials), each part's BVHs kept intact under their own
// roots.  roots/roots4/roots8/em_first/em_count describe part i inside the union.
void merge_scenes(const std::vector<const host_scene*>& parts, host_scene& out, std::vector<uint32_t>& roots,
                  std::vector<uint32_t>& roots4, std::vector<uint32_t>& roots8, std::vector<uint32_t>& em_first,
                  std::vector<uint32_t>& em_count);

constexpr uint32_t k_default_width = 4;  // BVH width of a render that does not name one (the fastest on the icosphere)

struct dev_scene {
    dbuf nodes, prim_a, prim_b, prim_c, nodes4, nodes8, prim8_a, prim8_b, prim8_c, pos, nrm, uv, idx, tri_entity,
        tri_material, spheres, sphere_entity, ents, mats, emitters, media, ent_media, grid;
    uint32_t nt = 0, ns = 0, num_emitters = 0, num_media = 0, light_flags = 0;
    bool has_uv = false, has_tri_material = false, valid = false;
    void upload(const host_scene& h, hipStream_t s);
    dscene view(uint32_t width = k_default_width) const;
};

// 0 -> the default width; otherwise 2, 4 or 8
uint32_t resolve_width(uint32_t requested);

}  // namespace host
}  // namespace rbw

// The batched scheduler's device-resident state (batch.hip)
struct rbw_batch_state;

struct rbw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the scene of rbw_upload_scene
    rbw::host::dev_scene main;
    bool has_scene = false;
    uint32_t bvh_depth = 0, bvh_nodes = 0;
    // render buffers
    rbw::host::dbuf pf, pu, prng, q0, q1, cq, counts, film_sum, film_w, film_nan, rgb, dbg, replay_off, replay_val,
        counters;
    uint32_t path_cap = 0;
    uint32_t* host_count = nullptr;  // pinned
    rbw::host::kernel_timer timer;
    // guiding
    rbw::f3 scene_min{0, 0, 0}, scene_max{0, 0, 0};
    rbw::guide::model gm;
    bool guide_ready = false;
    rbw::host::dbuf g_surf, g_mat, g_q, g_query, g_mix, g_rec, g_flag, g_tin, g_tdir, g_tc, g_io_in, g_io_out,
        g_io_dir, g_io_c;
    uint32_t g_cap = 0, g_rec_cap = 0, g_t_cap = 0;
    double guide_last_ms = 0.0;  // network kernels of the last rbw_guide_forward / rbw_guide_train
    // the job networks of guided batches (job i: network i)
    rbw::guide::model gjobs;
    // batched jobs: scenes added with rbw_add_scene, and the scheduler
    std::vector<std::unique_ptr<rbw::host::host_scene>> pool;
    rbw_batch_state* batch = nullptr;  // owned; released by batch_free

    void ensure_guide();
    ~rbw_ctx();
};

namespace rbw {
namespace host {
void submit_jobs(rbw_ctx* ctx, const rbw_job* jobs, size_t n, const rbw_batch_opts* opts, rbw_job_stats* job_stats,
                 rbw_batch_stats* stats);
void autotune_jobs(rbw_ctx* ctx, const rbw_job* jobs, size_t n, rbw_batch_opts* inout, double* ms);
void batch_free(rbw_ctx* ctx);
void render_scheduled(rbw_ctx* ctx, const rbw_camera_desc* cam, const rbw_render_params& pr, float* out_rgb,
                      rbw_stats* stats);
}  // namespace host
}  // namespace rbw
