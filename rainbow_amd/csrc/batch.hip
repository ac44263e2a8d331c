// Batched rendering: many jobs (scene, camera, spp, depth) in one wavefront.
//
// Every job's scene goes into one merged scene (a shared BVH pool: each part
// keeps its own tree under its own root).  All jobs share a pool of path slots.
// The job queue is the concatenation of the jobs' camera samples, job after
// job, each job's samples in the order a single render makes them (sample
// index outer, pixel inner).  One iteration of the bounce loop:
//   plan     how many free slots get a new camera sample this iteration
//   regen    camera rays for those samples into the free slots (path regeneration)
//   extend   closest hit of every live path
//   shade    vertex work; survivors to the next queue, next-event estimation
//            to the connect queue, finished paths to the retire queue
//   connect  shadow and MIS rays of next-event estimation
//   retire   finished paths into their job's film, slots back to the free stack
// Two iterations (the live queue alternates) are captured as one hipGraph and
// replayed until the job queue is empty and every slot is free.  Grids are
// fixed at capture; each kernel reads its queue length from device memory.
//
// A job's image is bit-identical to rbw_render's: the same per-sample pcg32
// stream, the same BVH (built by the same builder from the same scene), and
// order-independent 64-bit fixed-point film sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "host.hpp"
#include "path.hpp"

using namespace rbw;
using namespace rbw::host;

namespace {

struct djob {
    dcamera cam;
    dfilm film;
    unsigned long long base;   // first index of the job in the batch's sample sequence
    unsigned long long count;  // npix * spp
    uint32_t root, em_first, em_count, spp;
    uint32_t max_depth, seed, npix, bw;
    float rr_threshold;
    uint32_t pad[3];
};

struct dctl {
    unsigned long long cursor;    // next sample of the job queue
    unsigned long long total;     // samples of all jobs
    unsigned long long gen_base;  // first sample regenerated this iteration
    uint32_t free_top;            // free-slot stack height
    uint32_t n_new;               // slots regenerated this iteration
    uint32_t slot_base;           // free_stack[slot_base, slot_base + n_new) are taken
    uint32_t iteration;
    uint32_t count[4];            // live queue A, live queue B, connect queue, retire queue
};

struct dbatch {
    dctl* ctl;
    const djob* jobs;
    uint32_t njobs;
    uint32_t* free_stack;
    uint32_t* q[2];
    uint32_t* cq;
    uint32_t* retire;
    unsigned long long* retired;   // per job
    uint32_t* first_iter;          // per job
    uint32_t* last_iter;           // per job
    unsigned long long* counters;  // extension, shadow, MIS, light casts
};

__device__ __forceinline__ uint32_t find_job(const djob* jobs, uint32_t n, unsigned long long g) {
    uint32_t lo = 0, hi = n;  // jobs[lo].base <= g < jobs[hi].base
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].base <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// the job's view of the merged scene and of the render parameters
__device__ __forceinline__ void job_view(const dscene& S, const dparams& base, const djob& J, dscene& Sj, dparams& pj) {
    Sj = S;
    Sj.root = J.root;
    Sj.emitters = S.emitters + J.em_first;
    Sj.num_emitters = J.em_count;
    pj = base;
    pj.spp = J.spp;
    pj.max_depth = J.max_depth;
    pj.rr_threshold = J.rr_threshold;
    pj.seed = J.seed;
}

__global__ void k_b_init(dbatch B, uint32_t pool, unsigned long long total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < pool) B.free_stack[i] = i;
    if (i < B.njobs) {
        B.retired[i] = 0;
        B.first_iter[i] = 0xFFFFFFFFu;
        B.last_iter[i] = 0xFFFFFFFFu;
    }
    if (i < 4) B.counters[i] = 0;
    if (i == 0) {
        dctl c;
        c.cursor = 0;
        c.total = total;
        c.gen_base = 0;
        c.free_top = pool;
        c.n_new = 0;
        c.slot_base = 0;
        c.iteration = 0;
        c.count[0] = c.count[1] = c.count[2] = c.count[3] = 0;
        *B.ctl = c;
    }
}

__global__ void k_b_plan(dbatch B) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    dctl& c = *B.ctl;
    const unsigned long long left = c.total - c.cursor;
    const uint32_t n = left < (unsigned long long)c.free_top ? (uint32_t)left : c.free_top;
    c.n_new = n;
    c.slot_base = c.free_top - n;
    c.gen_base = c.cursor;
    c.cursor += n;
    c.free_top -= n;
    c.iteration += 1;
}

__global__ void k_b_regen(dscene S, dpaths P, dparams base, dbatch B, int cur) {
    const uint32_t n = B.ctl->n_new;
    const uint32_t sbase = B.ctl->slot_base;
    const unsigned long long g0 = B.ctl->gen_base;
    const uint32_t iter = B.ctl->iteration;
    // grid-stride over [0, n) in whole blocks, so every lane of a wave takes the same trips
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        const bool act = i < n;
        uint32_t slot = 0;
        if (act) {
            slot = B.free_stack[sbase + i];
            const unsigned long long g = g0 + i;
            const uint32_t j = find_job(B.jobs, B.njobs, g);
            const djob& J = B.jobs[j];
            const unsigned long long l = g - J.base;
            const uint32_t s = (uint32_t)(l / J.npix), li = (uint32_t)(l % J.npix);
            dscene Sj;
            dparams pj;
            job_view(S, base, J, Sj, pj);
            raygen_path(J.cam, pj, P, slot, J.cam.bx0 + (int32_t)(li % J.bw), J.cam.by0 + (int32_t)(li / J.bw), s);
            P.U(U_JOB, slot) = j;
            if (l == 0) B.first_iter[j] = iter;
        }
        const uint32_t k = wave_append(&B.ctl->count[cur], act);
        if (act) B.q[cur][k] = slot;
    }
}

__global__ void k_b_extend(dscene S, dpaths P, dparams base, dbatch B, int cur) {
    const uint32_t n = B.ctl->count[cur];
    trace_counters cnt{0, 0};
    uint32_t rays = 0;
    // grid-stride over [0, n) in whole blocks, so every lane of a wave takes the same trips
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        const bool act = i < n;
        if (act) {
            const uint32_t slot = B.q[cur][i];
            const djob& J = B.jobs[P.U(U_JOB, slot)];
            dscene Sj;
            dparams pj;
            job_view(S, base, J, Sj, pj);
            extend_path<false>(Sj, P, pj, slot, cnt);
            rays++;
        }
    }
    wave_add64(B.counters + 0, rays);
}

__global__ void k_b_shade(dscene S, dpaths P, dparams base, dbatch B, int cur) {
    const uint32_t n = B.ctl->count[cur];
    const int nxt = cur ^ 1;
    // grid-stride over [0, n) in whole blocks, so every lane of a wave takes the same trips
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        const bool act = i < n;
        bool alive = false, need_connect = false, to_guide = false;
        uint32_t slot = 0, gmat = 0;
        surf gsi;
        if (act) {
            slot = B.q[cur][i];
            const djob& J = B.jobs[P.U(U_JOB, slot)];
            dscene Sj;
            dparams pj;
            job_view(S, base, J, Sj, pj);
            shade_path<false>(Sj, P, pj, slot, alive, need_connect, to_guide, gsi, gmat);
        }
        const uint32_t a = wave_append(&B.ctl->count[nxt], alive);
        if (alive) B.q[nxt][a] = slot;
        const uint32_t c = wave_append(&B.ctl->count[2], need_connect);
        if (need_connect) B.cq[c] = slot;
        const bool dead = act && !alive;
        const uint32_t r = wave_append(&B.ctl->count[3], dead);
        if (dead) B.retire[r] = slot;
    }
}

__global__ void k_b_connect(dscene S, dpaths P, dparams base, dbatch B) {
    const uint32_t n = B.ctl->count[2];
    trace_counters cnt{0, 0};
    uint32_t nshadow = 0, nmis = 0, nlight = 0;
    // grid-stride over [0, n) in whole blocks, so every lane of a wave takes the same trips
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        const bool act = i < n;
        if (act) {
            const uint32_t slot = B.cq[i];
            const djob& J = B.jobs[P.U(U_JOB, slot)];
            dscene Sj;
            dparams pj;
            job_view(S, base, J, Sj, pj);
            connect_path<false>(Sj, P, pj, slot, cnt, nshadow, nmis, nlight);
        }
    }
    wave_add64(B.counters + 1, nshadow);
    wave_add64(B.counters + 2, nmis);
    wave_add64(B.counters + 3, nlight);
}

__global__ void k_b_retire(dpaths P, dbatch B) {
    const uint32_t n = B.ctl->count[3];
    const uint32_t iter = B.ctl->iteration;
    // grid-stride over [0, n) in whole blocks, so every lane of a wave takes the same trips
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        const bool act = i < n;
        uint32_t slot = 0, j = 0xFFFFFFFFu;
        if (act) {
            slot = B.retire[i];
            j = P.U(U_JOB, slot);
            const djob& J = B.jobs[j];
            int32_t x0, y0, x1, y1;
            footprint(J.cam, P.F(F_PX, slot), P.F(F_PY, slot), x0, y0, x1, y1);
            film_add_span(J.cam, J.film, x0, y0, x1, y1, P.F3(F_VX, slot));
        }
        // per-job retired counts, one atomic per (wave, job)
        const int leader = __ffsll((unsigned long long)__ballot(act)) - 1;
        const uint32_t jl = __shfl(j, leader < 0 ? 0 : leader);
        const bool same = act && j == jl;
        const uint32_t nsame = (uint32_t)__popcll(__ballot(same));
        if (leader >= 0 && (int)__lane_id() == leader) {
            const unsigned long long before = atomicAdd(&B.retired[jl], (unsigned long long)nsame);
            if (before + nsame == B.jobs[jl].count) B.last_iter[jl] = iter;
        }
        if (act && !same) {
            const unsigned long long before = atomicAdd(&B.retired[j], 1ull);
            if (before + 1 == B.jobs[j].count) B.last_iter[j] = iter;
        }
        const uint32_t k = wave_append(&B.ctl->free_top, act);
        if (act) B.free_stack[k] = slot;
    }
}

// the consumed live queue and the connect/retire queues start empty next iteration
__global__ void k_b_reset(dbatch B, int cur) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    B.ctl->count[cur] = 0;
    B.ctl->count[2] = 0;
    B.ctl->count[3] = 0;
}

__global__ void k_b_resolve(dcamera cam, dfilm film, float* rgb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cam.width * cam.height) return;
    film_resolve_pixel(cam, film, p, rgb);
}

}  // namespace

struct rbw_batch_state {
    // merged scene of the last submission
    std::vector<uint32_t> scene_ids;
    dev_scene scene;
    std::vector<uint32_t> roots, em_first, em_count;
    uint32_t merged_tris = 0;
    // path pool
    dbuf pf, pu, prng, q0, q1, cq, retire, free_stack;
    uint32_t pool = 0;
    // per batch
    dbuf ctl, jobs, retired, first_iter, last_iter, counters, films, rgb;
    dctl* host_ctl = nullptr;  // pinned
    // captured bounce loop
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    ~rbw_batch_state() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
        if (host_ctl) (void)hipHostFree(host_ctl);
    }
};

namespace rbw {
namespace host {

void batch_free(rbw_ctx* ctx) {
    delete ctx->batch;
    ctx->batch = nullptr;
}

namespace {

struct batch_config {
    uint32_t block = 256, pool = 1u << 21, width = 2;
    bool graph = true;
    uint32_t tri_test = RBW_TRI_MOLLER_TRUMBORE, shadow = RBW_SHADOW_CLOSEST_VISIBLE;
};

void validate_jobs(rbw_ctx* ctx, const rbw_job* jobs, size_t n) {
    require(jobs != nullptr && n > 0, RBW_ERR_INVALID_ARGUMENT, "no jobs");
    require(n < (1u << 20), RBW_ERR_UNSUPPORTED, "too many jobs");
    for (size_t i = 0; i < n; i++) {
        const rbw_job& J = jobs[i];
        require(J.scene < ctx->pool.size() && ctx->pool[J.scene], RBW_ERR_INVALID_ARGUMENT,
                "job " + std::to_string(i) + ": unknown scene id");
        require(J.spp > 0 && J.max_depth > 0 && J.max_depth < (1u << 20), RBW_ERR_INVALID_ARGUMENT,
                "job " + std::to_string(i) + ": spp and max_depth must be positive");
        const rbw_camera_desc& c = J.camera;
        require(c.width > 0 && c.height > 0, RBW_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": empty film");
        require(c.bound_min[0] >= 0 && c.bound_min[1] >= 0 && c.bound_max[0] <= (int32_t)c.width &&
                    c.bound_max[1] <= (int32_t)c.height && c.bound_min[0] < c.bound_max[0] &&
                    c.bound_min[1] < c.bound_max[1],
                RBW_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": pixel bounds outside the film");
        require(uint64_t(c.width) * c.height <= (1u << 28), RBW_ERR_UNSUPPORTED,
                "job " + std::to_string(i) + ": film too large");
    }
}

batch_config config_of(const rbw_batch_opts* o) {
    batch_config c;
    if (!o) return c;
    if (o->block_size) c.block = o->block_size;
    if (o->pool_size) c.pool = o->pool_size;
    if (o->bvh_width) c.width = o->bvh_width;
    c.graph = o->use_graph != 0;
    c.tri_test = o->triangle_test;
    c.shadow = o->shadow_mode;
    require(c.block >= 64 && c.block <= 1024 && (c.block & 63) == 0, RBW_ERR_INVALID_ARGUMENT,
            "block_size must be a multiple of 64 in [64, 1024]");
    require(c.pool >= 1024 && c.pool <= (1u << 26), RBW_ERR_INVALID_ARGUMENT, "pool_size must be in [2^10, 2^26]");
    require(c.width == 2, RBW_ERR_UNSUPPORTED, "bvh_width: only the binary BVH is built");
    require(c.tri_test == RBW_TRI_MOLLER_TRUMBORE || c.tri_test == RBW_TRI_WATERTIGHT, RBW_ERR_INVALID_ARGUMENT,
            "unknown triangle test");
    require(c.shadow == RBW_SHADOW_CLOSEST_VISIBLE || c.shadow == RBW_SHADOW_ANY_HIT, RBW_ERR_INVALID_ARGUMENT,
            "unknown shadow mode");
    return c;
}

// the merged scene of the jobs' scenes (distinct ids in first-use order), cached
void prepare_scene(rbw_ctx* ctx, const rbw_job* jobs, size_t n, std::vector<uint32_t>& part_of_job) {
    rbw_batch_state& b = *ctx->batch;
    std::vector<uint32_t> ids;
    std::map<uint32_t, uint32_t> part;
    part_of_job.resize(n);
    for (size_t i = 0; i < n; i++) {
        auto it = part.find(jobs[i].scene);
        if (it == part.end()) {
            it = part.emplace(jobs[i].scene, uint32_t(ids.size())).first;
            ids.push_back(jobs[i].scene);
        }
        part_of_job[i] = it->second;
    }
    if (ids == b.scene_ids && b.scene.valid) return;
    std::vector<const host_scene*> parts;
    for (uint32_t id : ids) parts.push_back(ctx->pool[id].get());
    host_scene merged;
    merge_scenes(parts, merged, b.roots, b.em_first, b.em_count);
    b.scene.upload(merged, ctx->stream);
    b.scene_ids = ids;
    b.merged_tris = merged.nt;
}

void ensure_pool(rbw_batch_state& b, uint32_t pool) {
    if (b.pool == pool) return;
    b.pf.release();
    b.pu.release();
    b.prng.release();
    b.pf.ensure(size_t(F_COUNT) * pool * sizeof(float));
    b.pu.ensure(size_t(U_COUNT) * pool * sizeof(uint32_t));
    b.prng.ensure(size_t(pool) * sizeof(uint64_t));
    b.q0.ensure(size_t(pool) * 4);
    b.q1.ensure(size_t(pool) * 4);
    b.cq.ensure(size_t(pool) * 4);
    b.retire.ensure(size_t(pool) * 4);
    b.free_stack.ensure(size_t(pool) * 4);
    b.pool = pool;
}

struct run_result {
    double ms = 0.0;
    uint32_t iterations = 0, launches = 0;
    unsigned long long counters[4] = {0, 0, 0, 0};
    std::vector<double> launch_ms;  // end time of each launch from the start
    std::vector<uint32_t> first_iter, last_iter;
};

// one batch through the bounce loop; films must be cleared by the caller
void run_batch(rbw_ctx* ctx, const std::vector<djob>& hjobs, const batch_config& cfg, run_result& rr,
               bool want_job_times) {
    rbw_batch_state& b = *ctx->batch;
    hipStream_t s = ctx->stream;
    const uint32_t nj = uint32_t(hjobs.size());
    unsigned long long total = 0;
    uint32_t max_depth = 0;
    for (const djob& J : hjobs) {
        total += J.count;
        max_depth = std::max(max_depth, J.max_depth);
    }
    ensure_pool(b, cfg.pool);
    b.ctl.ensure(sizeof(dctl));
    b.jobs.upload(hjobs.data(), hjobs.size(), s);
    b.retired.ensure(size_t(nj) * 8);
    b.first_iter.ensure(size_t(nj) * 4);
    b.last_iter.ensure(size_t(nj) * 4);
    b.counters.ensure(4 * 8);
    if (!b.host_ctl) HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b.host_ctl), sizeof(dctl), hipHostMallocDefault));

    dbatch B;
    B.ctl = b.ctl.as<dctl>();
    B.jobs = b.jobs.as<djob>();
    B.njobs = nj;
    B.free_stack = b.free_stack.as<uint32_t>();
    B.q[0] = b.q0.as<uint32_t>();
    B.q[1] = b.q1.as<uint32_t>();
    B.cq = b.cq.as<uint32_t>();
    B.retire = b.retire.as<uint32_t>();
    B.retired = b.retired.as<unsigned long long>();
    B.first_iter = b.first_iter.as<uint32_t>();
    B.last_iter = b.last_iter.as<uint32_t>();
    B.counters = b.counters.as<unsigned long long>();
    dpaths P;
    P.f = b.pf.as<float>();
    P.u = b.pu.as<uint32_t>();
    P.rng = b.prng.as<uint64_t>();
    P.cap = cfg.pool;
    dparams base;
    std::memset(&base, 0, sizeof(base));
    base.sampler = RBW_SAMPLER_PER_SAMPLE_PCG;
    base.tri_test = cfg.tri_test;
    base.shadow_mode = cfg.shadow;
    const dscene S = b.scene.view();
    const uint32_t bs = cfg.block;
    // grid: every slot of the pool, at most 16 blocks of 256 threads per CU's worth
    const uint32_t grid = std::min(blocks(cfg.pool, bs), std::max(1u, (1u << 20) / bs));

    const uint32_t init_n = std::max(cfg.pool, nj);
    hipLaunchKernelGGL(k_b_init, dim3(blocks(init_n, 256)), dim3(256), 0, s, B, cfg.pool, total);
    HIP_CHECK(hipGetLastError());

    auto iteration = [&](int cur) {
        hipLaunchKernelGGL(k_b_plan, dim3(1), dim3(64), 0, s, B);
        hipLaunchKernelGGL(k_b_regen, dim3(grid), dim3(bs), 0, s, S, P, base, B, cur);
        hipLaunchKernelGGL(k_b_extend, dim3(grid), dim3(bs), 0, s, S, P, base, B, cur);
        hipLaunchKernelGGL(k_b_shade, dim3(grid), dim3(bs), 0, s, S, P, base, B, cur);
        hipLaunchKernelGGL(k_b_connect, dim3(grid), dim3(bs), 0, s, S, P, base, B);
        hipLaunchKernelGGL(k_b_retire, dim3(grid), dim3(bs), 0, s, P, B);
        hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, s, B, cur);
    };
    if (b.exec) {
        (void)hipGraphExecDestroy(b.exec);
        b.exec = nullptr;
    }
    if (b.graph) {
        (void)hipGraphDestroy(b.graph);
        b.graph = nullptr;
    }
    if (cfg.graph) {
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        iteration(0);
        iteration(1);
        const hipError_t ce = hipStreamEndCapture(s, &b.graph);
        HIP_CHECK(ce);
        HIP_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    }
    // a path lives at most max_depth bounces plus material-less pass-throughs; the
    // pool turns over at least once per (4 max_depth + 256) iterations
    const unsigned long long refills = (total + cfg.pool - 1) / cfg.pool;
    const unsigned long long max_launches = (refills + 2) * (4ull * max_depth + 256) / 2 + 16;
    hipEvent_t e0;
    if (b.events.empty()) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        b.events.push_back(e);
    }
    e0 = b.events[0];
    HIP_CHECK(hipEventRecord(e0, s));
    const uint32_t check_every = 4;
    uint32_t launches = 0;
    for (;;) {
        if (cfg.graph) HIP_CHECK(hipGraphLaunch(b.exec, s));
        else {
            iteration(0);
            iteration(1);
            HIP_CHECK(hipGetLastError());
        }
        launches++;
        if (b.events.size() <= launches) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            b.events.push_back(e);
        }
        HIP_CHECK(hipEventRecord(b.events[launches], s));
        if (launches % check_every == 0) {
            HIP_CHECK(hipMemcpyAsync(b.host_ctl, B.ctl, sizeof(dctl), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (b.host_ctl->cursor == total && b.host_ctl->free_top == cfg.pool) break;
            require(launches < max_launches, RBW_ERR_DEVICE, "batched bounce loop did not drain");
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    rr.launches = launches;
    rr.iterations = b.host_ctl->iteration;
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, b.events[launches]));
    rr.ms = ms;
    HIP_CHECK(hipMemcpy(rr.counters, B.counters, sizeof(rr.counters), hipMemcpyDeviceToHost));
    if (want_job_times) {
        rr.launch_ms.resize(launches + 1);
        rr.launch_ms[0] = 0.0;
        for (uint32_t l = 1; l <= launches; l++) {
            float t = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&t, e0, b.events[l]));
            rr.launch_ms[l] = t;
        }
        rr.first_iter.resize(nj);
        rr.last_iter.resize(nj);
        HIP_CHECK(hipMemcpy(rr.first_iter.data(), B.first_iter, size_t(nj) * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(rr.last_iter.data(), B.last_iter, size_t(nj) * 4, hipMemcpyDeviceToHost));
    }
}

// job records (films laid out one after another in b.films); returns the film bytes
std::vector<djob> make_jobs(rbw_ctx* ctx, const rbw_job* jobs, size_t n, const std::vector<uint32_t>& part_of_job,
                            uint32_t spp_override) {
    rbw_batch_state& b = *ctx->batch;
    std::vector<djob> out(n);
    size_t film_bytes = 0;
    std::vector<size_t> off(n);
    for (size_t i = 0; i < n; i++) {
        const rbw_camera_desc& c = jobs[i].camera;
        off[i] = film_bytes;
        film_bytes += size_t(c.width) * c.height * (3 * 8 + 4 + 4);
        film_bytes = (film_bytes + 255) & ~size_t(255);
    }
    b.films.ensure(film_bytes);
    HIP_CHECK(hipMemsetAsync(b.films.p, 0, film_bytes, ctx->stream));
    unsigned long long base = 0;
    for (size_t i = 0; i < n; i++) {
        const rbw_job& J = jobs[i];
        djob& d = out[i];
        std::memset(&d, 0, sizeof(d));
        d.cam = make_dcamera(&J.camera);
        const size_t np = size_t(J.camera.width) * J.camera.height;
        char* f = b.films.as<char>() + off[i];
        d.film.sum = reinterpret_cast<unsigned long long*>(f);
        d.film.weight = reinterpret_cast<uint32_t*>(f + np * 24);
        d.film.nan = reinterpret_cast<uint32_t*>(f + np * 28);
        const uint32_t p = part_of_job[i];
        d.root = b.roots[p];
        d.em_first = b.em_first[p];
        d.em_count = b.em_count[p];
        d.bw = uint32_t(J.camera.bound_max[0] - J.camera.bound_min[0]);
        d.npix = d.bw * uint32_t(J.camera.bound_max[1] - J.camera.bound_min[1]);
        d.spp = spp_override ? spp_override : J.spp;
        d.max_depth = J.max_depth;
        d.seed = J.seed;
        d.rr_threshold = J.rr_threshold;
        d.base = base;
        d.count = (unsigned long long)d.npix * d.spp;
        base += d.count;
    }
    return out;
}

}  // namespace

void autotune_jobs(rbw_ctx* ctx, const rbw_job* jobs, size_t n, rbw_batch_opts* inout, double* ms_out) {
    require(inout != nullptr, RBW_ERR_INVALID_ARGUMENT, "opts is null");
    validate_jobs(ctx, jobs, n);
    HIP_CHECK(hipSetDevice(ctx->device));
    if (!ctx->batch) ctx->batch = new rbw_batch_state();
    std::vector<uint32_t> part;
    prepare_scene(ctx, jobs, n, part);
    batch_config cfg = config_of(inout);
    const auto t0 = std::chrono::steady_clock::now();
    // probe: every job at 1 spp; coordinate search over block size, then pool size
    auto probe = [&](uint32_t block, uint32_t pool) {
        batch_config c = cfg;
        c.block = block;
        c.pool = pool;
        std::vector<djob> hj = make_jobs(ctx, jobs, n, part, 1);
        run_result rr;
        run_batch(ctx, hj, c, rr, false);
        return rr.ms;
    };
    uint32_t best_block = 256, best_pool = 1u << 21;
    double best = 1e300;
    for (uint32_t block : {64u, 128u, 256u, 512u}) {
        const double t = probe(block, best_pool);
        if (t < best) {
            best = t;
            best_block = block;
        }
    }
    for (uint32_t pool : {1u << 19, 1u << 20, 1u << 22}) {
        const double t = probe(best_block, pool);
        if (t < best) {
            best = t;
            best_pool = pool;
        }
    }
    inout->block_size = best_block;
    inout->pool_size = best_pool;
    inout->bvh_width = 2;  // the only width built
    if (ms_out)
        *ms_out = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void submit_jobs(rbw_ctx* ctx, const rbw_job* jobs, size_t n, const rbw_batch_opts* opts, rbw_job_stats* job_stats,
                 rbw_batch_stats* stats) {
    validate_jobs(ctx, jobs, n);
    HIP_CHECK(hipSetDevice(ctx->device));
    if (!ctx->batch) ctx->batch = new rbw_batch_state();
    rbw_batch_opts o;
    std::memset(&o, 0, sizeof(o));
    o.use_graph = 1;
    if (opts) o = *opts;
    double tune_ms = 0.0;
    if (o.autotune) autotune_jobs(ctx, jobs, n, &o, &tune_ms);
    const batch_config cfg = config_of(&o);
    std::vector<uint32_t> part;
    prepare_scene(ctx, jobs, n, part);
    rbw_batch_state& b = *ctx->batch;
    std::vector<djob> hj = make_jobs(ctx, jobs, n, part, 0);
    run_result rr;
    run_batch(ctx, hj, cfg, rr, job_stats != nullptr);
    // films -> caller
    hipStream_t s = ctx->stream;
    size_t max_px = 0;
    for (size_t i = 0; i < n; i++) max_px = std::max(max_px, size_t(jobs[i].camera.width) * jobs[i].camera.height);
    b.rgb.ensure(max_px * 3 * sizeof(float));
    for (size_t i = 0; i < n; i++) {
        if (!jobs[i].out_rgb) continue;
        const size_t np = size_t(jobs[i].camera.width) * jobs[i].camera.height;
        hipLaunchKernelGGL(k_b_resolve, dim3(blocks(np, 256)), dim3(256), 0, s, hj[i].cam, hj[i].film,
                           b.rgb.as<float>());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(jobs[i].out_rgb, b.rgb.p, np * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    unsigned long long samples = 0;
    for (const djob& J : hj) samples += J.count;
    if (job_stats) {
        for (size_t i = 0; i < n; i++) {
            rbw_job_stats& js = job_stats[i];
            std::memset(&js, 0, sizeof(js));
            js.samples = hj[i].count;
            js.first_iteration = rr.first_iter[i];
            js.last_iteration = rr.last_iter[i];
            // launch l (1-based) runs iterations 2l-1 and 2l
            const uint32_t l0 = (rr.first_iter[i] + 1) / 2, l1 = (rr.last_iter[i] + 1) / 2;
            if (l0 >= 1 && l1 >= l0 && l1 < rr.launch_ms.size()) {
                js.active_ms = rr.launch_ms[l1] - rr.launch_ms[l0 - 1];
                js.msamples_per_s = js.active_ms > 0 ? double(js.samples) / js.active_ms * 1e-3 : 0.0;
            }
        }
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->render_ms = rr.ms;
        stats->samples = samples;
        stats->msamples_per_s = rr.ms > 0 ? double(samples) / rr.ms * 1e-3 : 0.0;
        stats->extension_rays = rr.counters[0];
        stats->shadow_rays = rr.counters[1];
        stats->mis_rays = rr.counters[2];
        stats->mrays_per_s =
            rr.ms > 0 ? double(rr.counters[0] + rr.counters[1] + rr.counters[2]) / rr.ms * 1e-3 : 0.0;
        stats->autotune_ms = tune_ms;
        stats->iterations = rr.iterations;
        stats->graph_launches = cfg.graph ? rr.launches : 0;
        stats->block_size = cfg.block;
        stats->pool_size = cfg.pool;
        stats->bvh_width = cfg.width;
        stats->scene_triangles = b.merged_tris;
    }
}

}  // namespace host
}  // namespace rbw
