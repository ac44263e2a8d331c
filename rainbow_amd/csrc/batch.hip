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
// fixed at capture (one thread per pool slot); each kernel reads its queue
// length from device memory and blocks past it leave at once.
//
// A job's image is bit-identical to rbw_render's: the same per-sample pcg32
// stream, the same BVH (built by the same builder from the same scene), and
// order-independent 64-bit fixed-point film sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "guided.hpp"
#include "guiding.hpp"
#include "host.hpp"
#include "path.hpp"

using namespace rbw;
using namespace rbw::host;

namespace {

// largest block size of the wavefront kernels (the autotuner's candidates are 64..512);
// bounding it lets the register allocator give each lane more than 128 VGPRs
constexpr int k_batch_max_block = 512;
// regeneration hands out samples per group of this many consecutive slots (one
// block of the regeneration kernel); block sizes of the other kernels divide it
constexpr uint32_t k_group = 512;

struct djob {
    dcamera cam;
    dfilm film;
    unsigned long long base;   // first index of the job in the batch's sample sequence
    unsigned long long count;  // npix * spp
    uint32_t root, em_first, em_count, spp;
    uint32_t max_depth, seed, npix, bw;
    float rr_threshold;
    uint32_t root8;
    uint32_t light_flags;  // k_lights_* of the job's scene
    uint32_t root4;
    f3 gmin, ginv;         // guided jobs: the network query's position normalisation (scene box)
};

struct dctl {
    unsigned long long cursor;  // next sample of the job queue
    unsigned long long total;   // samples of all jobs
    unsigned long long live;    // slots holding a path after this iteration's regeneration
    uint32_t iteration;
    // 1: the queue is exhausted and no slot holds a path after this iteration's regeneration
    // (the stage kernels return at once); 2: that regeneration has run too (it returns too)
    uint32_t done;
};

struct dbatch {
    dctl* ctl;
    const djob* jobs;
    uint32_t njobs;
    uint32_t pool;
    uint32_t ngroups;               // groups of k_group consecutive slots
    uint32_t* block_free;           // per group: slots free at the next regeneration
    uint32_t* block_n;              // per group: camera samples granted this iteration
    unsigned long long* block_base; // per group: first granted sample
    uint32_t* wave_job;             // per wave of slots: a job that retired a path in the last regeneration
    uint32_t nwaves;
    uint32_t* first_iter;           // per job
    uint32_t* last_iter;            // per job
    unsigned long long* counters;   // extension, shadow, MIS, light casts (striped)
    uint32_t* fetch;                // next item of the persistent extension [0] and occlusion [1] walks, MIS queue
                                    // length [2] (zeroed by the plan)
    uint32_t* mqueue;               // slots whose MIS ray reaches an emitter triangle (queue_mis)
    // guided jobs (job j uses network j): this iteration's guided vertices per job, as slots
    // (rows[j * pool + k], k < gcount[j]), and the training rows closed at retirement
    uint32_t guided, train;
    uint32_t* grows;
    uint32_t* gcount;
    float* t_in;                    // [j * t_cap + k] x 9
    float* t_dir;                   // x 3
    float* t_c;
    uint32_t* t_count;              // per job (rows offered; at most t_cap kept)
    uint32_t t_cap;
};

// guided jobs: floats of a guided vertex's surface record (by slot): point, normal, wo, the
// shading frame, the material's bits, one pad
constexpr uint32_t G_REC = 20;
static_assert(G_COUNT + 2 == G_REC, "record layout");

// slot state (U_STATE): 0 = free
constexpr uint32_t ST_ALIVE = 1u;    // trace an extension ray this iteration
constexpr uint32_t ST_CONNECT = 2u;  // next-event estimation rays pending
constexpr uint32_t ST_DYING = 4u;    // finished: into the film at the next regeneration
constexpr uint32_t ST_GUIDE = 8u;    // a guided vertex waits for its network outputs

// appends one entry per lane with `take` to its job's list (counts[job], slots of list job at
// list + job * stride): lanes of one job in a wave share one atomic; returns the lane's position
__device__ __forceinline__ uint32_t append_by_job(bool take, uint32_t job, uint32_t* counts) {
    uint64_t pending = __ballot(take);
    uint32_t pos = 0;
    const uint64_t below = __lane_id() == 0 ? 0ull : (~0ull >> (64 - __lane_id()));
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t jl = __shfl(job, leader);
        const uint64_t same = __ballot(take && job == jl);
        uint32_t base = 0;
        if ((int)__lane_id() == leader) base = atomicAdd(&counts[jl], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if (take && job == jl) pos = base + (uint32_t)__popcll(same & below);
        pending &= ~same;
    }
    return pos;
}

// reserves n consecutive entries of its job's list for every lane (n may be 0), one atomic
// per job and wave; returns the lane's first position.  Call with the whole wave converged.
__device__ __forceinline__ uint32_t append_by_job_n(uint32_t n, uint32_t job, uint32_t* counts) {
    uint64_t pending = __ballot(n > 0);
    uint32_t pos = 0;
    const int lane = (int)__lane_id();
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t jl = __shfl(job, leader);
        const bool mine = n > 0 && job == jl;
        uint32_t x = mine ? n : 0u;  // inclusive scan of the job's counts over the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)x, (unsigned)off);
            if (lane >= off) x += v;
        }
        const uint32_t total = (uint32_t)__shfl((int)x, 63);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counts[jl], total);
        base = (uint32_t)__shfl((int)base, leader);
        if (mine) pos = base + x - n;
        pending &= ~__ballot(mine);
    }
    return pos;
}

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
    Sj.root4 = J.root4;
    Sj.root8 = J.root8;
    Sj.emitters = S.emitters + J.em_first;
    Sj.num_emitters = J.em_count;
    Sj.light_flags = J.light_flags;
    pj = base;
    pj.spp = J.spp;
    pj.max_depth = J.max_depth;
    pj.rr_threshold = J.rr_threshold;
    pj.seed = J.seed;
}

// number of lanes of the block with `pred`, and the rank of this lane among them
// (slot order); every thread of the block must call it
__device__ __forceinline__ uint32_t block_rank(bool pred, uint32_t& total, uint32_t* lds) {
    const uint64_t m = __ballot(pred);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (lane == 0) lds[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < nwaves; w++) {
        const uint32_t c = lds[w];
        off += w < wave ? c : 0u;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + (uint32_t)__popcll(m & below);
}

__global__ void k_b_init(dbatch B, dpaths P, unsigned long long total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B.pool) P.U(U_STATE, i) = 0u;
    if (i < B.ngroups) {
        const uint32_t first = i * k_group;
        B.block_free[i] = B.pool - first < k_group ? B.pool - first : k_group;
    }
    if (i < B.nwaves) B.wave_job[i] = RBW_NONE;
    if (i < B.njobs) {
        B.first_iter[i] = 0xFFFFFFFFu;
        B.last_iter[i] = 0u;
    }
    if (i < RBW_STRIPES * RBW_STRIPE_WORDS) B.counters[i] = 0;
    if (i == 0) {
        dctl c;
        c.cursor = 0;
        c.total = total;
        c.live = 0;
        c.iteration = 0;
        c.done = 0;
        *B.ctl = c;
    }
}

// Folds the last regeneration's per-wave retire records into the jobs' last
// iterations (plain stores: every record of one regeneration carries the same
// iteration), with the threads of one block: four records per 16-byte load, the loads
// of a thread's next groups in flight together, and a job stored once per run of
// records that name it (a render's records mostly name one job).
__device__ __forceinline__ void fold_retire_records(const dbatch& B, uint32_t iter) {
    uint32_t last = RBW_NONE;
    auto fold = [&](uint32_t j) {
        if (j != RBW_NONE && j != last) {
            B.last_iter[j] = iter;
            last = j;
        }
    };
    uint4* rec = reinterpret_cast<uint4*>(B.wave_job);  // device allocations are 256-byte aligned
    const uint32_t n4 = B.nwaves / 4;
    const uint4 none = make_uint4(RBW_NONE, RBW_NONE, RBW_NONE, RBW_NONE);
#pragma unroll 4
    for (uint32_t q = threadIdx.x; q < n4; q += blockDim.x) {
        const uint4 w = rec[q];
        if ((w.x & w.y & w.z & w.w) == RBW_NONE) continue;  // RBW_NONE is all ones
        fold(w.x);
        fold(w.y);
        fold(w.z);
        fold(w.w);
        rec[q] = none;
    }
    for (uint32_t w = 4 * n4 + threadIdx.x; w < B.nwaves; w += blockDim.x) {
        fold(B.wave_job[w]);
        B.wave_job[w] = RBW_NONE;
    }
}

__global__ __launch_bounds__(1024) void k_b_fold(dbatch B) { fold_retire_records(B, B.ctl->iteration); }

// Hands out the next camera samples of the job queue: block b's free slots get
// the range [block_base[b], block_base[b] + block_n[b]), blocks in order.  One
// block of 1024 threads scans the per-block free counts.
__global__ __launch_bounds__(1024) void k_b_plan(dbatch B) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x, n = B.ngroups;
    if (t < 3) B.fetch[t] = 0u;
    if (B.guided)
        for (uint32_t j = t; j < B.njobs; j += 1024) B.gcount[j] = 0u;
    // the work is done: the iterations the host replays before it sees that cost only their
    // launches (the last regeneration's retire records are folded after the loop)
    if (B.ctl->done) {
        __syncthreads();
        if (t == 0) B.ctl->done = 2u;
        return;
    }
    fold_retire_records(B, B.ctl->iteration);
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b0 = t * per < n ? t * per : n, b1 = b0 + per < n ? b0 + per : n;
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += B.block_free[b];
    part[t] = sum;
    __syncthreads();
    // inclusive scan of the 1024 partial sums (Hillis-Steele)
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const uint32_t total_free = part[1023];
    const unsigned long long cursor = B.ctl->cursor, total = B.ctl->total;
    const unsigned long long left = total > cursor ? total - cursor : 0ull;
    uint32_t prefix = part[t] - sum;
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t f = B.block_free[b];
        B.block_free[b] = 0u;  // the connect kernel counts the next iteration's free slots
        const unsigned long long avail = left > prefix ? left - prefix : 0ull;
        B.block_base[b] = cursor + prefix;
        B.block_n[b] = avail < f ? (uint32_t)avail : f;
        prefix += f;
    }
    __syncthreads();
    if (t == 0) {
        const unsigned long long granted = left < total_free ? left : total_free;
        B.ctl->cursor = cursor + granted;
        B.ctl->live = (unsigned long long)B.pool - (total_free - granted);
        B.ctl->iteration += 1;
        if (B.ctl->live == 0 && cursor + granted >= total) B.ctl->done = 1u;
    }
}

// Retires finished paths into their job's film, then regenerates the free slots
// in place with the block's granted samples, in slot order, so consecutive slots
// get consecutive samples (neighbouring pixels of one job).
__global__ __launch_bounds__(k_batch_max_block) void k_b_regen(dscene S, dpaths P, dparams base, dbatch B) {
    __shared__ uint32_t lds[k_batch_max_block / 64];
    if (B.ctl->done == 2u) return;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = slot < B.pool;
    const uint32_t iter = B.ctl->iteration;
    uint32_t st = in ? P.U(U_STATE, slot) : ST_ALIVE;
    const bool dying = (st & ST_DYING) != 0;
    uint32_t j = 0xFFFFFFFFu;
    if (dying) {
        j = P.U(U_JOB, slot);
        const djob& J = B.jobs[j];
        const uint32_t pixel = P.U(U_PIXEL, slot);
        film_add_sample(J.cam, J.film, (int32_t)(pixel % J.cam.width), (int32_t)(pixel / J.cam.width),
                        P.F(F_PX, slot), P.F(F_PY, slot), P.F3(F_VX, slot), false);
        st = 0u;
    }
    if (B.train) {
        // the finished path's closed training records into its job's training rows: count
        // them, reserve the wave's rows of each job at once, then write them
        const dguide& G = base.g;
        uint32_t n = 0;
        if (dying)
            for (uint32_t b = 0; b < G.rec_depth; b++) {
                const size_t ri = (size_t)slot * G.rec_depth + b;
                if (G.rec_flag[ri] == 2 && record_target(P, G, slot, ri) > 0.0f) n++;
            }
        uint32_t k = append_by_job_n(n, j, B.t_count);
        if (dying)
            for (uint32_t b = 0; b < G.rec_depth; b++) {
                const size_t ri = (size_t)slot * G.rec_depth + b;
                const float c = G.rec_flag[ri] == 2 ? record_target(P, G, slot, ri) : 0.0f;
                if (c > 0.0f) {
                    if (k < B.t_cap) record_to_row(P, G, ri, c, B.t_in, B.t_dir, B.t_c, (size_t)j * B.t_cap + k);
                    k++;
                }
                G.rec_flag[ri] = 0;
            }
    }
    // the iteration in which a job's last path retired: the wave records its first
    // retiring job (folded by the next plan); each other job with retiring lanes in the
    // wave takes one atomic from its first lane.  (Slots regenerate in place, so a wave
    // of a shared launch holds paths of several jobs; an atomic per lane on the jobs'
    // few counters serialised the kernel.)
    {
        uint64_t pending = __ballot(dying);
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t jl = __shfl(j, leader < 0 ? 0 : leader);
        if (leader >= 0 && (int)__lane_id() == leader) B.wave_job[slot >> 6] = jl;
        pending &= ~__ballot(dying && j == jl);
        while (pending) {
            const int first = __ffsll((unsigned long long)pending) - 1;
            const uint32_t jo = __shfl(j, first);
            if ((int)__lane_id() == first) atomicMax(&B.last_iter[jo], iter);
            pending &= ~__ballot(dying && j == jo);
        }
    }
    uint32_t nfree = 0;
    const bool free_slot = in && st == 0u;
    const uint32_t r = block_rank(free_slot, nfree, lds);
    const bool gen = free_slot && r < B.block_n[blockIdx.x];
    if (gen) {
        const unsigned long long g = B.block_base[blockIdx.x] + r;
        const uint32_t jn = find_job(B.jobs, B.njobs, g);
        const djob& J = B.jobs[jn];
        const unsigned long long l = g - J.base;
        const uint32_t s = (uint32_t)(l / J.npix), li = (uint32_t)(l % J.npix);
        dscene Sj;
        dparams pj;
        job_view(S, base, J, Sj, pj);
        raygen_path(J.cam, pj, P, slot, J.cam.bx0 + (int32_t)(li % J.bw), J.cam.by0 + (int32_t)(li / J.bw), s);
        P.U(U_JOB, slot) = jn;
        if (l == 0) B.first_iter[jn] = iter;
        st = ST_ALIVE;
    }
    if (in && (gen || dying)) P.U(U_STATE, slot) = st;
}

// The extension walks of every live slot on the four-wide BVH, persistent: each wave takes
// ranges of k_fetch_chunk consecutive slots from one counter and hands them to its lanes as
// their walks end, so the lanes keep walking instead of waiting for the wave's longest walk
// (one slot per lane for the whole launch kept ~20 % of the lanes busy).  A slot's walk is
// traverse4's, so its hit is the same.  One wave per block; launch at most the resident
// waves (more are harmless: they find the counter past the pool and leave).
constexpr uint32_t k_fetch_chunk = 256;  // items per range (1024, 128, 64 measured slower)
constexpr uint32_t k_refill = 32;        // idle lanes that make a wave refill (8, 16: the same)
// CONNECT: the walks are the occlusion queries of the slots waiting for the connect stage
// (ST_CONNECT), connect_path's FAST forms: items [0, n) are the MIS rays that reach a
// nearest visible emitter triangle (the n slots of the MIS queue, filled by shading;
// U_MPRIM, F_MLEN), walked up to it with its entity the target, items [n, n + pool) the
// slots' shadow rays (C_SHADOW; the light's entity is the target).  A shadow ray with
// nothing in the way sets C_LIT in the slot's connect flags, an MIS ray with something in
// the way C_MOCC; the connect kernel then skips the walks.  Counter B.fetch[1].  Otherwise
// the extension walks of the live slots (closest hits), counter B.fetch[0].
template <bool CONNECT>
__device__ __forceinline__ void persistent_walk(const dscene& S, const dpaths& P, const dparams& base,
                                                const dbatch& B) {
    const uint32_t lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t tri_test = base.tri_test;
    const uint32_t nmis = CONNECT ? B.fetch[2] : 0u;
    const uint32_t items = nmis + B.pool;
    uint32_t rb = 0, re = 0;  // the wave's range of items not yet handed out
    bool exhausted = false;
    bool active = false;      // the lane walks the ray of `slot`
    bool mis = false;         // CONNECT: the lane's ray is the MIS ray of `slot`
    uint32_t slot = 0, rays = 0;
    // the walk (traverse4's state)
    local_ray lr;
    wt_setup wt;
    trav_state st;
    f3 o_w = f3{0.0f, 0.0f, 0.0f}, inv = o_w, oi = o_w;
    uint32_t stack[RBW_STACK];
    int sp = 0;
    int32_t cur = 0, leaf = 0;
    bool done = true, has_leaf = false;
    uint32_t target = RBW_NONE;  // CONNECT: the entity the ray may hit
    uint32_t* ring = rbw_lds_stack + threadIdx.x;
    auto push = [&](uint32_t v) {
        uint32_t* e = ring + (uint32_t)(sp & (RBW_LDS_STACK - 1)) * 64u;
        if (sp >= RBW_LDS_STACK) stack[sp - RBW_LDS_STACK] = *e;
        *e = v;
        sp++;
    };
    auto pop = [&]() -> uint32_t {
        sp--;
        uint32_t* e = ring + (uint32_t)(sp & (RBW_LDS_STACK - 1)) * 64u;
        const uint32_t v = *e;
        if (sp >= RBW_LDS_STACK) *e = stack[sp - RBW_LDS_STACK];
        return v;
    };
    auto advance = [&]() {
        if (sp == 0) done = true;
        else cur = (int32_t)pop();
    };
    // the walk of the ray (o, d) up to world length len_w (the caller sets `cur`, the root)
    auto start = [&](f3 o, f3 d, float len_w) {
        lr = make_local_ray(o, d);
        if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
        o_w = o;
        inv = lr.inv;
        oi = f3{lr.o.x * inv.x, lr.o.y * inv.y, lr.o.z * inv.z};
        st.best.prim = RBW_MISS;
        st.best.t = 0.0f;
        st.best.b1 = st.best.b2 = 0.0f;
        st.best_inst_entity = RBW_NONE;
        st.len_w = len_w;
        st.thr_local = st.len_w * lr.L0;
        sp = 0;
        done = false;
        has_leaf = false;
        active = true;
    };
    trace_counters cnt{0, 0};
    for (;;) {
        // idle lanes take the next items; an item without a ray leaves its lane idle
        uint64_t idle = __ballot(!active);
        if (!exhausted && (idle == ~0ull || (uint32_t)__popcll(idle) >= k_refill)) {
            while (idle && !exhausted) {
                if (rb == re) {
                    uint32_t g = 0;
                    if (lane == 0) g = atomicAdd(B.fetch + (CONNECT ? 1 : 0), k_fetch_chunk);
                    g = (uint32_t)__shfl((int)g, 0);
                    if (g >= items) {
                        exhausted = true;
                        break;
                    }
                    rb = g;
                    re = g + k_fetch_chunk < items ? g + k_fetch_chunk : items;
                }
                const uint32_t take = min((uint32_t)__popcll(idle), re - rb);
                const uint32_t rank = (uint32_t)__popcll(idle & below);
                if (!active && rank < take) {
                    const uint32_t item = rb + rank;
                    if constexpr (CONNECT) {
                        mis = item < nmis;
                        if (mis) {
                            slot = B.mqueue[item];
                            f3 eo, ed;
                            spawn_ray(P.F3(F_MPX, slot), P.F3(F_MNX, slot), P.F3(F_MWX, slot), eo, ed);
                            start(eo, ed, P.F(F_MLEN, slot));
                            target = S.tri_entity[P.U(U_MPRIM, slot)];
                            cur = (int32_t)B.jobs[P.U(U_JOB, slot)].root4;
                        } else {
                            slot = item - nmis;
                            if ((P.U(U_STATE, slot) & ST_CONNECT) && (P.U(U_CFLAGS, slot) & C_SHADOW)) {
                                start(P.F3(F_SOX, slot), P.F3(F_SDX, slot), P.F(F_SLEN, slot));
                                target = P.U(U_SENT, slot);
                                cur = (int32_t)B.jobs[P.U(U_JOB, slot)].root4;
                            }
                        }
                    } else {
                        slot = item;
                        if (P.U(U_STATE, slot) & ST_ALIVE) {
                            const f3 o = P.F3(F_OX, slot), d = P.F3(F_DX, slot);
                            lr = make_local_ray(o, d);
                            if (tri_test == RBW_TRI_WATERTIGHT) wt = make_wt(lr.d);
                            o_w = o;
                            inv = lr.inv;
                            oi = f3{lr.o.x * inv.x, lr.o.y * inv.y, lr.o.z * inv.z};
                            st.best.prim = RBW_MISS;
                            st.best.t = 0.0f;
                            st.best.b1 = st.best.b2 = 0.0f;
                            st.best_inst_entity = RBW_NONE;
                            st.len_w = __builtin_inff();
                            st.thr_local = st.len_w * lr.L0;
                            cur = (int32_t)B.jobs[P.U(U_JOB, slot)].root4;
                            sp = 0;
                            done = false;
                            has_leaf = false;
                            active = true;
                            rays++;
                        }
                    }
                }
                rb += take;
                idle = __ballot(!active);
            }
        }
        if (__ballot(active) == 0) break;  // exhausted, every walk finished
        // walk inner nodes until every walking lane holds a leaf or is done (traverse4)
        for (;;) {
            if (active && !done && !has_leaf && cur < 0) {
                if (cur != k_empty_child) {
                    leaf = cur;
                    has_leaf = true;
                }
                advance();
            }
            const bool walk = active && !done && cur >= 0;
            if (__ballot(walk) == 0) break;
            if (walk && !visit4(S, cur, inv, oi, st.thr_local * 1.0000002f, push)) advance();
            if (__ballot(active && !(has_leaf || done)) == 0) break;
        }
        while (has_leaf) {
            has_leaf = false;
            const uint32_t first = (uint32_t)leaf & 0x07FFFFFFu;
            const uint32_t count = (((uint32_t)leaf >> 27) & 15u) + 1u;
            if constexpr (CONNECT) {
                if (test_leaf<true, true, false>(S, S.prim_a, S.prim_b, S.prim_c, first, count, lr, o_w, wt,
                                                 tri_test, target, cnt, st)) {
                    done = true;  // occluded
                    break;
                }
            } else {
                test_leaf<false, false, false>(S, S.prim_a, S.prim_b, S.prim_c, first, count, lr, o_w, wt, tri_test,
                                               RBW_NONE, cnt, st);
            }
            if (!done && cur < 0) {
                if (cur != k_empty_child) {
                    leaf = cur;
                    has_leaf = true;
                }
                advance();
            }
        }
        if (active && done) {
            if constexpr (CONNECT) {
                // (the slot's other item may be finishing at the same time)
                const bool open = st.best.prim == RBW_MISS;
                if (open != mis) atomicOr(&P.U(U_CFLAGS, slot), mis ? C_MOCC : C_LIT);
            } else {
                P.U(U_HPRIM, slot) = st.best.prim;
                P.F(F_HT, slot) = st.best.t;
            }
            active = false;
        }
    }
    if (!CONNECT) stripe_add64(B.counters, 0, rays);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RBW_EXTEND4_WAVES)))
void k_b_extend_dyn(dscene S, dpaths P, dparams base, dbatch B) {
    if (B.ctl->done) return;
    persistent_walk<false>(S, P, base, B);
}

// the occlusion walks of the connect stage, persistent (four-wide BVH, every job's scene with
// few planar lights, no debug hits)
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RBW_EXTEND4_WAVES)))
void k_b_occlusion_dyn(dscene S, dpaths P, dparams base, dbatch B) {
    if (B.ctl->done) return;
    persistent_walk<true>(S, P, base, B);
}

template <int W>
__global__ __launch_bounds__(k_batch_max_block) RBW_EXTEND_OCCUPANCY void k_b_extend(dscene S, dpaths P, dparams base, dbatch B) {
    if (B.ctl->done) return;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    trace_counters cnt{0, 0};
    uint32_t rays = 0;
    if (slot < B.pool && (P.U(U_STATE, slot) & ST_ALIVE)) {
        dscene Sj;
        dparams pj;
        job_view(S, base, B.jobs[P.U(U_JOB, slot)], Sj, pj);
        extend_path<W, false, true>(Sj, P, pj, slot, cnt);
        rays = 1;
    }
    stripe_add64(B.counters, 0, rays);
}

// the slots whose MIS ray reaches an emitter triangle join the MIS queue (one atomic per wave)
__device__ __forceinline__ void queue_mis(const dbatch& B, bool reach, uint32_t slot) {
    const uint64_t m = __ballot(reach);
    if (m == 0) return;
    const uint32_t lane = __lane_id();
    const int first = __ffsll((unsigned long long)m) - 1;
    uint32_t at = 0;
    if ((int)lane == first) at = atomicAdd(B.fetch + 2, (uint32_t)__popcll(m));
    at = (uint32_t)__shfl((int)at, first);
    if (reach) B.mqueue[at + (uint32_t)__popcll(m & (lane == 0 ? 0ull : (~0ull >> (64 - lane))))] = slot;
}

// The vertex's material kind, for grouping vertices by the code they run: 0 no surface
// (missed, invisible entity), else 1 + the material type (matte, metal, plastic).  A GGX
// lobe's evaluation and samples run several times the instructions of a Lambert lobe, and
// a wave runs every kind any of its lanes needs.
__device__ __forceinline__ uint32_t shade_kind(const dscene& S, uint32_t prim) {
    if (prim == RBW_MISS) return 0u;
    uint32_t mid;
    if (prim < S.num_tris) {
        mid = S.tri_material ? S.tri_material[prim] : RBW_NONE;
        if (mid == RBW_NONE) mid = S.ents[S.tri_entity[prim]].material;
    } else {
        mid = S.ents[S.sphere_entity[prim - S.num_tris]].material;
    }
    if (mid == RBW_NONE) return 0u;
    const uint32_t t = S.mats[mid].type;
    return t < 1u ? 1u : (t > 3u ? 3u : t);
}

// The slot whose vertex this thread shades when each block of k_sort_window threads takes
// its window's live slots grouped by material kind (a stable counting sort of the slot
// indices in LDS), so most waves run one kind's code.  A slot's shading does not depend on
// the thread that runs it.  Every thread of the block calls it.
// (measured on the icosphere workload, shading per render: unsorted 51.6 ms; windows of
// 128 / 256 / 512 / 1024 slots 51.3 / 49.8 / 55.5 / 60.3 ms: wider windows scatter each
// wave's path-state accesses over more cache lines)
constexpr uint32_t k_sort_window = 256;
__device__ __forceinline__ uint32_t grouped_slot(const dscene& S, const dpaths& P, const dbatch& B,
                                                 uint32_t (*counts)[k_sort_window / 64], uint32_t* order) {
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t own = blockIdx.x * k_sort_window + t;
    uint32_t kind = 0;
    if (own < B.pool && (P.U(U_STATE, own) & ST_ALIVE)) kind = shade_kind(S, P.U(U_HPRIM, own));
    uint32_t rank = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4; c++) {
        const uint64_t m = __ballot(kind == c);
        if (lane == 0) counts[c][wave] = (uint32_t)__popcll(m);
        if (kind == c) rank = (uint32_t)__popcll(m & below);
    }
    __syncthreads();
    uint32_t at = rank;
    for (uint32_t c = 0; c < kind; c++)
        for (uint32_t w = 0; w < k_sort_window / 64; w++) at += counts[c][w];
    for (uint32_t w = 0; w < wave; w++) at += counts[kind][w];
    order[at] = own;
    __syncthreads();
    return order[t];
}

// The vertices of the live slots (shade_path), grouped by material kind (grouped_slot).
// MIS_REACH: the connect stage's occlusion walks run in k_b_occlusion_dyn, and the MIS ray
// of a vertex is tested against the emitter triangles here (connect_path's FAST form):
// U_MPRIM (RBW_MISS: none) and F_MLEN, the ray's world length up to it, and the slots with
// one join the MIS queue.  The test reads the ray back from the path state after shading
// (the lines are in L2): inside shade_path it made the register allocator spill.
template <bool MIS_REACH>
__global__ __launch_bounds__(k_sort_window) RBW_SHADE_OCCUPANCY void k_b_shade_sorted(dscene S, dpaths P, dparams base,
                                                                                      dbatch B) {
    __shared__ uint32_t counts[4][k_sort_window / 64];
    __shared__ uint32_t order[k_sort_window];
    if (B.ctl->done) return;
    const uint32_t slot = grouped_slot(S, P, B, counts, order);
    bool test = false;
    if (slot < B.pool && (P.U(U_STATE, slot) & ST_ALIVE)) {
        bool alive = false, need_connect = false, to_guide = false;
        uint32_t gmat = 0;
        surf gsi;
        dscene Sj;
        dparams pj;
        job_view(S, base, B.jobs[P.U(U_JOB, slot)], Sj, pj);
        shade_path<false>(Sj, P, pj, slot, alive, need_connect, to_guide, gsi, gmat);
        P.U(U_STATE, slot) = (alive ? ST_ALIVE : ST_DYING) | (need_connect ? ST_CONNECT : 0u);
        test = MIS_REACH && need_connect;
    }
    if (MIS_REACH) {
        // the slot index and the job's emitter range are read again (LDS, L2): held across
        // shading they spilled
        __asm__ volatile("" ::: "memory");
        const uint32_t ms = order[threadIdx.x];
        bool reach = false;
        if (test && (P.U(U_CFLAGS, ms) & C_MIS)) {
            const djob& J = B.jobs[P.U(U_JOB, ms)];
            f3 eo, ed;
            spawn_ray(P.F3(F_MPX, ms), P.F3(F_MNX, ms), P.F3(F_MWX, ms), eo, ed);
            const hitrec h = nearest_emitter_hit(S, S.emitters + J.em_first, J.em_count, eo, ed, base.tri_test);
            P.U(U_MPRIM, ms) = h.prim;
            if (h.prim != RBW_MISS) {
                P.F(F_MLEN, ms) = h.t * make_local_ray(eo, ed).L1;
                reach = true;
            }
        }
        queue_mis(B, reach, ms);
    }
}

// guided jobs: the vertex up to the guided stage.  A guided vertex stores its surface
// record, material and network query by slot and joins its job's list of guided slots;
// the others shade as the plain estimator (shade_path).  (Grouping the slots by material
// kind, as k_b_shade_sorted does, left the 16 guided rooms at 1024 spp unchanged: 29.9 s
// against 29.8 s.)
__global__ __launch_bounds__(k_batch_max_block) RBW_SHADE_OCCUPANCY void k_b_shade_guided(dscene S, dpaths P,
                                                                                         dparams base, dbatch B) {
    if (B.ctl->done) return;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool to_guide = false;
    uint32_t j = 0;
    if (slot < B.pool && (P.U(U_STATE, slot) & ST_ALIVE)) {
        bool alive = false, need_connect = false;
        uint32_t gmat = 0;
        surf gsi;
        dscene Sj;
        dparams pj;
        j = P.U(U_JOB, slot);
        const djob& J = B.jobs[j];
        job_view(S, base, J, Sj, pj);
        shade_path<true>(Sj, P, pj, slot, alive, need_connect, to_guide, gsi, gmat);
        if (to_guide) {
            const dguide& G = base.g;
            // the surface record by slot, 80 contiguous bytes (k_b_guided reads the scattered
            // guided slots back: five 16-byte loads each instead of 19 gathers)
            float4* r = reinterpret_cast<float4*>(G.surf + (size_t)slot * G_REC);
            r[0] = make_float4(gsi.point.x, gsi.point.y, gsi.point.z, gsi.normal.x);
            r[1] = make_float4(gsi.normal.y, gsi.normal.z, gsi.wo.x, gsi.wo.y);
            r[2] = make_float4(gsi.wo.z, gsi.sh.x.x, gsi.sh.x.y, gsi.sh.x.z);
            r[3] = make_float4(gsi.sh.y.x, gsi.sh.y.y, gsi.sh.y.z, gsi.sh.z.x);
            r[4] = make_float4(gsi.sh.z.y, gsi.sh.z.z, __uint_as_float(gmat), 0.0f);
            float q[9];
            guide_query(gsi, J.gmin, J.ginv, q);
#pragma unroll
            for (int k = 0; k < 9; k++) G.query[(size_t)slot * 9 + k] = q[k];
            P.U(U_STATE, slot) = ST_GUIDE;
        } else {
            P.U(U_STATE, slot) = (alive ? ST_ALIVE : ST_DYING) | (need_connect ? ST_CONNECT : 0u);
        }
    }
    const uint32_t k = append_by_job(to_guide, j, B.gcount);
    if (to_guide) B.grows[(size_t)j * B.pool + k] = slot;
}

// the guided vertex of the slot's path with its job's network outputs
__device__ __forceinline__ void guided_slot(const dscene& S, const dpaths& P, const dparams& base, const dbatch& B,
                                            uint32_t slot) {
    const dguide& G = base.g;
    const float4* r = reinterpret_cast<const float4*>(G.surf + (size_t)slot * G_REC);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4];
    surf si;
    si.point = f3{r0.x, r0.y, r0.z};
    si.normal = f3{r0.w, r1.x, r1.y};
    si.wo = f3{r1.z, r1.w, r2.x};
    si.sh.x = f3{r2.y, r2.z, r2.w};
    si.sh.y = f3{r3.x, r3.y, r3.z};
    si.sh.z = f3{r3.w, r4.x, r4.y};
    const uint32_t mat = __float_as_uint(r4.z);
    dscene Sj;
    dparams pj;
    job_view(S, base, B.jobs[P.U(U_JOB, slot)], Sj, pj);
    bool alive = false, need_connect = false;
    guided_vertex(Sj, P, pj, slot, si, mat, G.mix + (size_t)slot * guide::N_OUT_PAD,
                  G.query + (size_t)slot * 9, 0u, alive, need_connect);
    P.U(U_STATE, slot) = (alive ? ST_ALIVE : ST_DYING) | (need_connect ? ST_CONNECT : 0u);
}

// guided jobs: the guided vertex (guided.hpp) of every slot that waits for it, with its job's
// network outputs.  (Grouping each block's rows by material kind, as the plain shading
// kernel does, left the 16 guided rooms at 1024 spp unchanged, 29.8 s.)  The threads walk the jobs' lists of guided slots (rows of job 0, then
// job 1, ...) rather than the pool: the guided vertices of an iteration are a scattered
// minority of the slots (with the first vertex guided, the paths regenerated in it), so a
// walk over the pool would run each wave's few of them with most lanes idle.
__global__ __launch_bounds__(k_batch_max_block) RBW_SHADE_OCCUPANCY void k_b_guided(dscene S, dpaths P, dparams base,
                                                                                   dbatch B) {
    __shared__ uint32_t first_row[guide::MAX_NETS + 1];
    if (B.ctl->done) return;
    if (threadIdx.x == 0) {
        uint32_t a = 0;
        for (uint32_t j = 0; j < B.njobs; j++) {
            first_row[j] = a;
            const uint32_t c = B.gcount[j];
            a += c < B.pool ? c : B.pool;
        }
        first_row[B.njobs] = a;
    }
    __syncthreads();
    const uint32_t rows = first_row[B.njobs];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        uint32_t j = 0;
        while (first_row[j + 1] <= r) j++;
        guided_slot(S, P, base, B, B.grows[(size_t)j * B.pool + (r - first_row[j])]);
    }
}

// next-event estimation rays; then the block's count of slots the next
// regeneration will find free (free or finished)
// FAST: every job's scene has few planar lights (connect_path's occlusion-only form);
// WALKED: k_b_occlusion_dyn ran the occlusion walks
template <int W, bool FAST, bool WALKED = false>
__global__ __launch_bounds__(k_batch_max_block) RBW_CONNECT_OCCUPANCY void k_b_connect(dscene S, dpaths P, dparams base, dbatch B) {
    __shared__ uint32_t lds[k_batch_max_block / 64];
    if (B.ctl->done) return;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    trace_counters cnt{0, 0};
    uint32_t nshadow = 0, nmis = 0, nlight = 0;
    uint32_t st = ST_ALIVE;
    if (slot < B.pool) {
        st = P.U(U_STATE, slot);
        if (st & ST_CONNECT) {
            dscene Sj;
            dparams pj;
            job_view(S, base, B.jobs[P.U(U_JOB, slot)], Sj, pj);
            connect_path<W, false, true, FAST, WALKED>(Sj, P, pj, slot, cnt, nshadow, nmis, nlight);
            st &= ~ST_CONNECT;
            P.U(U_STATE, slot) = st;
        }
    }
    stripe_add64(B.counters, 1, nshadow);
    stripe_add64(B.counters, 2, nmis);
    stripe_add64(B.counters, 3, nlight);
    uint32_t nfree = 0;
    block_rank(slot < B.pool && (st == 0u || (st & ST_DYING)), nfree, lds);
    if (threadIdx.x == 0 && nfree) atomicAdd(&B.block_free[blockIdx.x * blockDim.x / k_group], nfree);
}

// every pixel inside the bounds has spp own samples; the film's weights count the spills
__global__ void k_b_resolve(dcamera cam, dfilm film, uint32_t spp, float* rgb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cam.width * cam.height) return;
    film_resolve_pixel(cam, film, p, rgb, spp);
}

}  // namespace

struct rbw_batch_state {
    // merged scene of the last submission
    std::vector<uint32_t> scene_ids;
    dev_scene scene;
    std::vector<uint32_t> roots, roots4, roots8, em_first, em_count, light_flags;
    uint32_t merged_tris = 0;
    // path pool
    dbuf pf, pu, prng;
    uint32_t pool = 0;
    // per batch
    dbuf ctl, jobs, block_free, block_n, block_base, wave_job, first_iter, last_iter, counters, fetch, mqueue, films, rgb;
    dctl* host_ctl = nullptr;  // pinned
    // guided jobs: per-slot guided vertex data and training records, per-job lists and rows
    dbuf g_surf, g_query, g_mix, g_rec, g_flag, g_rows, g_count, t_in, t_dir, t_c, t_count;
    uint32_t g_pool = 0, g_depth = 0, g_jobs = 0, g_tcap = 0;
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

// the single render's schedule (measured best on the 1.3M-triangle icosphere at 1024^2, 64 spp:
// block 64 / 128 / 256 with 2^23 slots: 348 / 332 / 309 Msamples/s; block 64 with 2^22 / 2^24 /
// 2^25 slots: 332 / 337 / 312, one slot per lane; with the persistent extension and occlusion
// walks and grouped shading, 2^23 / 2^24 slots: 430 / 445: the persistent walks gain from
// more items in flight)
constexpr uint32_t k_render_block = 64;
constexpr uint32_t k_render_pool = 1u << 24;

struct batch_config {
    uint32_t block = 256, pool = 1u << 21, width = k_default_width;
    bool graph = true;
    uint32_t tri_test = RBW_TRI_MOLLER_TRUMBORE, shadow = RBW_SHADOW_CLOSEST_VISIBLE;
    // RBW_SAMPLER_PER_SAMPLE_PCG or RBW_SAMPLER_STRATIFIED (a single render's sampler)
    uint32_t sampler = RBW_SAMPLER_PER_SAMPLE_PCG, strata_x = 1, strata_y = 1, dims_1d = 0, dims_2d = 0;
    // the autotuner's probes: stop after this many graph launches (0: run to the end)
    uint32_t probe_launches = 0;
    // guided jobs (job j: network j of `model`)
    const rbw_guiding_opts* guiding = nullptr;
    guide::model* model = nullptr;
    uint32_t seed = 0;
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
    c.width = resolve_width(o->bvh_width);
    c.graph = o->use_graph != 0;
    c.tri_test = o->triangle_test;
    c.shadow = o->shadow_mode;
    require(c.block == 64 || c.block == 128 || c.block == 256 || c.block == 512, RBW_ERR_INVALID_ARGUMENT,
            "block_size must be 64, 128, 256 or 512");
    require(c.pool >= 1024 && c.pool <= (1u << 26), RBW_ERR_INVALID_ARGUMENT, "pool_size must be in [2^10, 2^26]");
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
    merge_scenes(parts, merged, b.roots, b.roots4, b.roots8, b.em_first, b.em_count);
    b.light_flags.clear();
    for (const host_scene* h : parts) b.light_flags.push_back(h->light_flags);
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
    b.pool = pool;
}

struct run_result {
    double ms = 0.0;
    uint32_t iterations = 0, launches = 0;
    unsigned long long started = 0;  // camera samples started (all of them unless cut short)
    unsigned long long guide_samples = 0;
    uint32_t guide_steps = 0;
    unsigned long long counters[4] = {0, 0, 0, 0};
    std::vector<double> launch_ms;  // end time of each launch from the start
    std::vector<uint32_t> first_iter, last_iter;
};

// one batch through the bounce loop; films must be cleared by the caller
// blocks of a persistent kernel: as many as the device holds at once (cached per kernel)
static uint32_t persistent_blocks(const void* kernel, uint32_t block, size_t lds) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, uint32_t> cache;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(dev, kernel);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, lds));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t n = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    cache[key] = n;
    return n;
}

void run_batch(rbw_ctx* ctx, const dscene& S, const std::vector<djob>& hjobs, const batch_config& cfg, run_result& rr,
               bool want_job_times, kernel_timer* timer = nullptr) {
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
    const uint32_t bs = cfg.block;
    const uint32_t nblocks = blocks(cfg.pool, bs);
    const uint32_t ngroups = blocks(cfg.pool, k_group);
    b.block_free.ensure(size_t(ngroups) * 4);
    b.block_n.ensure(size_t(ngroups) * 4);
    b.block_base.ensure(size_t(ngroups) * 8);
    const uint32_t nwaves = blocks(cfg.pool, 64);
    b.wave_job.ensure(size_t(nwaves) * 4);
    b.first_iter.ensure(size_t(nj) * 4);
    b.last_iter.ensure(size_t(nj) * 4);
    b.counters.ensure(RBW_STRIPES * RBW_STRIPE_WORDS * 8);
    b.fetch.ensure(256);
    b.mqueue.ensure(size_t(cfg.pool) * 4);
    if (!b.host_ctl) HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b.host_ctl), sizeof(dctl), hipHostMallocDefault));

    dbatch B{};
    B.ctl = b.ctl.as<dctl>();
    B.jobs = b.jobs.as<djob>();
    B.njobs = nj;
    B.pool = cfg.pool;
    B.ngroups = ngroups;
    B.block_free = b.block_free.as<uint32_t>();
    B.block_n = b.block_n.as<uint32_t>();
    B.block_base = b.block_base.as<unsigned long long>();
    B.wave_job = b.wave_job.as<uint32_t>();
    B.nwaves = nwaves;
    B.first_iter = b.first_iter.as<uint32_t>();
    B.last_iter = b.last_iter.as<uint32_t>();
    B.counters = b.counters.as<unsigned long long>();
    B.fetch = b.fetch.as<uint32_t>();
    B.mqueue = b.mqueue.as<uint32_t>();
    dpaths P;
    P.f = b.pf.as<float>();
    P.u = b.pu.as<uint32_t>();
    P.rng = b.prng.as<uint64_t>();
    P.cap = cfg.pool;
    dparams base;
    std::memset(&base, 0, sizeof(base));
    base.sampler = cfg.sampler;
    base.strata_x = cfg.strata_x;
    base.strata_y = cfg.strata_y;
    base.dims_1d = cfg.dims_1d;
    base.dims_2d = cfg.dims_2d;
    base.tri_test = cfg.tri_test;
    base.shadow_mode = cfg.shadow;
    // grid: one thread per slot of the pool
    const uint32_t grid = nblocks;
    // guided jobs: per-slot vertex data and training records, per-job lists and training rows
    const bool guided = cfg.guiding != nullptr;
    const bool train = guided && cfg.guiding->train;
    B.guided = guided ? 1u : 0u;
    B.train = train ? 1u : 0u;
    B.grows = nullptr;
    B.gcount = nullptr;
    B.t_in = B.t_dir = B.t_c = nullptr;
    B.t_count = nullptr;
    B.t_cap = 0;
    std::vector<uint32_t> host_tcount;
    if (guided) {
        const rbw_guiding_opts& go = *cfg.guiding;
        const uint32_t pool = cfg.pool;
        const uint32_t tcap = go.batch * 2;  // training rows kept per job between training points
        if (b.g_pool != pool || b.g_depth < max_depth || b.g_jobs < nj || b.g_tcap < tcap) {
            const uint32_t depth = std::max(max_depth, b.g_depth);
            const uint32_t jobs = std::max(nj, b.g_jobs);
            b.g_surf.ensure(size_t(G_REC) * pool * 4);
            b.g_query.ensure(size_t(pool) * 9 * 4);
            b.g_mix.ensure(size_t(pool) * guide::N_OUT_PAD * 4);
            b.g_rec.ensure(size_t(R_STRIDE) * pool * depth * 4);
            b.g_flag.ensure(size_t(pool) * depth * 4);
            b.g_rows.ensure(size_t(jobs) * pool * 4);
            b.g_count.ensure(size_t(jobs) * 4);
            b.t_in.ensure(size_t(jobs) * tcap * 9 * 4);
            b.t_dir.ensure(size_t(jobs) * tcap * 3 * 4);
            b.t_c.ensure(size_t(jobs) * tcap * 4);
            b.t_count.ensure(size_t(jobs) * 4);
            b.g_pool = pool;
            b.g_depth = depth;
            b.g_jobs = jobs;
            b.g_tcap = tcap;
        }
        HIP_CHECK(hipMemsetAsync(b.g_flag.p, 0, size_t(pool) * b.g_depth * 4, s));
        HIP_CHECK(hipMemsetAsync(b.t_count.p, 0, size_t(nj) * 4, s));
        dguide& G = base.g;
        G.enable = 1;
        G.alpha = go.mis_fraction;
        G.train = train ? 1u : 0u;
        G.light_records = 0;
        G.depth = go.guide_depth ? go.guide_depth : 0xFFFFFFFFu;
        G.product = go.product_targets ? 1u : 0u;
        G.surf = b.g_surf.as<float>();
        G.surf_mat = nullptr;  // the material is in the slot's surface record
        G.query = b.g_query.as<float>();
        G.mix = b.g_mix.as<float>();
        G.rec = b.g_rec.as<float>();
        G.rec_flag = b.g_flag.as<uint32_t>();
        G.rec_depth = b.g_depth;
        G.cap = pool;
        B.grows = b.g_rows.as<uint32_t>();
        B.gcount = b.g_count.as<uint32_t>();
        B.t_in = b.t_in.as<float>();
        B.t_dir = b.t_dir.as<float>();
        B.t_c = b.t_c.as<float>();
        B.t_count = b.t_count.as<uint32_t>();
        B.t_cap = b.g_tcap;
        host_tcount.resize(nj);
        guide::prepare_kernels();
    }

    const uint32_t init_n = std::max(std::max(cfg.pool, nj), RBW_STRIPES * RBW_STRIPE_WORDS);
    hipLaunchKernelGGL(k_b_init, dim3(blocks(init_n, 256)), dim3(256), 0, s, B, P, total);
    HIP_CHECK(hipGetLastError());

    // the walks' LDS stacks
    const size_t trav_lds = lds_stack_bytes(bs);
    // per-kernel HIP events only without graph capture (kernel_ms of rbw_stats)
    kernel_timer* tm = cfg.graph ? nullptr : timer;
    // the occlusion-only connect kernel when every job's scene has few planar lights
    bool fast = base.debug == nullptr;
    for (const djob& J : hjobs) fast = fast && (J.light_flags & (k_lights_few | k_lights_planar)) == (k_lights_few | k_lights_planar);
    void (*connect_kernel)(dscene, dpaths, dparams, dbatch) =
        S.width == 8   ? (fast ? k_b_connect<8, true> : k_b_connect<8, false>)
        : S.width == 4 ? (fast ? k_b_connect<4, true> : k_b_connect<4, false>)
                       : (fast ? k_b_connect<2, true> : k_b_connect<2, false>);
    void (*extend_kernel)(dscene, dpaths, dparams, dbatch) =
        S.width == 8 ? k_b_extend<8> : (S.width == 4 ? k_b_extend<4> : k_b_extend<2>);
    // the four-wide extension walk runs persistent when the pool gives every resident wave a
    // range or more (debug-hit renders keep one slot per lane); on the 64-job workload's small
    // jobs, rendered one at a time, the persistent walk took 0.27 ms per launch against 0.14
    const uint32_t dyn_grid = S.width == 4 && base.debug == nullptr
                                  ? persistent_blocks(reinterpret_cast<const void*>(&k_b_extend_dyn), 64,
                                                      lds_stack_bytes(64))
                                  : 0u;
    const bool dyn = dyn_grid && cfg.pool >= k_fetch_chunk * dyn_grid;
    // the connect stage's walks run persistent too where they are occlusion queries, except in
    // guided batches (the narrow-aperture rooms, 16 guided jobs at 1024 spp: 30.8 s against
    // 29.8 s with the one-slot-per-lane connect kernel when a kernel of its own filled the MIS
    // queue, 29.6 s against 29.7 s when the guided shading kernel did)
    const uint32_t occl_grid =
        dyn && fast && !guided
            ? persistent_blocks(reinterpret_cast<const void*>(&k_b_occlusion_dyn), 64, lds_stack_bytes(64))
            : 0u;
    if (occl_grid) connect_kernel = k_b_connect<4, true, true>;
    auto iteration = [&]() {
        if (tm) tm->begin(RBW_KERNEL_RAYGEN, s);
        hipLaunchKernelGGL(k_b_plan, dim3(1), dim3(1024), 0, s, B);
        hipLaunchKernelGGL(k_b_regen, dim3(ngroups), dim3(k_group), 0, s, S, P, base, B);
        if (tm) tm->end(s);
        if (tm) tm->begin(RBW_KERNEL_EXTEND, s);
        if (dyn)
            hipLaunchKernelGGL(k_b_extend_dyn, dim3(dyn_grid), dim3(64), lds_stack_bytes(64), s, S, P, base, B);
        else
            hipLaunchKernelGGL(extend_kernel, dim3(grid), dim3(bs), trav_lds, s, S, P, base, B);
        if (tm) tm->end(s);
        if (tm) tm->begin(RBW_KERNEL_SHADE, s);
        if (guided) {
            hipLaunchKernelGGL(k_b_shade_guided, dim3(grid), dim3(bs), 0, s, S, P, base, B);
            if (tm) tm->end(s);
            // every job's guided vertices through its network, in one launch
            if (tm) tm->begin(RBW_KERNEL_GUIDE_INFER, s);
            guide::forward_segments(*cfg.model, base.g.query, B.gcount, cfg.pool, cfg.pool, base.g.mix, s, B.grows,
                                    nj);
            if (tm) tm->end(s);
            if (tm) tm->begin(RBW_KERNEL_SHADE_GUIDED, s);
            // a grid-stride walk over the guided rows: 8192 blocks cover an iteration's typical rows
            hipLaunchKernelGGL(k_b_guided, dim3(std::min<uint32_t>(grid, 8192u)), dim3(bs), 0, s, S, P, base, B);
        } else {
            hipLaunchKernelGGL(occl_grid ? k_b_shade_sorted<true> : k_b_shade_sorted<false>,
                               dim3(blocks(cfg.pool, k_sort_window)), dim3(k_sort_window), 0, s, S, P, base, B);
        }
        if (tm) tm->end(s);
        if (tm) tm->begin(RBW_KERNEL_CONNECT, s);
        if (occl_grid)
            hipLaunchKernelGGL(k_b_occlusion_dyn, dim3(occl_grid), dim3(64), lds_stack_bytes(64), s, S, P, base, B);
        hipLaunchKernelGGL(connect_kernel, dim3(grid), dim3(bs), trav_lds, s, S, P, base, B);
        if (tm) tm->end(s);
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
        iteration();
        iteration();
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
            iteration();
            iteration();
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
            if (train) HIP_CHECK(hipMemcpyAsync(host_tcount.data(), B.t_count, size_t(nj) * 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            const bool done = b.host_ctl->cursor >= total && b.host_ctl->live == 0;
            if (train) {
                // every network with enough closed rows takes steps_per_pass Adam steps on up to
                // `batch` of them, all networks in one batched step
                const rbw_guiding_opts& go = *cfg.guiding;
                std::vector<guide::train_rows> rows(cfg.model->nets);
                bool any = false;
                for (uint32_t j = 0; j < nj; j++) {
                    const uint32_t avail = std::min(host_tcount[j], B.t_cap);
                    rr.guide_samples += host_tcount[j];
                    if (avail < 256) continue;
                    rows[j].base = j * B.t_cap;
                    rows[j].n = std::min(avail, go.batch);
                    any = true;
                }
                if (any) {
                    guide::adam_opts ao;
                    ao.lr = go.learning_rate;
                    for (uint32_t st = 0; st < go.steps_per_pass; st++) {
                        for (uint32_t j = 0; j < nj; j++)
                            if (rows[j].n)
                                rows[j].map = guide::make_sample_map(std::min(host_tcount[j], B.t_cap),
                                                                     (uint64_t(cfg.seed) << 32) ^ (uint64_t(launches) << 12) ^
                                                                         (uint64_t(j) << 6) ^ st);
                        guide::train_step_batched(*cfg.model, B.t_in, B.t_dir, B.t_c, rows, ao, true, s);
                    }
                    rr.guide_steps += go.steps_per_pass;
                }
                HIP_CHECK(hipMemsetAsync(B.t_count, 0, size_t(nj) * 4, s));
            }
            if (done || (cfg.probe_launches && launches >= cfg.probe_launches)) break;
            require(launches < max_launches, RBW_ERR_DEVICE, "batched bounce loop did not drain");
        }
    }
    hipLaunchKernelGGL(k_b_fold, dim3(1), dim3(1024), 0, s, B);  // the final regeneration's records
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    rr.launches = launches;
    rr.iterations = b.host_ctl->iteration;
    rr.started = b.host_ctl->cursor < total ? b.host_ctl->cursor : total;
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, b.events[launches]));
    rr.ms = ms;
    {
        std::vector<unsigned long long> st(RBW_STRIPES * RBW_STRIPE_WORDS);
        HIP_CHECK(hipMemcpy(st.data(), B.counters, st.size() * 8, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < 4; k++) {
            rr.counters[k] = 0;
            for (uint32_t t = 0; t < RBW_STRIPES; t++) rr.counters[k] += st[t * RBW_STRIPE_WORDS + k];
        }
    }
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
        d.root4 = b.roots4[p];
        d.root8 = b.roots8[p];
        d.em_first = b.em_first[p];
        d.em_count = b.em_count[p];
        d.light_flags = b.light_flags[p];
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
    // Probes run the real jobs at their real spp through the real schedule, each launch cut
    // after k_probe_launches graph launches (the first 24 bounce iterations: past the
    // camera-ray ramp of the deepest jobs); a probe's rate is the camera samples it started
    // per millisecond.  Jobs that fill the pool run alone (submit_jobs), so the pool decides
    // which jobs share the launch as well as its size: each candidate pool is scored by the
    // estimated time of the whole batch, the shared jobs' samples at their probed rate plus,
    // for the jobs that run alone under some candidates and share under others, their
    // samples at the rate each one reaches alone.  Jobs alone under every candidate do not
    // change the choice.
    constexpr uint32_t k_probe_launches = 12;
    const std::vector<djob> all = make_jobs(ctx, jobs, n, part, 0);
    auto probe = [&](const std::vector<size_t>& idx, uint32_t block, uint32_t pool, uint32_t width) {
        std::vector<djob> sub;
        unsigned long long base = 0;
        for (size_t i : idx) {
            djob d = all[i];
            d.base = base;
            base += d.count;
            sub.push_back(d);
        }
        batch_config c = cfg;
        c.block = block;
        c.pool = pool;
        c.width = width;
        c.probe_launches = k_probe_launches;
        run_result rr;
        run_batch(ctx, ctx->batch->scene.view(c.width), sub, c, rr, false);
        return double(rr.started) / std::max(rr.ms, 1e-3);  // samples per ms
    };
    auto shared_of = [&](uint32_t pool) {
        std::vector<size_t> idx;
        for (size_t i = 0; i < n; i++)
            if (n == 1 || all[i].count < pool) idx.push_back(i);
        return idx;
    };
    auto alone_pool = [&](size_t i, uint32_t pool) {
        const unsigned long long want = (all[i].count + 1023) & ~1023ull;
        return std::max<uint32_t>(pool, (uint32_t)std::min<unsigned long long>(k_render_pool, want));
    };
    // coordinate search on the shared launch at the given pool: BVH width, then block size
    uint32_t best_block = cfg.block, best_pool = cfg.pool, best_width = k_default_width;
    {
        std::vector<size_t> idx = shared_of(cfg.pool);
        if (idx.empty()) idx.push_back(0);
        const uint32_t pool = idx.size() == 1 && n > 1 ? alone_pool(idx[0], cfg.pool) : cfg.pool;
        double best = -1.0;
        for (uint32_t width : {2u, 4u, 8u}) {
            const double r = probe(idx, best_block, pool, width);
            if (r > best) {
                best = r;
                best_width = width;
            }
        }
        for (uint32_t block : {64u, 128u, 256u, 512u}) {
            if (block == best_block) continue;
            const double r = probe(idx, block, pool, best_width);
            if (r > best) {
                best = r;
                best_block = block;
            }
        }
    }
    // the pool: 2^20, 2^21 or 2^22 slots (adding 2^23 on the 64-job workload: chosen, the same
    // 476 Msamples/s, autotuning 0.91 s against 0.59)
    if (n > 1) {
        const uint32_t cands[3] = {1u << 20, 1u << 21, 1u << 22};
        std::vector<double> alone_rate(n, 0.0);
        for (size_t i = 0; i < n; i++)
            if (all[i].count >= cands[0] && all[i].count < cands[2])
                alone_rate[i] = probe({i}, best_block, alone_pool(i, cands[0]), best_width);
        double best_t = 1e300;
        for (uint32_t pool : cands) {
            const std::vector<size_t> idx = shared_of(pool);
            double t = 0.0;
            if (!idx.empty()) {
                unsigned long long shared = 0;
                for (size_t i : idx) shared += all[i].count;
                t += double(shared) / probe(idx, best_block, pool, best_width);
            }
            for (size_t i = 0; i < n; i++)
                if (all[i].count >= pool && all[i].count < cands[2]) t += double(all[i].count) / alone_rate[i];
            if (t < best_t) {
                best_t = t;
                best_pool = pool;
            }
        }
    }
    inout->block_size = best_block;
    inout->pool_size = best_pool;
    inout->bvh_width = best_width;
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
    const rbw_guiding_opts* go = o.guiding && o.guiding->enable ? o.guiding : nullptr;
    if (go) {
        require(ctx->gjobs.nets >= n, RBW_ERR_INVALID_ARGUMENT,
                "a guided batch needs a network per job (rbw_guide_jobs_reset with at least as many networks as jobs)");
        require(go->mis_fraction >= 0.0f && go->mis_fraction <= 1.0f, RBW_ERR_INVALID_ARGUMENT,
                "guiding mis_fraction must be in [0, 1]");
        require(!go->train || (go->batch >= 256 && go->batch <= (1u << 20)), RBW_ERR_INVALID_ARGUMENT,
                "guiding batch must be in [256, 2^20]");
        require(!go->light_records, RBW_ERR_UNSUPPORTED, "guided batches take no light-sample training records");
        require(!o.autotune, RBW_ERR_UNSUPPORTED, "the autotuner probes plain batches only");
    }
    double tune_ms = 0.0;
    if (o.autotune) autotune_jobs(ctx, jobs, n, &o, &tune_ms);
    batch_config cfg = config_of(&o);
    if (go) {
        cfg.guiding = go;
        cfg.model = &ctx->gjobs;
        cfg.seed = jobs[0].seed;
    }
    std::vector<uint32_t> part;
    prepare_scene(ctx, jobs, n, part);
    rbw_batch_state& b = *ctx->batch;
    std::vector<djob> hj = make_jobs(ctx, jobs, n, part, 0);
    if (go) {
        // each job's network query normalises positions by its scene's box
        for (size_t i = 0; i < n; i++) {
            const host_scene& h = *ctx->pool[jobs[i].scene];
            hj[i].gmin = h.smin;
            hj[i].ginv = f3{1.0f / std::max(h.smax.x - h.smin.x, 1e-6f), 1.0f / std::max(h.smax.y - h.smin.y, 1e-6f),
                            1.0f / std::max(h.smax.z - h.smin.z, 1e-6f)};
        }
    }
    // Jobs with at least as many samples as the pool has slots fill it on their own: they run
    // alone, one after another in submission order, and the other jobs then share one
    // persistent launch.  (Measured on the 64-job workload: letting large jobs share the
    // pool only mixes their scenes in the caches -- the shared launch ran 14 % slower than
    // the same jobs rendered one by one.)  Images do not depend on the grouping: a sample's
    // stream and film sums depend only on its job, pixel and index.
    std::vector<std::vector<size_t>> groups;
    {
        std::vector<size_t> small;
        for (size_t i = 0; i < n; i++) {
            // (guided jobs all share the launch: their networks are addressed by job index)
            if (!go && n > 1 && hj[i].count >= cfg.pool) groups.push_back({i});
            else small.push_back(i);
        }
        if (!small.empty()) groups.push_back(small);
    }
    run_result rr;
    std::vector<double> job_active_ms(n, 0.0);
    rr.first_iter.assign(n, 0);
    rr.last_iter.assign(n, 0);
    const dscene Sv = b.scene.view(cfg.width);
    for (const auto& g : groups) {
        std::vector<djob> sub;
        unsigned long long base = 0;
        for (size_t i : g) {
            djob d = hj[i];
            d.base = base;
            base += d.count;
            sub.push_back(d);
        }
        // a job alone gets at least the single render's slot count (capped by its samples)
        batch_config gcfg = cfg;
        if (g.size() == 1 && groups.size() > 1) {
            const unsigned long long want = (sub[0].count + 1023) & ~1023ull;
            gcfg.pool = std::max<uint32_t>(cfg.pool, (uint32_t)std::min<unsigned long long>(k_render_pool, want));
        }
        run_result r;
        run_batch(ctx, Sv, sub, gcfg, r, job_stats != nullptr);
        if (job_stats) {
            for (size_t k = 0; k < g.size(); k++) {
                const size_t i = g[k];
                rr.first_iter[i] = r.first_iter[k] + rr.iterations;
                rr.last_iter[i] = r.last_iter[k] + rr.iterations;
                // launch l (1-based) runs iterations 2l-1 and 2l
                const uint32_t l0 = (r.first_iter[k] + 1) / 2, l1 = (r.last_iter[k] + 1) / 2;
                if (l0 >= 1 && l1 >= l0 && l1 < r.launch_ms.size()) job_active_ms[i] = r.launch_ms[l1] - r.launch_ms[l0 - 1];
            }
        }
        rr.ms += r.ms;
        rr.iterations += r.iterations;
        rr.launches += r.launches;
        rr.guide_samples += r.guide_samples;
        rr.guide_steps += r.guide_steps;
        for (int k = 0; k < 4; k++) rr.counters[k] += r.counters[k];
    }
    // films -> caller
    hipStream_t s = ctx->stream;
    size_t max_px = 0;
    for (size_t i = 0; i < n; i++) max_px = std::max(max_px, size_t(jobs[i].camera.width) * jobs[i].camera.height);
    b.rgb.ensure(max_px * 3 * sizeof(float));
    for (size_t i = 0; i < n; i++) {
        if (!jobs[i].out_rgb) continue;
        const size_t np = size_t(jobs[i].camera.width) * jobs[i].camera.height;
        hipLaunchKernelGGL(k_b_resolve, dim3(blocks(np, 256)), dim3(256), 0, s, hj[i].cam, hj[i].film, hj[i].spp,
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
            js.active_ms = job_active_ms[i];
            js.msamples_per_s = js.active_ms > 0 ? double(js.samples) / js.active_ms * 1e-3 : 0.0;
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
        stats->guide_samples = rr.guide_samples;
        stats->guide_steps = rr.guide_steps;
    }
}

// rbw_render's plain estimator: the scene of rbw_upload_scene as a one-job batch
void render_scheduled(rbw_ctx* ctx, const rbw_camera_desc* cam, const rbw_render_params& pr, float* out_rgb,
                      rbw_stats* stats) {
    if (!ctx->batch) ctx->batch = new rbw_batch_state();
    rbw_batch_state& b = *ctx->batch;
    hipStream_t s = ctx->stream;
    const size_t np = size_t(cam->width) * cam->height;
    b.films.ensure(np * 32);
    HIP_CHECK(hipMemsetAsync(b.films.p, 0, np * 32, s));
    djob J;
    std::memset(&J, 0, sizeof(J));
    J.cam = make_dcamera(cam);
    char* f = b.films.as<char>();
    J.film.sum = reinterpret_cast<unsigned long long*>(f);
    J.film.weight = reinterpret_cast<uint32_t*>(f + np * 24);
    J.film.nan = reinterpret_cast<uint32_t*>(f + np * 28);
    J.root = 0;
    J.root4 = 0;
    J.root8 = 0;
    J.em_first = 0;
    J.em_count = ctx->main.num_emitters;
    J.light_flags = ctx->main.light_flags;
    J.bw = uint32_t(cam->bound_max[0] - cam->bound_min[0]);
    J.npix = J.bw * uint32_t(cam->bound_max[1] - cam->bound_min[1]);
    J.spp = pr.spp;
    J.max_depth = pr.max_depth;
    J.seed = pr.seed;
    J.rr_threshold = pr.rr_threshold;
    J.base = 0;
    J.count = (unsigned long long)J.npix * pr.spp;
    batch_config cfg;
    cfg.block = k_render_block;
    // no more slots than samples (rounded to whole blocks)
    const unsigned long long want = (J.count + 1023) & ~1023ull;
    cfg.pool = (uint32_t)std::min<unsigned long long>(k_render_pool, std::max<unsigned long long>(1024, want));
    cfg.graph = !pr.time_kernels;
    cfg.tri_test = pr.triangle_test;
    cfg.shadow = pr.shadow_mode;
    cfg.width = resolve_width(pr.bvh_width);
    cfg.sampler = pr.sampler;
    cfg.strata_x = pr.strata_x;
    cfg.strata_y = pr.strata_y;
    cfg.dims_1d = pr.dims_1d;
    cfg.dims_2d = pr.dims_2d;
    ctx->timer.reset();
    run_result rr;
    std::vector<djob> jobs(1, J);
    run_batch(ctx, ctx->main.view(cfg.width), jobs, cfg, rr, false, stats ? &ctx->timer : nullptr);
    b.rgb.ensure(np * 3 * sizeof(float));
    hipEvent_t ea = ctx->timer.get(), eb = ctx->timer.get();
    HIP_CHECK(hipEventRecord(ea, s));
    hipLaunchKernelGGL(k_b_resolve, dim3(blocks(np, 256)), dim3(256), 0, s, J.cam, J.film, J.spp, b.rgb.as<float>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(eb, s));
    if (out_rgb) HIP_CHECK(hipMemcpyAsync(out_rgb, b.rgb.p, np * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float resolve_ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&resolve_ms, ea, eb));
        stats->render_ms = rr.ms + resolve_ms;
        stats->bvh_width = cfg.width;
        stats->samples = J.count;
        stats->camera_rays = J.count;
        stats->extension_rays = rr.counters[0];
        stats->shadow_rays = rr.counters[1];
        stats->mis_rays = rr.counters[2];
        stats->light_casts = rr.counters[3];
        for (const auto& r : ctx->timer.recs) {
            float kms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&kms, r.a, r.b));
            stats->kernel_ms[r.kind] += kms;
            stats->kernel_launches[r.kind] += 1;
        }
        stats->kernel_ms[RBW_KERNEL_FILM] += resolve_ms;
        stats->kernel_launches[RBW_KERNEL_FILM] += 1;
        const double sec = stats->render_ms * 1e-3;
        stats->msamples_per_s = sec > 0 ? double(stats->samples) / sec * 1e-6 : 0.0;
        const uint64_t rays = stats->extension_rays + stats->shadow_rays + stats->mis_rays;
        stats->mrays_per_s = sec > 0 ? double(rays) / sec * 1e-6 : 0.0;
    }
}

}  // namespace host
}  // namespace rbw
