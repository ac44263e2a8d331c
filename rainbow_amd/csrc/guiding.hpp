// Learned path-guiding proposal: a small MLP maps frequency-encoded
// (position, normal, outgoing direction) to an 8-lobe von Mises-Fisher
// mixture over incident directions.  FP16 MFMA kernels (CDNA4's
// v_mfma_f32_16x16x32_f16) with FP32 accumulation for batched inference and
// for one fused training step (forward, mixture gradient, backward, weight
// gradients), FP32 Adam.  A model holds several networks (one per render job
// of a batch); inference and training take the networks' rows in segments.
// No counterpart exists in the reference; the continuation sample of the path
// integrator mixes this proposal with BSDF sampling (one-sample MIS).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rbw {
namespace guide {

constexpr int N_IN = 9;
constexpr int N_FREQ = 12;
constexpr int N_FEAT = 216;      // N_IN * N_FREQ * 2
constexpr int N_FEAT_PAD = 224;
constexpr int HIDDEN = 64;
constexpr int N_LAYERS = 6;      // 224->64, 4 x 64->64, 64->40
constexpr int N_LOBES = 8;
constexpr int N_OUT = 40;
constexpr int N_OUT_PAD = 48;
constexpr float KAPPA_MIN = 0.01f;
constexpr float KAPPA_MAX = 500.0f;
constexpr uint32_t MAX_NETS = 32;  // networks of one model

__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? N_FEAT_PAD : HIDDEN; }
__host__ __device__ constexpr int layer_out(int l) { return l == N_LAYERS - 1 ? N_OUT_PAD : HIDDEN; }
// parameter layout (padded): per layer W[out][in] row-major, then b[out]
__host__ __device__ constexpr int layer_w_offset(int l) {
    int o = 0;
    for (int i = 0; i < l; i++) o += layer_out(i) * layer_in(i) + layer_out(i);
    return o;
}
__host__ __device__ constexpr int layer_b_offset(int l) { return layer_w_offset(l) + layer_out(l) * layer_in(l); }
constexpr int N_PARAMS = layer_w_offset(N_LAYERS);  // 34160 padded floats (33128 live)

// The kernels read each network's half weights as MFMA A fragments in fragment
// order (64 lanes x 8 halfs, 1 KiB a fragment): the forward product W_l x of
// every layer (66 fragments) and the backward product W_l^T dz of layers 1..5
// (40 fragments); biases separately in FP32.
constexpr int FWD_FRAGS = 28 + 4 * 8 + 6;  // layer 0: 4 x 7, hidden: 4 x 2 each, last: 3 x 2
constexpr int BWD_FRAGS = 5 * 8;
constexpr int FRAG_HALFS = 64 * 8;
constexpr int PACK_HALFS = (FWD_FRAGS + BWD_FRAGS) * FRAG_HALFS;  // 54,272
constexpr int N_BIAS = 5 * HIDDEN + N_OUT_PAD;                     // 368
__host__ __device__ constexpr int bias_offset(int l) { return l * HIDDEN; }

struct model {
    uint32_t nets = 0;
    float* w32 = nullptr;      // [nets][N_PARAMS] master weights
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    _Float16* pack = nullptr;  // [nets][PACK_HALFS] half weights in fragment order
    float* bias = nullptr;     // [nets][N_BIAS]
    float* grad = nullptr;     // [nets][N_PARAMS]: the mean gradient of the last training step
    float* loss = nullptr;     // [nets]: the loss of the last step (before its update)
    std::vector<uint32_t> step;  // Adam steps taken per network
    // training scratch
    float* part = nullptr;       // [nets * blocks][N_PARAMS] per-block gradient partial sums
    float* loss_part = nullptr;  // [nets * blocks]
    float* csum_part = nullptr;  // [nets][SUM_BLOCKS] partial sums of the records' weights
    uint32_t* idx = nullptr;     // the training rows' mapped record indices (all networks)
    uint32_t idx_cap = 0;
};

struct adam_opts {
    float lr = 1e-2f, b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
};

// training rows r = 0..n-1 read sample (a * r + b) mod n_src of the collected set; with
// gcd(a, n_src) = 1 and n <= n_src the rows are distinct samples
struct sample_map {
    uint32_t n_src = 0, a = 1, b = 0;
    __host__ __device__ uint32_t operator()(uint32_t r) const {
        return (uint32_t)(((uint64_t)a * r + b) % n_src);
    }
};
sample_map make_sample_map(uint32_t n_src, uint64_t seed);

// one network's training rows: records base + map(r), r < n, of the collected arrays
struct train_rows {
    uint32_t base = 0, n = 0;
    sample_map map;
};

// sets the kernels' launch attributes; call once before capturing launches in a graph
void prepare_kernels();
void model_init(model& M, uint32_t seed, hipStream_t s, uint32_t nets = 1);  // allocates, He-uniform init
void model_free(model& M);
void model_get(model& M, float* host_params, hipStream_t s, uint32_t net = 0);  // N_PARAMS floats (padded layout)
void model_set(model& M, const float* host_params, hipStream_t s, uint32_t net = 0);

// inputs: n x 9 floats in [0,1] (device); out: n x N_OUT_PAD raw outputs (device); network 0
void forward(model& M, const float* inputs, uint32_t n, float* out, hipStream_t s);
// the same with the row count in device memory (at most n_max rows; no host synchronisation)
void forward_dev(model& M, const float* inputs, const uint32_t* n_dev, uint32_t n_max, float* out, hipStream_t s);
// networks 0..nets-1 (0: all) at once: network k's rows are inputs/out rows [k * stride,
// k * stride + n_dev[k]), n_dev[k] <= n_max (device memory); with `rows`, network k's row r
// is record rows[k * stride + r] of inputs and out
void forward_segments(model& M, const float* inputs, const uint32_t* n_dev, uint32_t stride, uint32_t n_max,
                      float* out, hipStream_t s, const uint32_t* rows = nullptr, uint32_t nets = 0);
// one training step of every network k with rows[k].n > 0 on its rows: inputs x 9, dirs x 3,
// weights (radiance / pdf), all device.  Leaves each network's mean gradient in M.grad and
// its loss in M.loss and, when apply, takes an Adam step.  Three launches, no memsets.
void train_step_batched(model& M, const float* inputs, const float* dirs, const float* weights,
                        const std::vector<train_rows>& rows, const adam_opts& opt, bool apply, hipStream_t s);
// network 0 on n rows through map (default: rows 0..n-1)
void train_step(model& M, const float* inputs, const float* dirs, const float* weights, uint32_t n,
                const adam_opts& opt, bool apply, hipStream_t s, sample_map map = sample_map());

}  // namespace guide
}  // namespace rbw
