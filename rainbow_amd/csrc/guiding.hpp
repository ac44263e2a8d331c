// Learned path-guiding proposal: a small MLP maps frequency-encoded
// (position, normal, outgoing direction) to an 8-lobe von Mises-Fisher
// mixture over incident directions.  FP16 MFMA kernels with FP32
// accumulation for batched inference, training forward/backward, and Adam.
// No counterpart exists in the reference; the continuation sample of the
// path integrator mixes this proposal with BSDF sampling (one-sample MIS).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

struct model {
    _Float16* w16 = nullptr;   // N_PARAMS (biases included, unused in half)
    float* w32 = nullptr;      // master weights
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    float* grad = nullptr;     // gradient accumulator
    float* loss = nullptr;     // 1 float
    uint32_t step = 0;
    // activation scratch for training (capacity rows)
    _Float16* acts = nullptr;  // inputs of layers 1..5: [5*64] halfs per row
    float* part = nullptr;     // per-block partial weight gradients
    _Float16* dz = nullptr;    // gradients: 2 x [64] halfs per row (ping-pong) + output [48]
    float* out = nullptr;      // raw outputs, [48] floats per row
    uint32_t cap = 0;
};

struct adam_opts {
    float lr = 1e-2f, b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
};

void model_init(model& M, uint32_t seed, hipStream_t s);      // allocates, He-uniform init
void model_free(model& M);
void model_ensure(model& M, uint32_t rows);                   // training scratch capacity
void model_get(model& M, float* host_params, hipStream_t s);  // N_PARAMS floats (padded layout)
void model_set(model& M, const float* host_params, hipStream_t s);

// training rows r = 0..n-1 read sample (a * r + b) mod n_src of the collected set; with
// gcd(a, n_src) = 1 and n <= n_src the rows are distinct samples
struct sample_map {
    uint32_t n_src = 0, a = 1, b = 0;
    __host__ __device__ uint32_t operator()(uint32_t r) const {
        return (uint32_t)(((uint64_t)a * r + b) % n_src);
    }
};
sample_map make_sample_map(uint32_t n_src, uint64_t seed);

// inputs: n x 9 floats in [0,1] (device); out: n x N_OUT_PAD raw outputs (device)
void forward(model& M, const float* inputs, uint32_t n, float* out, hipStream_t s);
// the same with the row count in device memory (at most n_max rows; no host synchronisation)
void forward_dev(model& M, const float* inputs, const uint32_t* n_dev, uint32_t n_max, float* out, hipStream_t s);
// one training step on n samples: inputs n x 9, dirs n x 3, weights n (radiance / pdf), all device.
// Leaves the gradient in M.grad (scaled to the mean loss) and, when apply, takes an Adam step.
void train_step(model& M, const float* inputs, const float* dirs, const float* weights, uint32_t n,
                const adam_opts& opt, bool apply, hipStream_t s, sample_map map = sample_map());
// Adam step on the accumulated gradient of n samples (bias corrections c1, c2)
void adam_apply(model& M, uint32_t n, const adam_opts& opt, float c1, float c2, hipStream_t s);

}  // namespace guide
}  // namespace rbw
