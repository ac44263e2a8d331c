// FP16-MFMA kernels of the path-guiding MLP (inference, training forward,
// backward, Adam).  One wavefront owns 16 rows at a time and runs them through
// the whole network with v_mfma_f32_16x16x16_f16 (FP32 accumulation); layer
// outputs go through a per-wave LDS tile to become the next layer's A operand.
//
// MFMA operand layout (wave64, 16x16x16):
//   A (16x16, M x K): lane l holds A[l % 16][4 * (l / 16) + 0..3]
//   B (16x16, K x N): lane l holds B[4 * (l / 16) + 0..3][l % 16]
//   D (16x16):        lane l holds D[4 * (l / 16) + 0..3][l % 16]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "dmath.hpp"
#include "guiding.hpp"

namespace rbw {
namespace guide {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// stored activations: the inputs of layers 1..5 (layer 0's input, the encoding, is
// recomputed from the 9 raw inputs where the backward pass needs it)
constexpr int ACT_STRIDE = 5 * HIDDEN;  // 320 halfs per row
__host__ __device__ constexpr int act_offset(int l) { return (l - 1) * HIDDEN; }
constexpr int TILE_PAD = HIDDEN + 8;  // LDS row pitch (halfs)

#define GCHECK(expr)                                                                                \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

__device__ __forceinline__ f4 mfma16(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ float pick9(const float* x, int v) {
    float r = x[0];
#pragma unroll
    for (int i = 1; i < N_IN; i++) r = (v == i) ? x[i] : r;
    return r;
}

// four consecutive encoded features starting at f0 (a multiple of 4): features of one
// input scalar v = f0 / 24 at frequencies k and k+1 as (sin, cos, sin, cos)
__device__ __forceinline__ h4 encode4(const float* x, int f0) {
    h4 a;
    if (f0 >= N_FEAT) {
        a[0] = a[1] = a[2] = a[3] = (_Float16)0.0f;
        return a;
    }
    const int v = f0 / 24;
    const int k = (f0 % 24) >> 1;
    const float xv = pick9(x, v);
    // sin/cos(2^k pi x) = sin/cos(2 pi t) with t = 2^(k-1) x; v_sin/v_cos take turns
    float t0 = xv * (float)(1 << k) * 0.5f;
    float t1 = xv * (float)(1 << (k + 1)) * 0.5f;
    t0 = t0 - floorf(t0);
    t1 = t1 - floorf(t1);
    a[0] = (_Float16)__builtin_amdgcn_sinf(t0);
    a[1] = (_Float16)__builtin_amdgcn_cosf(t0);
    a[2] = (_Float16)__builtin_amdgcn_sinf(t1);
    a[3] = (_Float16)__builtin_amdgcn_cosf(t1);
    return a;
}

__device__ __forceinline__ h4 load_h4(const _Float16* p) { return *reinterpret_cast<const h4*>(p); }

// LDS copy of the half weights: rows padded so that the 16 rows one MFMA B
// fragment reads fall into distinct banks (row pitch of 114 dwords for layer 0,
// 34 dwords for the others)
constexpr int W0_PITCH = N_FEAT_PAD + 4;  // halfs
constexpr int WH_PITCH = HIDDEN + 4;      // halfs
__host__ __device__ constexpr int lds_w_offset(int l) {
    return l == 0 ? 0 : HIDDEN * W0_PITCH + (l - 1) * HIDDEN * WH_PITCH;
}
constexpr int LDS_W_HALFS = lds_w_offset(N_LAYERS - 1) + N_OUT_PAD * WH_PITCH;  // 34,944 halfs
constexpr int FWD_WAVES = 4;

// Forward pass.  A block stages every layer's half weights in LDS once, then its
// four waves run 16-row tiles through the network (one wave per tile), persistent
// over the rows: the B fragments of all MFMAs come from LDS.  STORE writes every
// layer's input (training) for all rows up to the padded count; rows >= n use the
// last valid input row.
template <bool STORE>
__global__ __launch_bounds__(FWD_WAVES * 64) void k_forward(const _Float16* __restrict__ w16,
                                                            const float* __restrict__ w32,
                                                            const float* __restrict__ inputs, uint32_t n,
                                                            uint32_t n_pad, float* __restrict__ out,
                                                            _Float16* __restrict__ acts, sample_map map,
                                                            const uint32_t* __restrict__ n_dev = nullptr) {
    // (n_dev: the row count lives in device memory, written by an earlier kernel of the stream)
    if (n_dev) {
        n = *n_dev;
        n_pad = (n + 15) & ~15u;
        map.n_src = n;
        if (n == 0) return;
    }
    extern __shared__ __align__(16) _Float16 lds[];
    _Float16* wl = lds;                                  // LDS_W_HALFS
    _Float16* tiles = lds + LDS_W_HALFS;                 // FWD_WAVES x 16 x TILE_PAD
    // stage the weights: 4 halfs per thread and step, row by row
    for (int l = 0; l < N_LAYERS; l++) {
        const int in = layer_in(l), outs = layer_out(l), pitch = l == 0 ? W0_PITCH : WH_PITCH;
        const _Float16* src = w16 + layer_w_offset(l);
        _Float16* dst = wl + lds_w_offset(l);
        const int quads = outs * in / 4;
        for (int q = threadIdx.x; q < quads; q += blockDim.x) {
            const int row = (q * 4) / in, col = (q * 4) % in;
            *reinterpret_cast<h4*>(dst + row * pitch + col) = load_h4(src + row * in + col);
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lg = lane >> 4;
    _Float16* tile = tiles + wave * 16 * TILE_PAD;
    const uint32_t nwaves = gridDim.x * FWD_WAVES;
    for (uint32_t chunk = blockIdx.x * FWD_WAVES + wave; chunk * 16 < n_pad; chunk += nwaves) {
        const uint32_t r0 = chunk * 16;
        const uint32_t arow = r0 + li;  // the row this lane feeds as A operand
        const uint32_t src = map(arow < n ? arow : n - 1);
        float x[N_IN];
#pragma unroll
        for (int i = 0; i < N_IN; i++) x[i] = inputs[(size_t)src * N_IN + i];
        f4 acc[4];
        // ---- layer 0: encoded inputs (224) -> 64
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        const _Float16* W0 = wl + lds_w_offset(0);
        for (int s2 = 0; s2 < N_FEAT_PAD / 16; s2++) {
            const h4 a = encode4(x, 16 * s2 + 4 * lg);
#pragma unroll
            for (int t = 0; t < 4; t++)
                acc[t] = mfma16(a, load_h4(W0 + (16 * t + li) * W0_PITCH + 16 * s2 + 4 * lg), acc[t]);
        }
        // ---- hidden layers 1..4 and the output layer 5
        for (int l = 1; l < N_LAYERS; l++) {
            // bias + ReLU of layer l-1, to the wave's LDS tile (and the activation store)
            const float* bprev = w32 + layer_b_offset(l - 1);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float b = bprev[16 * t + li];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float z = acc[t][j] + b;
                    const _Float16 h = (_Float16)(z > 0.0f ? z : 0.0f);
                    tile[(4 * lg + j) * TILE_PAD + 16 * t + li] = h;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (STORE) {
                // the tile's 16 rows x 64 halfs to the activation store, 16 bytes per lane and step
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int id = lane + 64 * h, row = id >> 3, c8 = id & 7;
                    *reinterpret_cast<h8*>(acts + (size_t)(r0 + row) * ACT_STRIDE + act_offset(l) + 8 * c8) =
                        *reinterpret_cast<const h8*>(&tile[row * TILE_PAD + 8 * c8]);
                }
            }
            const int nt = layer_out(l) / 16;
            const _Float16* Wl = wl + lds_w_offset(l);
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s2 = 0; s2 < HIDDEN / 16; s2++) {
                const h4 a = load_h4(&tile[li * TILE_PAD + 16 * s2 + 4 * lg]);
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (t < nt) acc[t] = mfma16(a, load_h4(Wl + (16 * t + li) * WH_PITCH + 16 * s2 + 4 * lg), acc[t]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // ---- raw outputs (48 columns, 40 live) with bias
        const float* b5 = w32 + layer_b_offset(N_LAYERS - 1);
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const float b = b5[16 * t + li];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t r = r0 + 4 * lg + j;
                if (r < n_pad) out[(size_t)r * N_OUT_PAD + 16 * t + li] = acc[t][j] + b;
            }
        }
    }
}

constexpr size_t FWD_LDS_BYTES = (size_t(LDS_W_HALFS) + size_t(FWD_WAVES) * 16 * TILE_PAD) * sizeof(_Float16);

__device__ __forceinline__ float softplus(float x) { return x > 20.0f ? x : log1pf(expf(x)); }

// d loss / d raw output of the vMF-mixture head for one row; returns log q
__device__ __forceinline__ float mixture_grad(const float* o, const float* d, float scale, float* g) {
    float w[N_LOBES], vv[N_LOBES], kap[N_LOBES], mu[N_LOBES][3], nm[N_LOBES], cosv[N_LOBES], sig[N_LOBES];
    bool kcap[N_LOBES];
    float amax = -3.4e38f;
    for (int j = 0; j < N_LOBES; j++) amax = fmaxf(amax, o[5 * j + 4]);
    float wsum = 0.0f;
    for (int j = 0; j < N_LOBES; j++) {
        w[j] = expf(o[5 * j + 4] - amax);
        wsum += w[j];
    }
    float q = 0.0f;
    for (int j = 0; j < N_LOBES; j++) {
        w[j] /= wsum;
        const float mx = o[5 * j], my = o[5 * j + 1], mz = o[5 * j + 2];
        nm[j] = fmaxf(sqrtf(mx * mx + my * my + mz * mz), 1e-12f);
        mu[j][0] = mx / nm[j];
        mu[j][1] = my / nm[j];
        mu[j][2] = mz / nm[j];
        const float sp = softplus(o[5 * j + 3]) + KAPPA_MIN;
        kcap[j] = sp >= KAPPA_MAX;
        kap[j] = fminf(sp, KAPPA_MAX);
        sig[j] = 1.0f / (1.0f + expf(-o[5 * j + 3]));
        cosv[j] = mu[j][0] * d[0] + mu[j][1] * d[1] + mu[j][2] * d[2];
        const float norm = kap[j] / (2.0f * 3.14159265f * (-expm1f(-2.0f * kap[j])));
        vv[j] = norm * expf(kap[j] * (cosv[j] - 1.0f));
        q += w[j] * vv[j];
    }
    q = fmaxf(q, 1e-30f);
    const float gq = -scale;  // d loss / d log q (scaled)
    for (int j = 0; j < N_LOBES; j++) {
        const float r = w[j] * vv[j] / q;
        const float e = expf(-2.0f * kap[j]);
        const float dlv = 1.0f / kap[j] - 2.0f * e / (-expm1f(-2.0f * kap[j])) + (cosv[j] - 1.0f);
        g[5 * j + 3] = kcap[j] ? 0.0f : gq * r * dlv * sig[j];
        g[5 * j + 4] = gq * (r - w[j]);
        float gm[3];
        for (int c = 0; c < 3; c++) gm[c] = gq * r * kap[j] * d[c];
        const float p = gm[0] * mu[j][0] + gm[1] * mu[j][1] + gm[2] * mu[j][2];
        for (int c = 0; c < 3; c++) g[5 * j + c] = (gm[c] - p * mu[j][c]) / nm[j];
    }
    return logf(q);
}

// sum of the batch's weights: SUM_PER_THREAD rows per thread, one atomic per block
constexpr uint32_t SUM_PER_THREAD = 8;
__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ c, uint32_t n, sample_map map,
                                             float* __restrict__ total) {
    __shared__ float part[4];
    const uint32_t first = blockIdx.x * blockDim.x * SUM_PER_THREAD + threadIdx.x;
    float v = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < SUM_PER_THREAD; j++) {
        const uint32_t i = first + j * blockDim.x;
        if (i < n) v += c[map(i)];
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float t = part[0] + part[1] + part[2] + part[3];
        if (t != 0.0f) atomicAdd(total, t);
    }
}

// per row: the clamped, loss-scaled output gradient (FP16, 48 columns) and the loss term
__global__ __launch_bounds__(256) void k_mixture_grad(const float* __restrict__ out, const float* __restrict__ dirs,
                                                      const float* __restrict__ c, const float* __restrict__ csum,
                                                      uint32_t n, uint32_t n_pad, sample_map map,
                                                      _Float16* __restrict__ dz, float* __restrict__ loss) {
    __shared__ float part[4];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    float lterm = 0.0f;
    if (r < n_pad) {
        float g[N_OUT_PAD];
#pragma unroll
        for (int k = 0; k < N_OUT_PAD; k++) g[k] = 0.0f;
        if (r < n) {
            const float total = *csum;
            const float cmean = total / (float)n;
            const uint32_t src = map(r);
            const float wr = cmean > 0.0f ? fminf(c[src] / cmean, 64.0f) : 0.0f;
            if (wr > 0.0f) {
                float o[N_OUT_PAD];
                const float4* orow = reinterpret_cast<const float4*>(out + (size_t)r * N_OUT_PAD);
#pragma unroll
                for (int k = 0; k < N_OUT_PAD / 4; k++) {
                    const float4 v = orow[k];
                    o[4 * k] = v.x;
                    o[4 * k + 1] = v.y;
                    o[4 * k + 2] = v.z;
                    o[4 * k + 3] = v.w;
                }
                const float lq = mixture_grad(o, dirs + (size_t)src * 3, wr, g);
                lterm = -(c[src] / total) * lq;
            }
        }
        h8* drow = reinterpret_cast<h8*>(dz + (size_t)r * N_OUT_PAD);
#pragma unroll
        for (int k = 0; k < N_OUT_PAD / 8; k++) {
            h8 v;
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (_Float16)fminf(fmaxf(g[8 * k + e], -30000.0f), 30000.0f);
            drow[k] = v;
        }
    }
    for (int off = 32; off > 0; off >>= 1) lterm += __shfl_xor(lterm, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = lterm;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float t = part[0] + part[1] + part[2] + part[3];
        if (t != 0.0f) atomicAdd(loss, t);
    }
}

// dZ_{l-1} = (dZ_l W_l) masked by the ReLU of layer l's input.  The product is formed
// transposed, dZ_{l-1}^T = W_l^T dZ_l^T (A = W_l^T, B = dZ_l^T), so that a lane's four
// accumulators are four consecutive features of one row: 8-byte loads of the stored
// activations and 8-byte stores of the result.
__global__ __launch_bounds__(256) void k_backward_data(const _Float16* __restrict__ w16, int l,
                                                       const _Float16* __restrict__ dz_in, int in_stride,
                                                       const _Float16* __restrict__ acts, uint32_t n_pad,
                                                       _Float16* __restrict__ dz_out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int K = layer_out(l);     // reduction over layer l's outputs
    const int N = layer_in(l);      // 64
    const _Float16* W = w16 + layer_w_offset(l);
    // the B fragments (W, at most 4 x 4 tiles) stay in registers for all of the wave's rows
    h4 wb[4][4];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                wb[s][t][i] = s < K / 16 ? W[(16 * s + 4 * lg + i) * N + 16 * t + li] : (_Float16)0.0f;
    const uint32_t nwaves = gridDim.x * 4;
    for (uint32_t chunk = blockIdx.x * 4 + wave; chunk * 16 < n_pad; chunk += nwaves) {
        const uint32_t r0 = chunk * 16;
        f4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 4; s++) {
            if (s < K / 16) {
                // B fragment of dZ_l^T: k = 16 s + 4 lg + i (a feature), column li (a row)
                const h4 b = load_h4(dz_in + (size_t)(r0 + li) * in_stride + 16 * s + 4 * lg);
#pragma unroll
                for (int t = 0; t < 4; t++) acc[t] = mfma16(wb[s][t], b, acc[t]);
            }
        }
        // acc[t][j] = dZ_{l-1}[row r0 + li][feature 16 t + 4 lg + j]
        const uint32_t r = r0 + li;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const h4 a = load_h4(acts + (size_t)r * ACT_STRIDE + act_offset(l) + 16 * t + 4 * lg);
            h4 o;
#pragma unroll
            for (int j = 0; j < 4; j++)
                o[j] = (_Float16)((float)a[j] > 0.0f ? fminf(fmaxf(acc[t][j], -30000.0f), 30000.0f) : 0.0f);
            *reinterpret_cast<h4*>(dz_out + (size_t)r * HIDDEN + 16 * t + 4 * lg) = o;
        }
    }
}

// Weight gradient of layer L over a block's rows: grad(W)[n][k] = sum_r dZ[r][n] in[r][k],
// grad(b)[n] = sum_r dZ[r][n].  The rows are the MFMA reduction, so a lane's operand is
// four consecutive rows of one column: the block stages 64-row slabs of dZ and of the
// layer input in LDS transposed (row index fastest; pitch 36 dwords, so the 8-byte
// operand reads of a half-wave hit distinct banks).  Layer 0's input, the frequency
// encoding, is recomputed from the 9 raw inputs.  Output tiles (N/16 x (K/16 + 1), the
// last k tile the bias as a column of ones) are dealt to the 4 waves; each block
// writes its partial sums (weights then biases, the parameter layout of the layer)
// to part[blockIdx.x], and k_reduce_partials adds them into grad.
constexpr int BW_ROWS = 64;
constexpr int BW_PITCH = BW_ROWS + 8;  // halfs
constexpr uint32_t BW_MAX_BLOCKS = 512;
__host__ __device__ constexpr int bw_count(int l) { return layer_out(l) * layer_in(l) + layer_out(l); }

template <int L>
__global__ __launch_bounds__(256) void k_backward_weight(const _Float16* __restrict__ dz,
                                                         const _Float16* __restrict__ acts,
                                                         const float* __restrict__ inputs, uint32_t n,
                                                         uint32_t n_pad, uint32_t rows_per_block, sample_map map,
                                                         float* __restrict__ part) {
    constexpr int K = layer_in(L), N = layer_out(L);
    constexpr int KT = K / 16 + 1, NT = N / 16, T = NT * KT, TPW = (T + 3) / 4;
    __shared__ __align__(16) _Float16 actT[K * BW_PITCH];
    __shared__ __align__(16) _Float16 dzT[N * BW_PITCH];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lg = lane >> 4;
    f4 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; j++) acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t rbeg = blockIdx.x * rows_per_block;
    const uint32_t rend = rbeg + rows_per_block < n_pad ? rbeg + rows_per_block : n_pad;
    const h8 zero8 = h8{0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t s0 = rbeg; s0 < rend; s0 += BW_ROWS) {
        // stage the slab: lane = row, waves take turns over 8-column groups
        const uint32_t r = s0 + (uint32_t)lane;
        const bool valid = r < rend;  // rows past the end: zero dZ, no contribution
#pragma unroll
        for (int c8 = wave; c8 < N / 8; c8 += 4) {
            const h8 v = valid ? *reinterpret_cast<const h8*>(dz + (size_t)r * N + 8 * c8) : zero8;
#pragma unroll
            for (int e = 0; e < 8; e++) dzT[(8 * c8 + e) * BW_PITCH + lane] = v[e];
        }
        if (L == 0) {
            const uint32_t src = map(valid && r < n ? r : n - 1);
            float x[N_IN];
#pragma unroll
            for (int i = 0; i < N_IN; i++) x[i] = inputs[(size_t)src * N_IN + i];
            for (int q = wave; q < K / 4; q += 4) {
                const h4 a = encode4(x, 4 * q);
#pragma unroll
                for (int e = 0; e < 4; e++) actT[(4 * q + e) * BW_PITCH + lane] = a[e];
            }
        } else {
#pragma unroll
            for (int c8 = wave; c8 < K / 8; c8 += 4) {
                const h8 v = valid ? *reinterpret_cast<const h8*>(acts + (size_t)r * ACT_STRIDE + act_offset(L) + 8 * c8)
                                   : zero8;
#pragma unroll
                for (int e = 0; e < 8; e++) actT[(8 * c8 + e) * BW_PITCH + lane] = v[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int st = 0; st < BW_ROWS / 16; st++) {
#pragma unroll
            for (int j = 0; j < TPW; j++) {
                const int ti = wave + 4 * j;
                if (ti < T) {
                    const int tn = ti / KT, tk = ti % KT;
                    const h4 a = load_h4(&dzT[(16 * tn + li) * BW_PITCH + 16 * st + 4 * lg]);
                    h4 b;
                    if (tk == KT - 1) {
                        const _Float16 one = (_Float16)(li == 0 ? 1.0f : 0.0f);
                        b = h4{one, one, one, one};
                    } else {
                        b = load_h4(&actT[(16 * tk + li) * BW_PITCH + 16 * st + 4 * lg]);
                    }
                    acc[j] = mfma16(a, b, acc[j]);
                }
            }
        }
        __syncthreads();
    }
    float* dst = part + (size_t)blockIdx.x * bw_count(L);
#pragma unroll
    for (int j = 0; j < TPW; j++) {
        const int ti = wave + 4 * j;
        if (ti >= T) break;
        const int tn = ti / KT, tk = ti % KT;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int nrow = 16 * tn + 4 * lg + e;
            if (tk == KT - 1) {
                if (li == 0) dst[N * K + nrow] = acc[j][e];
            } else {
                dst[nrow * K + 16 * tk + li] = acc[j][e];
            }
        }
    }
}

// grad[layer_w_offset(l) + p] += sum over blocks of part[b][p]; blockIdx.y takes a
// group of BW_GROUP blocks
constexpr uint32_t BW_GROUP = 16;
__global__ void k_reduce_partials(const float* __restrict__ part, uint32_t nblocks, uint32_t count,
                                  float* __restrict__ grad) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const uint32_t b0 = blockIdx.y * BW_GROUP;
    const uint32_t b1 = b0 + BW_GROUP < nblocks ? b0 + BW_GROUP : nblocks;
    float s = 0.0f;
    for (uint32_t b = b0; b < b1; b++) s += part[(size_t)b * count + p];
    atomicAdd(&grad[p], s);
}

__device__ __forceinline__ bool live_param(uint32_t p) {
    // padded input columns of layer 0 and padded outputs of the last layer stay zero
    if (p < (uint32_t)layer_b_offset(0)) return (p % N_FEAT_PAD) < N_FEAT;
    const uint32_t w5 = layer_w_offset(N_LAYERS - 1), b5 = layer_b_offset(N_LAYERS - 1);
    if (p >= w5 && p < b5) return (p - w5) / HIDDEN < N_OUT;
    if (p >= b5) return p - b5 < N_OUT;
    return true;
}

__global__ void k_adam(float* __restrict__ w32, _Float16* __restrict__ w16, float* __restrict__ m,
                       float* __restrict__ v, float* __restrict__ grad, float inv_n, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint32_t apply) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)N_PARAMS) return;
    if (!apply) return;
    if (!live_param(p)) {
        grad[p] = 0.0f;
        return;
    }
    const float g = grad[p] * inv_n;
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = w32[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    w32[p] = w;
    w16[p] = (_Float16)w;
    grad[p] = 0.0f;
}

__global__ void k_to_half(const float* __restrict__ w32, _Float16* __restrict__ w16) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)N_PARAMS) w16[p] = (_Float16)w32[p];
}

// ------------------------------------------------------------------ host

void model_free(model& M) {
    for (void* p : {(void*)M.w16, (void*)M.w32, (void*)M.m, (void*)M.v, (void*)M.grad, (void*)M.loss,
                    (void*)M.acts, (void*)M.dz, (void*)M.out, (void*)M.part})
        if (p) (void)hipFree(p);
    M = model();
}

void model_set(model& M, const float* host, hipStream_t s) {
    GCHECK(hipMemcpyAsync(M.w32, host, sizeof(float) * N_PARAMS, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_to_half, dim3((N_PARAMS + 255) / 256), dim3(256), 0, s, M.w32, M.w16);
    GCHECK(hipMemsetAsync(M.m, 0, sizeof(float) * N_PARAMS, s));
    GCHECK(hipMemsetAsync(M.v, 0, sizeof(float) * N_PARAMS, s));
    GCHECK(hipMemsetAsync(M.grad, 0, sizeof(float) * N_PARAMS, s));
    M.step = 0;
}

void model_get(model& M, float* host, hipStream_t s) {
    GCHECK(hipMemcpyAsync(host, M.w32, sizeof(float) * N_PARAMS, hipMemcpyDeviceToHost, s));
    GCHECK(hipStreamSynchronize(s));
}

void model_init(model& M, uint32_t seed, hipStream_t s) {
    if (!M.w32) {
        GCHECK(hipMalloc(&M.w16, sizeof(_Float16) * N_PARAMS));
        GCHECK(hipMalloc(&M.w32, sizeof(float) * N_PARAMS));
        GCHECK(hipMalloc(&M.m, sizeof(float) * N_PARAMS));
        GCHECK(hipMalloc(&M.v, sizeof(float) * N_PARAMS));
        GCHECK(hipMalloc(&M.grad, sizeof(float) * N_PARAMS));
        GCHECK(hipMalloc(&M.loss, sizeof(float) * 4));
    }
    // He-uniform, zero biases; the last layer at half scale
    std::vector<float> p(N_PARAMS, 0.0f);
    pcg32 g;
    g.seed(seed);
    for (int l = 0; l < N_LAYERS; l++) {
        const int fan_in = l == 0 ? N_FEAT : HIDDEN;
        const int outs = l == N_LAYERS - 1 ? N_OUT : HIDDEN;
        const float lim = std::sqrt(6.0f / fan_in) * (l == N_LAYERS - 1 ? 0.5f : 1.0f);
        for (int o = 0; o < outs; o++)
            for (int i = 0; i < (l == 0 ? N_FEAT : HIDDEN); i++)
                p[layer_w_offset(l) + o * layer_in(l) + i] = (2.0f * g.uniform() - 1.0f) * lim;
    }
    model_set(M, p.data(), s);
    GCHECK(hipStreamSynchronize(s));
}

void model_ensure(model& M, uint32_t rows) {
    const uint32_t pad = (rows + 15) & ~15u;
    if (pad <= M.cap && M.acts) return;
    for (void* q : {(void*)M.acts, (void*)M.dz, (void*)M.out})
        if (q) (void)hipFree(q);
    GCHECK(hipMalloc(&M.acts, sizeof(_Float16) * (size_t)pad * ACT_STRIDE));
    GCHECK(hipMalloc(&M.dz, sizeof(_Float16) * (size_t)pad * (2 * HIDDEN + N_OUT_PAD)));
    GCHECK(hipMalloc(&M.out, sizeof(float) * (size_t)pad * N_OUT_PAD));
    if (!M.part) GCHECK(hipMalloc(&M.part, sizeof(float) * (size_t)BW_MAX_BLOCKS * bw_count(0)));
    M.cap = pad;
}

// the forward kernel's dynamic LDS (78 KiB) is above the default limit of a launch
static void allow_forward_lds() {
    static bool done = false;
    if (done) return;
    GCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_forward<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(FWD_LDS_BYTES)));
    GCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_forward<true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(FWD_LDS_BYTES)));
    done = true;
}

// persistent forward blocks: two fit a CU's 160 KiB of LDS
static uint32_t forward_blocks(uint32_t n_pad) {
    const uint32_t tiles = n_pad / 16;
    const uint32_t want = (tiles + FWD_WAVES - 1) / FWD_WAVES;
    return want < 512 ? (want ? want : 1) : 512;
}

static uint32_t grid_waves(uint32_t n_pad) {
    const uint32_t chunks = n_pad / 16;
    const uint32_t blocks = (chunks + 3) / 4;
    return blocks < 2048 ? (blocks ? blocks : 1) : 2048;
}

void forward(model& M, const float* inputs, uint32_t n, float* out, hipStream_t s) {
    if (n == 0) return;
    const uint32_t n_pad = (n + 15) & ~15u;
    sample_map id;
    id.n_src = n;
    allow_forward_lds();
    hipLaunchKernelGGL(k_forward<false>, dim3(forward_blocks(n_pad)), dim3(FWD_WAVES * 64), FWD_LDS_BYTES, s, M.w16,
                       M.w32, inputs, n, n_pad,
                       out, nullptr, id);
    GCHECK(hipGetLastError());
}

void forward_dev(model& M, const float* inputs, const uint32_t* n_dev, uint32_t n_max, float* out, hipStream_t s) {
    if (n_max == 0) return;
    sample_map id;
    id.n_src = n_max;
    allow_forward_lds();
    hipLaunchKernelGGL(k_forward<false>, dim3(forward_blocks((n_max + 15) & ~15u)), dim3(FWD_WAVES * 64),
                       FWD_LDS_BYTES, s, M.w16, M.w32, inputs, n_max, (n_max + 15) & ~15u, out, nullptr, id, n_dev);
    GCHECK(hipGetLastError());
}

sample_map make_sample_map(uint32_t n_src, uint64_t seed) {
    sample_map m;
    m.n_src = n_src;
    if (n_src <= 1) return m;
    pcg32 g;
    g.seed(seed);
    auto gcd = [](uint64_t x, uint64_t y) {
        while (y) {
            const uint64_t t = x % y;
            x = y;
            y = t;
        }
        return x;
    };
    uint64_t a = (uint64_t(g.next_u32()) % n_src) | 1u;
    while (gcd(a, n_src) != 1) a = (a + 2) % n_src;
    m.a = uint32_t(a);
    m.b = g.next_u32() % n_src;
    return m;
}

void train_step(model& M, const float* inputs, const float* dirs, const float* weights, uint32_t n,
                const adam_opts& opt, bool apply, hipStream_t s, sample_map map) {
    if (n == 0) return;
    if (map.n_src == 0) map.n_src = n;
    if (map.n_src < n) throw std::runtime_error("training batch larger than the sample set");
    model_ensure(M, n);
    const uint32_t n_pad = (n + 15) & ~15u;
    _Float16* dz5 = M.dz + (size_t)n_pad * 2 * HIDDEN;
    _Float16* dzA = M.dz;
    _Float16* dzB = M.dz + (size_t)n_pad * HIDDEN;
    GCHECK(hipMemsetAsync(M.loss, 0, sizeof(float) * 4, s));
    allow_forward_lds();
    hipLaunchKernelGGL(k_forward<true>, dim3(forward_blocks(n_pad)), dim3(FWD_WAVES * 64), FWD_LDS_BYTES, s, M.w16,
                       M.w32, inputs, n, n_pad,
                       M.out, M.acts, map);
    hipLaunchKernelGGL(k_sum, dim3((n + 256 * SUM_PER_THREAD - 1) / (256 * SUM_PER_THREAD)), dim3(256), 0, s, weights,
                       n, map, M.loss + 1);
    hipLaunchKernelGGL(k_mixture_grad, dim3((n_pad + 255) / 256), dim3(256), 0, s, M.out, dirs, weights,
                       M.loss + 1, n, n_pad, map, dz5, M.loss);
    // weight gradients: about BW_MAX_BLOCKS blocks of whole 64-row slabs
    uint32_t rows_per_block = (n_pad + BW_MAX_BLOCKS - 1) / BW_MAX_BLOCKS;
    rows_per_block = (rows_per_block + BW_ROWS - 1) / BW_ROWS * BW_ROWS;
    const uint32_t wblocks = (n_pad + rows_per_block - 1) / rows_per_block;
    const _Float16* dz = dz5;
    int dz_stride = N_OUT_PAD;
    for (int l = N_LAYERS - 1; l >= 0; l--) {
        switch (l) {
            case 0:
                hipLaunchKernelGGL(k_backward_weight<0>, dim3(wblocks), dim3(256), 0, s, dz, M.acts, inputs, n, n_pad,
                                   rows_per_block, map, M.part);
                break;
            case 1:
                hipLaunchKernelGGL(k_backward_weight<1>, dim3(wblocks), dim3(256), 0, s, dz, M.acts, inputs, n, n_pad,
                                   rows_per_block, map, M.part);
                break;
            case 2:
                hipLaunchKernelGGL(k_backward_weight<2>, dim3(wblocks), dim3(256), 0, s, dz, M.acts, inputs, n, n_pad,
                                   rows_per_block, map, M.part);
                break;
            case 3:
                hipLaunchKernelGGL(k_backward_weight<3>, dim3(wblocks), dim3(256), 0, s, dz, M.acts, inputs, n, n_pad,
                                   rows_per_block, map, M.part);
                break;
            case 4:
                hipLaunchKernelGGL(k_backward_weight<4>, dim3(wblocks), dim3(256), 0, s, dz, M.acts, inputs, n, n_pad,
                                   rows_per_block, map, M.part);
                break;
            default:
                hipLaunchKernelGGL(k_backward_weight<5>, dim3(wblocks), dim3(256), 0, s, dz, M.acts, inputs, n, n_pad,
                                   rows_per_block, map, M.part);
                break;
        }
        const uint32_t count = (uint32_t)bw_count(l);
        hipLaunchKernelGGL(k_reduce_partials, dim3((count + 255) / 256, (wblocks + BW_GROUP - 1) / BW_GROUP), dim3(256),
                           0, s, M.part, wblocks, count, M.grad + layer_w_offset(l));
        if (l > 0) {
            _Float16* nxt = (dz == dzA) ? dzB : dzA;
            hipLaunchKernelGGL(k_backward_data, dim3(grid_waves(n_pad)), dim3(256), 0, s, M.w16, l, dz, dz_stride,
                               M.acts, n_pad, nxt);
            dz = nxt;
            dz_stride = HIDDEN;
        }
    }
    if (apply) {
        M.step++;
        const float c1 = 1.0f - std::pow(opt.b1, (float)M.step);
        const float c2 = 1.0f - std::pow(opt.b2, (float)M.step);
        adam_apply(M, n, opt, c1, c2, s);
    }
    GCHECK(hipGetLastError());
}

void adam_apply(model& M, uint32_t n, const adam_opts& opt, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((N_PARAMS + 255) / 256), dim3(256), 0, s, M.w32, M.w16, M.m, M.v, M.grad,
                       1.0f / (float)n, opt.lr, opt.b1, opt.b2, opt.eps, c1, c2, 1u);
    GCHECK(hipGetLastError());
}

}  // namespace guide
}  // namespace rbw
