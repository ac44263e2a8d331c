// FP16-MFMA kernels of the path-guiding MLP for CDNA4 (gfx950): batched
// inference and one fused training step, with v_mfma_f32_16x16x32_f16 and FP32
// accumulation.
//
// MFMA operand layout (wave64, 16x16x32, lane l, g = l / 16):
//   A (16 x 32, M x K): lane l holds A[l % 16][8 g + i], i = 0..7
//   B (32 x 16, K x N): lane l holds B[8 g + i][l % 16]
//   D (16 x 16):        lane l holds D[4 g + j][l % 16], j = 0..3
//
// Layers run transposed: Z^T = W X^T, the weights are the A operand (M = output
// features), the records the N columns.  A layer's four 16-feature output tiles
// then sit in the lanes exactly where the next layer needs its B operand, if the
// 32 features of a K block are taken in the order
//   position 8 g + i  <->  feature 32 kb + 4 g + i (i < 4), 32 kb + 16 + 4 g + (i - 4)
// (tile 2 kb's four values of the lane, then tile 2 kb + 1's): the weights are
// stored in that order (fragment order, packed by k_adam_pack), so activations
// never leave the registers between layers.  Layer 0's B operand is the
// frequency encoding, computed in place: K block kb, lane group g holds input
// v = (4 kb + g) / 3 at the four frequencies 4 ((4 kb + g) % 3) + q, q = 0..3, as
// (sin, cos) pairs (9 inputs x 3 quads = 27 of the 28 lane groups; the last is zero).
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

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

#define GCHECK(expr)                                                                                \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

__device__ __forceinline__ f4 mfma32(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

// feature of K-block position (kb, g, i) in the chained layers' order
__host__ __device__ constexpr int hid_feat(int kb, int g, int i) { return 32 * kb + (i < 4 ? 4 * g + i : 12 + 4 * g + i); }
// fragment index of the forward product of layer l, M tile m, K block kb
__host__ __device__ constexpr int fwd_frag(int l, int m, int kb) {
    return l == 0 ? m * 7 + kb : (l < 5 ? 28 + (l - 1) * 8 + m * 2 + kb : 60 + m * 2 + kb);
}
// fragment index of the backward product W_l^T (l = 1..5), M tile m (input features), K block kb
__host__ __device__ constexpr int bwd_frag(int l, int m, int kb) { return FWD_FRAGS + (l - 1) * 8 + m * 2 + kb; }

// sin and cos of 2 pi t for t in [0, 1): the hardware's v_sin / v_cos take revolutions
__device__ __forceinline__ float sin_rev(float t) { return __builtin_amdgcn_sinf(t); }
__device__ __forceinline__ float cos_rev(float t) { return __builtin_amdgcn_cosf(t); }

// Layer 0's B fragment for K block kb of one record (inputs x at stride 1): lane group g
// holds input v at four consecutive frequencies k0..k0+3, sin(2^k pi x) = sin(2 pi t) with
// t = 2^(k-1) x.  The first frequency's sine and cosine come from the hardware units
// (v_sin_f32 / v_cos_f32 take revolutions), the next three by the double-angle identities
// sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin^2 a: two transcendentals per four frequencies
// instead of eight.  Three doublings scale the units' ~1e-6 error by at most 8, far below
// the half precision the features are rounded to.
__device__ __forceinline__ h8 encode_frag(float xv, int kb, int g) {
    h8 b;
    const int Q = 4 * kb + g;
    if (Q >= 27) {
#pragma unroll
        for (int i = 0; i < 8; i++) b[i] = (_Float16)0.0f;
        return b;
    }
    const int k0 = 4 * (Q % 3);
    const float t = xv * (float)(1 << k0) * 0.5f;
    const float f = t - floorf(t);
    float sn = sin_rev(f), cs = cos_rev(f);
    b[0] = (_Float16)sn;
    b[1] = (_Float16)cs;
#pragma unroll
    for (int q = 1; q < 4; q++) {
        const float s2 = sn + sn;
        const float s_next = s2 * cs;
        cs = __builtin_fmaf(-s2, sn, 1.0f);
        sn = s_next;
        b[2 * q] = (_Float16)sn;
        b[2 * q + 1] = (_Float16)cs;
    }
    return b;
}
__device__ __forceinline__ int encode_input(int kb, int g) { return (4 * kb + g) / 3; }  // input v (< 9 when live)

__device__ __forceinline__ h8 lds_h8(const _Float16* p) { return *reinterpret_cast<const h8*>(p); }
__device__ __forceinline__ h8 frag(const _Float16* wl, int f, int lane) { return lds_h8(wl + (f * 64 + lane) * 8); }
// clamp to +-30000 (inside half's range) in one v_med3
__device__ __forceinline__ float clamp_h(float v) { return __builtin_amdgcn_fmed3f(v, -30000.0f, 30000.0f); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ inference

#ifndef RBW_INFER_WAVES
#define RBW_INFER_WAVES 16
#endif
#ifndef RBW_INFER_TILES
#define RBW_INFER_TILES 2
#endif
constexpr int INF_WAVES = RBW_INFER_WAVES;  // waves per block (one block per CU)
constexpr int INF_R = RBW_INFER_TILES;      // 16-record tiles per wave
constexpr uint32_t INF_MAX_BLOCKS = 256;
constexpr size_t INF_XS = size_t(FWD_FRAGS) * FRAG_HALFS * 2 + N_BIAS * 4;  // per-wave input rows
constexpr size_t INF_LDS = INF_XS + size_t(INF_WAVES) * 16 * INF_R * N_IN * 4;
static_assert(INF_XS % 16 == 0, "aligned input rows");

// R record tiles (16 rows each) per wave share every weight fragment read from LDS.
// Network k (k < nets) has rows [k * stride, k * stride + n_k), n_k from n_dev[k] (or n_host
// with one network); with `rows`, row r of network k is record rows[k * stride + r] of
// inputs and out.  The networks' rows are cut into steps of INF_WAVES groups of 16 R rows
// (a step never spans two networks) and each block takes a contiguous range of the steps
// of all networks, so the work spreads over the whole grid however unevenly the networks'
// rows fall; a block stages a network's weights in LDS when its range enters that network.
constexpr uint32_t INF_STEP_ROWS = INF_WAVES * 16 * INF_R;
template <int R>
__global__ __launch_bounds__(INF_WAVES * 64) void k_infer(const _Float16* __restrict__ pack,
                                                          const float* __restrict__ bias,
                                                          const float* __restrict__ inputs, uint32_t n_host,
                                                          const uint32_t* __restrict__ n_dev, uint32_t nets,
                                                          uint32_t stride, float* __restrict__ out,
                                                          const uint32_t* __restrict__ rows = nullptr) {
    static_assert(INF_WAVES * 16 * R == INF_STEP_ROWS, "step size");
    static_assert(MAX_NETS <= 64, "one network per lane");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    // lane k: the steps of networks 0..k (an inclusive scan over the wave)
    uint32_t c = 0;
    if ((uint32_t)lane < nets) {
        const uint32_t nk = n_dev ? n_dev[lane] : n_host;
        c = (nk + INF_STEP_ROWS - 1) / INF_STEP_ROWS;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)c, off);
        if (lane >= off) c += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)c, 63);
    const uint32_t s0 = (uint32_t)((uint64_t)blockIdx.x * total / gridDim.x);
    const uint32_t s1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * total / gridDim.x);
    if (s0 >= s1) return;
    extern __shared__ __align__(16) unsigned char smem[];
    _Float16* wl = reinterpret_cast<_Float16*>(smem);
    float* bl = reinterpret_cast<float*>(smem + size_t(FWD_FRAGS) * FRAG_HALFS * 2);
    // the first step of network k and the network of step st (both wave-uniform)
    auto first_step = [&](uint32_t k) { return k == 0 ? 0u : (uint32_t)__shfl((int)c, (int)k - 1); };
    auto net_of = [&](uint32_t st) {
        return (uint32_t)__popcll(__ballot((uint32_t)lane + 1 < nets && c <= st));
    };
    // the wave's 16 R rows of a step are staged in LDS (coalesced loads, the next step's in
    // flight while this one computes)
    constexpr int XW = 16 * R * N_IN;          // floats per wave
    constexpr int XPL = (XW + 63) / 64;        // per lane
    float* xs = reinterpret_cast<float*>(smem + INF_XS) + wave * XW;
    float xpre[XPL];
    auto load_step = [&](uint32_t st) {
        const uint32_t k = st < s1 ? net_of(st) : 0u;
        const uint32_t nk = st < s1 ? (n_dev ? n_dev[k] : n_host) : 0u;
        const uint32_t r0 = (st - first_step(k)) * INF_STEP_ROWS + (uint32_t)wave * 16 * R;
        const float* in = rows ? inputs : inputs + (size_t)k * stride * N_IN;
#pragma unroll
        for (int q = 0; q < XPL; q++) {
            const uint32_t e = (uint32_t)lane + 64u * q, r = r0 + e / N_IN;
            if (st < s1 && e < (uint32_t)XW && r < nk) {
                const size_t rec = rows ? (size_t)rows[(size_t)k * stride + r] : (size_t)r;
                xpre[q] = in[rec * N_IN + e % N_IN];
            } else {
                xpre[q] = 0.0f;
            }
        }
    };
    load_step(s0);
    uint32_t staged = 0xFFFFFFFFu;
    for (uint32_t st = s0; st < s1; st++) {
        const uint32_t net = net_of(st);
        if (net != staged) {
            __syncthreads();  // every wave is done with the previous network's weights
            const h8* src = reinterpret_cast<const h8*>(pack + (size_t)net * PACK_HALFS);
            h8* dst = reinterpret_cast<h8*>(wl);
            for (int q = threadIdx.x; q < FWD_FRAGS * 64; q += blockDim.x) dst[q] = src[q];
            for (int q = threadIdx.x; q < N_BIAS; q += blockDim.x) bl[q] = bias[(size_t)net * N_BIAS + q];
            __syncthreads();
            staged = net;
        }
        const uint32_t n = n_dev ? n_dev[net] : n_host;
        float* o = rows ? out : out + (size_t)net * stride * N_OUT_PAD;
        const uint32_t* rl = rows ? rows + (size_t)net * stride : nullptr;
        const uint32_t r0 = (st - first_step(net)) * INF_STEP_ROWS + (uint32_t)wave * 16 * R;
#pragma unroll
        for (int k = 0; k < XPL; k++)
            if (lane + 64 * k < XW) xs[lane + 64 * k] = xpre[k];
        wave_sync();
        load_step(st + 1);
        uint32_t row[R];
#pragma unroll
        for (int t = 0; t < R; t++) row[t] = r0 + 16 * t + li;
        f4 acc[R][4];
        h8 X[R][2];
        // every layer's sums start from its biases
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const f4 b0 = *reinterpret_cast<const f4*>(bl + bias_offset(0) + 16 * m + 4 * g);
#pragma unroll
            for (int t = 0; t < R; t++) acc[t][m] = b0;
        }
        // layer 0: the encoding
#pragma unroll 1
        for (int kb = 0; kb < 7; kb++) {
            const int v = encode_input(kb, g);
            h8 b[R];
#pragma unroll
            for (int t = 0; t < R; t++) b[t] = encode_frag(v < N_IN ? xs[(16 * t + li) * N_IN + v] : 0.0f, kb, g);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const h8 a = frag(wl, fwd_frag(0, m, kb), lane);
#pragma unroll
                for (int t = 0; t < R; t++) acc[t][m] = mfma32(a, b[t], acc[t][m]);
            }
        }
        // hidden layers (rolled: unrolled, every layer's fragment loads are hoisted together)
#pragma unroll 1
        for (int l = 1; l < N_LAYERS; l++) {
            // ReLU on half pairs (rounding commutes with max(., 0)) into the next B fragments
#pragma unroll
            for (int t = 0; t < R; t++)
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int j = 0; j < 4; j += 2) {
                        h2 p = {(_Float16)acc[t][m][j], (_Float16)acc[t][m][j + 1]};
                        p = __builtin_elementwise_max(p, h2{(_Float16)0.0f, (_Float16)0.0f});
                        X[t][m >> 1][(m & 1) * 4 + j] = p.x;
                        X[t][m >> 1][(m & 1) * 4 + j + 1] = p.y;
                    }
            const int mt = l == N_LAYERS - 1 ? 3 : 4;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                if (m >= mt) break;
                const f4 bm = *reinterpret_cast<const f4*>(bl + bias_offset(l) + 16 * m + 4 * g);
#pragma unroll
                for (int t = 0; t < R; t++) acc[t][m] = bm;
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    const h8 a = frag(wl, fwd_frag(l, m, kb), lane);
#pragma unroll
                    for (int t = 0; t < R; t++) acc[t][m] = mfma32(a, X[t][kb], acc[t][m]);
                }
            }
        }
        // raw outputs (48 columns, 40 live; biases included)
#pragma unroll
        for (int t = 0; t < R; t++) {
            if (row[t] >= n) continue;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const int c0 = 16 * m + 4 * g;
                float4 v;
                v.x = acc[t][m][0];
                v.y = acc[t][m][1];
                v.z = acc[t][m][2];
                v.w = acc[t][m][3];
                const size_t orow = rl ? (size_t)rl[row[t]] : (size_t)row[t];
                *reinterpret_cast<float4*>(o + orow * N_OUT_PAD + c0) = v;
            }
        }
    }
}

// ------------------------------------------------------------------ training

// the training head's transcendentals on the hardware units (v_exp_f32 / v_log_f32 / v_rcp_f32,
// ~1 ulp): the head's gradients are rounded to half afterwards
__device__ __forceinline__ float hexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
__device__ __forceinline__ float hlog(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
__device__ __forceinline__ float hrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hsoftplus(float x) { return x > 20.0f ? x : hlog(1.0f + hexp(x)); }

struct train_seg {
    uint32_t base, n, n_src, a, b;
    float c1, c2;
    uint32_t idx_off;  // the network's first entry in the mapped-index array
    __device__ uint32_t src(uint32_t r) const {
        return base + (uint32_t)(((uint64_t)a * r + b) % n_src);
    }
};
struct train_segs {
    train_seg s[MAX_NETS];
};

constexpr uint32_t SUM_BLOCKS = 256;

// partial sums of each network's record weights (block (b, net) writes csum_part[net][b]) and
// the records' mapped indices, idx[idx_off + r] = src(r), for the training kernel
__global__ __launch_bounds__(256) void k_csum(const float* __restrict__ c, train_segs segs,
                                              float* __restrict__ csum_part, uint32_t* __restrict__ idx) {
    __shared__ float part[4];
    const train_seg sg = segs.s[blockIdx.y];
    float v = 0.0f;
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < sg.n; r += SUM_BLOCKS * 256) {
        const uint32_t src = sg.src(r);
        idx[sg.idx_off + r] = src;
        v += c[src];
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) csum_part[blockIdx.y * SUM_BLOCKS + blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

#ifndef RBW_TRAIN_WAVES
#define RBW_TRAIN_WAVES 8
#endif
constexpr int TR_WAVES = RBW_TRAIN_WAVES;
constexpr int CHUNK = 16 * TR_WAVES;  // records per block iteration
constexpr int XB_PITCH = CHUNK + 8;   // halfs per exchange row
constexpr size_t L_FRAGS = 0;
constexpr size_t L_BIAS = L_FRAGS + size_t(FWD_FRAGS + BWD_FRAGS) * FRAG_HALFS * 2;
constexpr size_t L_XIN = L_BIAS + N_BIAS * 4;
constexpr size_t L_XB_DZ = L_XIN + CHUNK * N_IN * 4;
constexpr size_t L_XB_X = L_XB_DZ + size_t(HIDDEN) * XB_PITCH * 2;
constexpr size_t L_RED = L_XB_X + size_t(HIDDEN) * XB_PITCH * 2;
constexpr size_t TRAIN_LDS = L_RED + 64;  // 130,816 bytes
// the head's staging (raw outputs, then their gradients as halfs) aliases the exchange rows,
// which are free between a chunk's last weight-gradient read and its first exchange write
constexpr size_t L_HEAD = L_XB_DZ;
constexpr size_t L_DZ5 = L_HEAD + size_t(TR_WAVES) * 16 * N_OUT * 4;
static_assert(L_DZ5 + size_t(TR_WAVES) * 16 * N_OUT_PAD * 2 <= L_RED, "head staging inside the exchange rows");
static_assert(TRAIN_LDS <= 160 * 1024, "a CU's LDS");
static_assert(L_XIN % 16 == 0 && L_HEAD % 16 == 0 && L_DZ5 % 16 == 0 && L_XB_DZ % 16 == 0 && L_XB_X % 16 == 0 &&
                  (XB_PITCH * 2) % 16 == 0,
              "16-byte sections");
constexpr int XIN_PER_THREAD = (CHUNK * N_IN + TR_WAVES * 64 - 1) / (TR_WAVES * 64);
// weight-gradient tiles per wave: TR_NW waves share an input tile of layers 1..5 and split its
// output tiles (TR_MS of layers 1..4, TR_S5 of layer 5); layer 0's input tiles go round-robin
constexpr int TR_NW = TR_WAVES / 4;
constexpr int TR_MS = 4 / TR_NW;
constexpr int TR_S5 = (3 + TR_NW - 1) / TR_NW;
constexpr int TR_S0 = (14 + TR_WAVES - 1) / TR_WAVES;
static_assert(TR_WAVES == 4 || TR_WAVES == 8, "four or eight waves");

// The vMF-mixture head's gradient for one record, four lanes per record (lane sub takes
// lobes 2 sub and 2 sub + 1): o the record's raw outputs, d the sampled direction, wr the
// record's weight (radiance / pdf over the batch mean, at most 64).  Writes d loss / d o
// (scaled by wr) for the lane's ten outputs as halfs; returns log q.
__device__ __forceinline__ float mixture_grad4(const float* o, const float* d, float wr, int sub, _Float16* g_out) {
    float a[2], e[2], w[2], kap[2], sig[2], cosv[2], nm[2], mu[2][3], vv[2], om[2];
    bool kcap[2];
#pragma unroll
    for (int t = 0; t < 2; t++) a[t] = o[5 * (2 * sub + t) + 4];
    float amax = fmaxf(a[0], a[1]);
    amax = fmaxf(amax, __shfl_xor(amax, 1));
    amax = fmaxf(amax, __shfl_xor(amax, 2));
#pragma unroll
    for (int t = 0; t < 2; t++) e[t] = hexp(a[t] - amax);
    float ws = e[0] + e[1];
    ws += __shfl_xor(ws, 1);
    ws += __shfl_xor(ws, 2);
    float qp = 0.0f;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const float* oj = o + 5 * (2 * sub + t);
        w[t] = e[t] * hrcp(ws);
        nm[t] = fmaxf(sqrtf(oj[0] * oj[0] + oj[1] * oj[1] + oj[2] * oj[2]), 1e-12f);
        const float inm = hrcp(nm[t]);
        mu[t][0] = oj[0] * inm;
        mu[t][1] = oj[1] * inm;
        mu[t][2] = oj[2] * inm;
        const float sp = hsoftplus(oj[3]) + KAPPA_MIN;
        kcap[t] = sp >= KAPPA_MAX;
        kap[t] = fminf(sp, KAPPA_MAX);
        sig[t] = hrcp(1.0f + hexp(-oj[3]));
        cosv[t] = mu[t][0] * d[0] + mu[t][1] * d[1] + mu[t][2] * d[2];
        // 1 - e^(-2 kappa) in FP32 keeps ~5e-6 relative accuracy down to kappa = KAPPA_MIN
        om[t] = 1.0f - hexp(-2.0f * kap[t]);
        const float norm = kap[t] * hrcp(2.0f * 3.14159265f * om[t]);
        vv[t] = norm * hexp(kap[t] * (cosv[t] - 1.0f));
        qp += w[t] * vv[t];
    }
    float q = qp + __shfl_xor(qp, 1);
    q += __shfl_xor(q, 2);
    q = fmaxf(q, 1e-30f);
    const float gq = -wr;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const float r = w[t] * vv[t] * hrcp(q);
        const float dlv = hrcp(kap[t]) - 2.0f * (1.0f - om[t]) * hrcp(om[t]) + (cosv[t] - 1.0f);
        float gm[3];
#pragma unroll
        for (int c = 0; c < 3; c++) gm[c] = gq * r * kap[t] * d[c];
        const float p = gm[0] * mu[t][0] + gm[1] * mu[t][1] + gm[2] * mu[t][2];
        _Float16* gj = g_out + 5 * (2 * sub + t);
#pragma unroll
        for (int c = 0; c < 3; c++) gj[c] = (_Float16)clamp_h((gm[c] - p * mu[t][c]) * hrcp(nm[t]));
        gj[3] = (_Float16)clamp_h(kcap[t] ? 0.0f : gq * r * dlv * sig[t]);
        gj[4] = (_Float16)clamp_h(gq * (r - w[t]));
    }
    return hlog(q);
}

// One training step's gradients.  Block (x, net) takes chunks of 128 records of network net
// (wave w records 16 w .. 16 w + 15 of a chunk): forward with the inputs of layers 1..5 kept
// in registers, the head's gradient, the data gradients down the layers, and, per layer, the
// weight gradient dW_l += dZ_l^T X_l over the chunk's records: the waves exchange dZ_l and
// X_l through LDS (feature-major, a record per column) and each accumulates its share of
// dW's 16 x 16 tiles in registers across all its chunks (layer 0: input tiles w and w + 8,
// all output tiles; the other layers: input tile w % 4 and half of the output tiles), the
// bias gradients as row sums.  The next chunk's inputs and the head's directions and
// weights are loaded while the current chunk computes.  Every block writes its full partial
// gradient (and loss) to part[net][x]; k_adam_pack adds them.
__global__ __launch_bounds__(TR_WAVES * 64) __attribute__((amdgpu_waves_per_eu(TR_NW, TR_NW))) void k_train(
    const _Float16* __restrict__ pack, const float* __restrict__ bias, const float* __restrict__ inputs,
    const float* __restrict__ dirs, const float* __restrict__ cw, train_segs segs,
    const float* __restrict__ csum_part, const uint32_t* __restrict__ idx, float* __restrict__ part,
    float* __restrict__ loss_part) {
    extern __shared__ __align__(16) unsigned char smem[];
    _Float16* wl = reinterpret_cast<_Float16*>(smem + L_FRAGS);
    float* bl = reinterpret_cast<float*>(smem + L_BIAS);
    float* xin = reinterpret_cast<float*>(smem + L_XIN);
    float* red = reinterpret_cast<float*>(smem + L_RED);
    _Float16* xb_dz = reinterpret_cast<_Float16*>(smem + L_XB_DZ);
    _Float16* xb_x = reinterpret_cast<_Float16*>(smem + L_XB_X);
    const uint32_t net = blockIdx.y;
    const train_seg sg = segs.s[net];
    const uint32_t n = sg.n;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    __builtin_assume(g >= 0 && g < 4);
    const int nq = wave & 3;              // input tile of layers 1..5
    const int mh = TR_MS * (wave >> 2);   // first of the TR_MS output tiles of layers 1..4
    const int m5 = wave >> 2;             // output tiles m5 + TR_NW s (< 3) of layer 5
    // the bias gradients (row sums of dZ) of output tile m are added by the wave with m % 4 == wave % 4
    // among those that read tile m (layers 1..5: the waves with input tile 0 .. TR_NW - 1 wrap)
    float* hs = reinterpret_cast<float*>(smem + L_HEAD) + wave * 16 * N_OUT;
    _Float16* dz5s = reinterpret_cast<_Float16*>(smem + L_DZ5) + wave * 16 * N_OUT_PAD;
    {
        const h8* src = reinterpret_cast<const h8*>(pack + (size_t)net * PACK_HALFS);
        h8* dst = reinterpret_cast<h8*>(wl);
        for (int q = threadIdx.x; q < (FWD_FRAGS + BWD_FRAGS) * 64; q += blockDim.x) dst[q] = src[q];
        for (int q = threadIdx.x; q < N_BIAS; q += blockDim.x) bl[q] = bias[(size_t)net * N_BIAS + q];
    }
    float total = 0.0f;
    for (uint32_t b = 0; b < SUM_BLOCKS; b++) total += csum_part[net * SUM_BLOCKS + b];
    const float cmean = n ? total / (float)n : 0.0f;
    f4 dw0[TR_S0][4], dwh[4][TR_MS], dw5[TR_S5];
    float db0[4], dbh[4][TR_MS], db5[TR_S5];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        db0[m] = 0.0f;
#pragma unroll
        for (int s = 0; s < TR_S0; s++) dw0[s][m] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < TR_MS; s++) {
            dwh[m][s] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            dbh[m][s] = 0.0f;
        }
    }
#pragma unroll
    for (int s = 0; s < TR_S5; s++) {
        dw5[s] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        db5[s] = 0.0f;
    }
    float lsum = 0.0f;
    const uint32_t chunks = (n + CHUNK - 1) / CHUNK;
    // the first chunk's inputs (each thread XIN_PER_THREAD of the chunk's 128 x 9 floats)
    float xpre[XIN_PER_THREAD];
    auto load_inputs = [&](uint32_t ch) {
#pragma unroll
        for (int k = 0; k < XIN_PER_THREAD; k++) {
            const int q = threadIdx.x + k * TR_WAVES * 64;
            xpre[k] = 0.0f;
            if (q < CHUNK * N_IN && ch < chunks) {
                const uint32_t r = ch * CHUNK + (uint32_t)(q / N_IN);
                xpre[k] = inputs[(size_t)idx[sg.idx_off + (r < n ? r : n - 1)] * N_IN + q % N_IN];
            }
        }
    };
    load_inputs(blockIdx.x);
    for (uint32_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const uint32_t r0 = ch * CHUNK;
#pragma unroll
        for (int k = 0; k < XIN_PER_THREAD; k++) {
            const int q = threadIdx.x + k * TR_WAVES * 64;
            if (q < CHUNK * N_IN) xin[q] = xpre[k];
        }
        // the head's record data (four lanes per record), then the next chunk's inputs
        const int rec = lane >> 2, sub = lane & 3;
        const uint32_t rh = r0 + 16 * wave + rec;
        const bool valid = rh < n;
        const uint32_t src = idx[sg.idx_off + (valid ? rh : n - 1)];
        const float d[3] = {dirs[(size_t)src * 3], dirs[(size_t)src * 3 + 1], dirs[(size_t)src * 3 + 2]};
        const float c = cw[src];
        load_inputs(ch + gridDim.x);
        __syncthreads();
        // ---- forward: the wave's 16 records
        const float* xr = xin + (16 * wave + li) * N_IN;
        h8 X[6][2];  // X[l]: the input of layer l (l = 1..5) as B fragments
        f4 acc[4];
        // every layer's sums start from its biases
#pragma unroll
        for (int m = 0; m < 4; m++) acc[m] = *reinterpret_cast<const f4*>(bl + bias_offset(0) + 16 * m + 4 * g);
#pragma unroll
        for (int kb = 0; kb < 7; kb++) {
            const int v = encode_input(kb, g);
            const h8 b = encode_frag(v < N_IN ? xr[v] : 0.0f, kb, g);
#pragma unroll
            for (int m = 0; m < 4; m++) acc[m] = mfma32(frag(wl, fwd_frag(0, m, kb), lane), b, acc[m]);
        }
#pragma unroll
        for (int l = 1; l < N_LAYERS; l++) {
            // ReLU on half pairs (rounding commutes with max(., 0))
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int j = 0; j < 4; j += 2) {
                    h2 p = {(_Float16)acc[m][j], (_Float16)acc[m][j + 1]};
                    p = __builtin_elementwise_max(p, h2{(_Float16)0.0f, (_Float16)0.0f});
                    X[l][m >> 1][(m & 1) * 4 + j] = p.x;
                    X[l][m >> 1][(m & 1) * 4 + j + 1] = p.y;
                }
            const int mt = l == N_LAYERS - 1 ? 3 : 4;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                if (m < mt) {
                    acc[m] = *reinterpret_cast<const f4*>(bl + bias_offset(l) + 16 * m + 4 * g);
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) acc[m] = mfma32(frag(wl, fwd_frag(l, m, kb), lane), X[l][kb], acc[m]);
                }
            }
        }
        // ---- head: raw outputs (biases included) to the wave's staging rows, then four lanes
        // per record
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int col = 16 * m + 4 * g + j;
                if (col < N_OUT) hs[li * N_OUT + col] = acc[m][j];
            }
        wave_sync();
        {
            const float wr = valid && cmean > 0.0f ? fminf(c / cmean, 64.0f) : 0.0f;
            _Float16* grow = dz5s + rec * N_OUT_PAD;
            const float lq = mixture_grad4(hs + rec * N_OUT, d, wr, sub, grow);
            if (sub == 3)
#pragma unroll
                for (int k = N_OUT; k < N_OUT_PAD; k++) grow[k] = (_Float16)0.0f;
            if (sub == 0 && wr > 0.0f) lsum += -(c / total) * lq;
        }
        wave_sync();
        // dZ_5 as B fragments (K = the head's outputs in the chained order, N = records)
        h8 dz[2];
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int f = hid_feat(kb, g, i);
                dz[kb][i] = f < N_OUT_PAD ? dz5s[li * N_OUT_PAD + f] : (_Float16)0.0f;
            }
        __syncthreads();  // every wave's head staging is read before the exchange rows are written
        // ---- backward, one layer at a time from the head: exchange dZ_l and X_l, the next
        // data gradient, then the weight-gradient tiles
#pragma unroll
        for (int l = N_LAYERS - 1; l >= 0; l--) {
            const int mo = l == N_LAYERS - 1 ? 3 : 4;  // output tiles of layer l
            const int col = 16 * wave + li;
#pragma unroll
            for (int kb = 0; kb < 2; kb++)
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int f = hid_feat(kb, g, i);
                    if (f < 16 * mo) xb_dz[f * XB_PITCH + col] = dz[kb][i];
                    if (l >= 1) xb_x[f * XB_PITCH + col] = X[l][kb][i];
                }
            h8 nd[2];
            if (l >= 1) {
                f4 a[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    a[m] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) a[m] = mfma32(frag(wl, bwd_frag(l, m, kb), lane), dz[kb], a[m]);
                }
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int k = (m & 1) * 4 + j;
                        nd[m >> 1][k] = X[l][m >> 1][k] > (_Float16)0.0f ? (_Float16)clamp_h(a[m][j]) : (_Float16)0.0f;
                    }
            }
            __syncthreads();
            // the weight gradients' K dimension: the chunk's records, 32 per product
#pragma unroll
            for (int kb = 0; kb < CHUNK / 32; kb++) {
                const int kofs = 32 * kb + 8 * g;
                if (l >= 1) {
                    const h8 B = lds_h8(xb_x + (16 * nq + li) * XB_PITCH + kofs);
#pragma unroll
                    for (int s = 0; s < (l == N_LAYERS - 1 ? TR_S5 : TR_MS); s++) {
                        const int m = l == N_LAYERS - 1 ? m5 + TR_NW * s : mh + s;
                        if (m < mo) {
                            const h8 A = lds_h8(xb_dz + (16 * m + li) * XB_PITCH + kofs);
                            float rs = 0.0f;
                            if ((m & 3) == nq) {
#pragma unroll
                                for (int i = 0; i < 8; i++) rs += (float)A[i];
                            }
                            if (l == N_LAYERS - 1) {
                                dw5[s] = mfma32(A, B, dw5[s]);
                                db5[s] += rs;
                            } else {
                                dwh[l - 1][s] = mfma32(A, B, dwh[l - 1][s]);
                                dbh[l - 1][s] += rs;
                            }
                        }
                    }
                } else {
                    // layer 0: B is the encoding, feature 16 nt + li (the parameter layout's
                    // order, 24 v + 2 k + (0 sin | 1 cos)) of records 32 kb + 8 g + i
                    h8 A[4];
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        A[m] = lds_h8(xb_dz + (16 * m + li) * XB_PITCH + kofs);
                        if (m == wave % 4 && wave < 4) {
                            float rs = 0.0f;
#pragma unroll
                            for (int i = 0; i < 8; i++) rs += (float)A[m][i];
                            db0[m] += rs;
                        }
                    }
#pragma unroll
                    for (int s = 0; s < TR_S0; s++) {
                        const int nt = wave + TR_WAVES * s;
                        if (nt >= 14) break;
                        const int f = 16 * nt + li;
                        h8 B;
                        if (f < N_FEAT) {
                            const int v = f / 24, k = (f % 24) >> 1;
                            // cos(2 pi t) = sin(2 pi (t + 1/4)): one transcendental per lane
                            const float shift = (f & 1) ? 0.25f : 0.0f;
                            const float scale = (float)(1 << k) * 0.5f;
#pragma unroll
                            for (int i = 0; i < 8; i++) {
                                float t = xin[(kofs + i) * N_IN + v] * scale;
                                t = t - floorf(t) + shift;
                                B[i] = (_Float16)sin_rev(t - floorf(t));
                            }
                        } else {
#pragma unroll
                            for (int i = 0; i < 8; i++) B[i] = (_Float16)0.0f;
                        }
#pragma unroll
                        for (int m = 0; m < 4; m++) dw0[s][m] = mfma32(A[m], B, dw0[s][m]);
                    }
                }
            }
            __syncthreads();
            if (l >= 1) {
                dz[0] = nd[0];
                dz[1] = nd[1];
            }
        }
    }
    // ---- this block's partial gradient (every parameter written) and loss
    float* P = part + ((size_t)net * gridDim.x + blockIdx.x) * N_PARAMS;
#pragma unroll
    for (int s = 0; s < TR_S0; s++) {
        const int nt = wave + TR_WAVES * s;
        if (nt >= 14) break;
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                P[layer_w_offset(0) + (16 * m + 4 * g + j) * N_FEAT_PAD + 16 * nt + li] = dw0[s][m][j];
    }
#pragma unroll
    for (int l = 1; l < N_LAYERS; l++)
#pragma unroll
        for (int s = 0; s < (l == N_LAYERS - 1 ? TR_S5 : TR_MS); s++) {
            const int m = l == N_LAYERS - 1 ? m5 + TR_NW * s : mh + s;
            if (m >= (l == N_LAYERS - 1 ? 3 : 4)) continue;
            const f4 t = l == N_LAYERS - 1 ? dw5[s] : dwh[l - 1][s];
#pragma unroll
            for (int j = 0; j < 4; j++) P[layer_w_offset(l) + (16 * m + 4 * g + j) * HIDDEN + 16 * nq + li] = t[j];
            if ((m & 3) == nq) {
                float v = l == N_LAYERS - 1 ? db5[s] : dbh[l - 1][s];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (g == 0) P[layer_b_offset(l) + 16 * m + li] = v;
            }
        }
    if (wave < 4) {
        const int m = wave;
        float v = db0[m];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (g == 0) P[layer_b_offset(0) + 16 * m + li] = v;
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) red[wave] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < TR_WAVES; w++) t += red[w];
        loss_part[(size_t)net * gridDim.x + blockIdx.x] = t;
    }
}

// ------------------------------------------------------------------ Adam and packing

__device__ __forceinline__ bool live_param(uint32_t p) {
    // padded input columns of layer 0 and padded outputs of the last layer stay zero
    if (p < (uint32_t)layer_b_offset(0)) return (p % N_FEAT_PAD) < N_FEAT;
    const uint32_t w5 = layer_w_offset(N_LAYERS - 1), b5 = layer_b_offset(N_LAYERS - 1);
    if (p >= w5 && p < b5) return (p - w5) / HIDDEN < N_OUT;
    if (p >= b5) return p - b5 < N_OUT;
    return true;
}

// position (K block, lane group, element) of a chained layer's input feature f
__device__ __forceinline__ void chained_pos(int f, int& kb, int& g, int& i) {
    kb = f >> 5;
    const int r = f & 31;
    g = (r & 15) >> 2;
    i = (r & 3) + (r >= 16 ? 4 : 0);
}

// writes parameter p's half (and FP32 bias) where the kernels read it
__device__ __forceinline__ void pack_param(uint32_t p, float w, _Float16* __restrict__ pk, float* __restrict__ bs) {
    int l = 0;
    while (l < N_LAYERS - 1 && p >= (uint32_t)layer_w_offset(l + 1)) l++;
    const uint32_t wo = layer_w_offset(l), bo = layer_b_offset(l);
    if (p >= bo) {
        bs[bias_offset(l) + (p - bo)] = w;
        return;
    }
    const int in = layer_in(l);
    const int o = (int)((p - wo) / in), i = (int)((p - wo) % in);
    const _Float16 h = (_Float16)w;
    int kb, g, e;
    if (l == 0) {
        // feature i = 24 v + 2 k + sc sits in quad Q = 3 v + k / 4 (kb = Q / 4, g = Q % 4) at 2 (k % 4) + sc
        const int v = i / 24, k = (i % 24) >> 1, sc = i & 1;
        const int Q = 3 * v + (k >> 2);
        kb = Q >> 2;
        g = Q & 3;
        e = 2 * (k & 3) + sc;
    } else {
        chained_pos(i, kb, g, e);
    }
    pk[(size_t)(fwd_frag(l, o >> 4, kb) * 64 + g * 16 + (o & 15)) * 8 + e] = h;
    if (l >= 1) {
        // W_l^T: row = input feature i, K = output feature o in the chained order
        chained_pos(o, kb, g, e);
        pk[(size_t)(bwd_frag(l, i >> 4, kb) * 64 + g * 16 + (i & 15)) * 8 + e] = h;
    }
}

// mode 0: pack w32 into the kernels' layout; 1: gradient (sum of the blocks' partials over
// n) and Adam step, then pack; 2: gradient only.  A block takes 8 parameter quads: its 256
// threads add the partial rows of a quad in 32 interleaved splits with 16-byte loads (enough
// loads in flight to stream the partials), then one thread per quad adds the splits and
// updates the four parameters.  Block (0, net) also adds the loss partials.
constexpr int RED_QUADS = 8;
constexpr int RED_SPLITS = 32;
constexpr uint32_t ADAM_BLOCKS = (N_PARAMS / 4 + RED_QUADS - 1) / RED_QUADS;
static_assert(N_PARAMS % 4 == 0, "parameter quads");
__global__ __launch_bounds__(RED_QUADS * RED_SPLITS) void k_adam_pack(
    float* __restrict__ w32, float* __restrict__ m, float* __restrict__ v, _Float16* __restrict__ pack,
    float* __restrict__ bias, float* __restrict__ grad, float* __restrict__ loss, const float* __restrict__ part,
    const float* __restrict__ loss_part, uint32_t blocks, train_segs segs, float lr, float b1, float b2, float eps,
    uint32_t mode) {
    __shared__ float4 red[RED_SPLITS][RED_QUADS];
    const uint32_t net = blockIdx.y;
    const train_seg sg = segs.s[net];
    if (mode != 0 && sg.n == 0) return;
    const int qi = threadIdx.x % RED_QUADS, sp = threadIdx.x / RED_QUADS;
    const uint32_t p0 = (blockIdx.x * RED_QUADS + qi) * 4;
    if (mode != 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        float s = 0.0f;
        for (uint32_t b = 0; b < blocks; b++) s += loss_part[(size_t)net * blocks + b];
        loss[net] = s;
    }
    if (mode != 0) {
        float4 gs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (p0 < (uint32_t)N_PARAMS) {
            const float* pp = part + (size_t)net * blocks * N_PARAMS + p0;
#pragma unroll 4
            for (uint32_t b = sp; b < blocks; b += RED_SPLITS) {
                const float4 t = *reinterpret_cast<const float4*>(pp + (size_t)b * N_PARAMS);
                gs.x += t.x;
                gs.y += t.y;
                gs.z += t.z;
                gs.w += t.w;
            }
        }
        red[sp][qi] = gs;
        __syncthreads();
    }
    if (sp != 0 || p0 >= (uint32_t)N_PARAMS) return;
    float gq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (mode != 0) {
        for (int k = 0; k < RED_SPLITS; k++) {
            const float4 t = red[k][qi];
            gq[0] += t.x;
            gq[1] += t.y;
            gq[2] += t.z;
            gq[3] += t.w;
        }
    }
    float* W = w32 + (size_t)net * N_PARAMS;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint32_t p = p0 + e;
        float w = W[p];
        const bool live = live_param(p);
        if (mode != 0) {
            const float gr = live ? gq[e] / (float)sg.n : 0.0f;
            grad[(size_t)net * N_PARAMS + p] = gr;
            if (mode == 2 || !live) continue;
            const size_t q = (size_t)net * N_PARAMS + p;
            const float mm = b1 * m[q] + (1.0f - b1) * gr;
            const float vv = b2 * v[q] + (1.0f - b2) * gr * gr;
            m[q] = mm;
            v[q] = vv;
            w = w - lr * (mm / sg.c1) / (sqrtf(vv / sg.c2) + eps);
            W[p] = w;
        }
        if (live) pack_param(p, w, pack + (size_t)net * PACK_HALFS, bias + (size_t)net * N_BIAS);
    }
}

// ------------------------------------------------------------------ host

void model_free(model& M) {
    for (void* p : {(void*)M.w32, (void*)M.m, (void*)M.v, (void*)M.pack, (void*)M.bias, (void*)M.grad,
                    (void*)M.loss, (void*)M.part, (void*)M.loss_part, (void*)M.csum_part, (void*)M.idx})
        if (p) (void)hipFree(p);
    M = model();
}

static uint32_t train_blocks(uint32_t nets) {
    const uint32_t b = 256 / (nets ? nets : 1);
    return b < 4 ? 4 : b;
}

static train_segs no_segs() {
    train_segs t;
    std::memset(&t, 0, sizeof(t));
    return t;
}

static void pack_net(model& M, uint32_t net, hipStream_t s) {
    train_segs t = no_segs();
    hipLaunchKernelGGL(k_adam_pack, dim3(ADAM_BLOCKS, net + 1), dim3(RED_QUADS * RED_SPLITS), 0, s, M.w32, M.m, M.v, M.pack,
                       M.bias, M.grad, M.loss, M.part, M.loss_part, train_blocks(M.nets), t, 0.0f, 0.0f, 0.0f, 0.0f,
                       0u);
    GCHECK(hipGetLastError());
}

void model_set(model& M, const float* host, hipStream_t s, uint32_t net) {
    if (net >= M.nets) throw std::runtime_error("no such network");
    const size_t off = (size_t)net * N_PARAMS;
    GCHECK(hipMemcpyAsync(M.w32 + off, host, sizeof(float) * N_PARAMS, hipMemcpyHostToDevice, s));
    GCHECK(hipMemsetAsync(M.m + off, 0, sizeof(float) * N_PARAMS, s));
    GCHECK(hipMemsetAsync(M.v + off, 0, sizeof(float) * N_PARAMS, s));
    pack_net(M, net, s);  // packs networks 0..net (unchanged ones to the same values)
    M.step[net] = 0;
}

void model_get(model& M, float* host, hipStream_t s, uint32_t net) {
    if (net >= M.nets) throw std::runtime_error("no such network");
    GCHECK(hipMemcpyAsync(host, M.w32 + (size_t)net * N_PARAMS, sizeof(float) * N_PARAMS, hipMemcpyDeviceToHost, s));
    GCHECK(hipStreamSynchronize(s));
}

void prepare_kernels();
static void allow_lds() { prepare_kernels(); }

// the kernels' dynamic LDS above the default limit of a launch (before any graph capture)
void prepare_kernels() {
    static bool done = false;
    if (done) return;
    GCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_infer<INF_R>), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(INF_LDS)));
    GCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_train), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(TRAIN_LDS)));
    done = true;
}

void model_init(model& M, uint32_t seed, hipStream_t s, uint32_t nets) {
    if (nets == 0 || nets > MAX_NETS) throw std::runtime_error("a model holds 1 to 32 networks");
    if (M.nets != nets) {
        model_free(M);
        M.nets = nets;
        const size_t np = (size_t)nets * N_PARAMS;
        GCHECK(hipMalloc(&M.w32, sizeof(float) * np));
        GCHECK(hipMalloc(&M.m, sizeof(float) * np));
        GCHECK(hipMalloc(&M.v, sizeof(float) * np));
        GCHECK(hipMalloc(&M.grad, sizeof(float) * np));
        GCHECK(hipMalloc(&M.pack, sizeof(_Float16) * (size_t)nets * PACK_HALFS));
        GCHECK(hipMalloc(&M.bias, sizeof(float) * (size_t)nets * N_BIAS));
        GCHECK(hipMalloc(&M.loss, sizeof(float) * nets));
        const size_t nb = (size_t)nets * train_blocks(nets);
        GCHECK(hipMalloc(&M.part, sizeof(float) * nb * N_PARAMS));
        GCHECK(hipMalloc(&M.loss_part, sizeof(float) * nb));
        GCHECK(hipMalloc(&M.csum_part, sizeof(float) * (size_t)nets * SUM_BLOCKS));
        // the padded positions of the packed weights stay zero
        GCHECK(hipMemsetAsync(M.pack, 0, sizeof(_Float16) * (size_t)nets * PACK_HALFS, s));
        GCHECK(hipMemsetAsync(M.grad, 0, sizeof(float) * np, s));
        GCHECK(hipMemsetAsync(M.loss, 0, sizeof(float) * nets, s));
    }
    M.step.assign(nets, 0);
    allow_lds();
    // He-uniform, zero biases; the last layer at half scale; network k from seed + k
    std::vector<float> p(N_PARAMS);
    for (uint32_t k = 0; k < nets; k++) {
        std::fill(p.begin(), p.end(), 0.0f);
        pcg32 gen;
        gen.seed(seed + k);
        for (int l = 0; l < N_LAYERS; l++) {
            const int fan_in = l == 0 ? N_FEAT : HIDDEN;
            const int outs = l == N_LAYERS - 1 ? N_OUT : HIDDEN;
            const float lim = std::sqrt(6.0f / fan_in) * (l == N_LAYERS - 1 ? 0.5f : 1.0f);
            for (int o = 0; o < outs; o++)
                for (int i = 0; i < (l == 0 ? N_FEAT : HIDDEN); i++)
                    p[layer_w_offset(l) + o * layer_in(l) + i] = (2.0f * gen.uniform() - 1.0f) * lim;
        }
        GCHECK(hipMemcpyAsync(M.w32 + (size_t)k * N_PARAMS, p.data(), sizeof(float) * N_PARAMS, hipMemcpyHostToDevice,
                              s));
        GCHECK(hipStreamSynchronize(s));  // p is reused for the next network
    }
    GCHECK(hipMemsetAsync(M.m, 0, sizeof(float) * (size_t)nets * N_PARAMS, s));
    GCHECK(hipMemsetAsync(M.v, 0, sizeof(float) * (size_t)nets * N_PARAMS, s));
    pack_net(M, nets - 1, s);
    GCHECK(hipStreamSynchronize(s));
}

// persistent inference blocks: two fit a CU's LDS
static uint32_t infer_blocks(uint32_t rows) {
    const uint32_t want = (rows + INF_STEP_ROWS - 1) / INF_STEP_ROWS;
    return want < INF_MAX_BLOCKS ? (want ? want : 1) : INF_MAX_BLOCKS;
}

void forward(model& M, const float* inputs, uint32_t n, float* out, hipStream_t s) {
    if (n == 0) return;
    allow_lds();
    hipLaunchKernelGGL(k_infer<INF_R>, dim3(infer_blocks(n)), dim3(INF_WAVES * 64), INF_LDS, s, M.pack, M.bias, inputs, n,
                       nullptr, 1u, 0u, out);
    GCHECK(hipGetLastError());
}

void forward_dev(model& M, const float* inputs, const uint32_t* n_dev, uint32_t n_max, float* out, hipStream_t s) {
    if (n_max == 0) return;
    allow_lds();
    hipLaunchKernelGGL(k_infer<INF_R>, dim3(infer_blocks(n_max)), dim3(INF_WAVES * 64), INF_LDS, s, M.pack, M.bias,
                       inputs, 0u, n_dev, 1u, 0u, out);
    GCHECK(hipGetLastError());
}

void forward_segments(model& M, const float* inputs, const uint32_t* n_dev, uint32_t stride, uint32_t n_max,
                      float* out, hipStream_t s, const uint32_t* rows, uint32_t nets) {
    if (n_max == 0) return;
    if (nets == 0 || nets > M.nets) nets = M.nets;
    allow_lds();
    // (the rows of all networks together number at most n_max * nets)
    hipLaunchKernelGGL(k_infer<INF_R>, dim3(infer_blocks((uint32_t)std::min<uint64_t>(uint64_t(n_max) * nets, 1u << 30))),
                       dim3(INF_WAVES * 64), INF_LDS, s, M.pack, M.bias, inputs, 0u, n_dev, nets, stride, out, rows);
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

void train_step_batched(model& M, const float* inputs, const float* dirs, const float* weights,
                        const std::vector<train_rows>& rows, const adam_opts& opt, bool apply, hipStream_t s) {
    if (rows.size() != M.nets) throw std::runtime_error("one row set per network");
    train_segs t = no_segs();
    bool any = false;
    for (uint32_t k = 0; k < M.nets; k++) {
        const train_rows& r = rows[k];
        train_seg& sg = t.s[k];
        sg.base = r.base;
        sg.n = r.n;
        sg.n_src = r.map.n_src ? r.map.n_src : r.n;
        sg.a = r.map.n_src ? r.map.a : 1u;
        sg.b = r.map.n_src ? r.map.b : 0u;
        if (r.n > sg.n_src) throw std::runtime_error("training batch larger than the sample set");
        if (r.n && apply) M.step[k]++;
        sg.c1 = 1.0f - std::pow(opt.b1, (float)M.step[k]);
        sg.c2 = 1.0f - std::pow(opt.b2, (float)M.step[k]);
        any = any || r.n;
    }
    if (!any) return;
    uint32_t rows_total = 0;
    for (uint32_t k = 0; k < M.nets; k++) {
        t.s[k].idx_off = rows_total;
        rows_total += t.s[k].n;
    }
    if (rows_total > M.idx_cap) {  // grows once to the largest batch a model trains on
        if (M.idx) GCHECK(hipFree(M.idx));
        M.idx = nullptr;
        GCHECK(hipMalloc(&M.idx, sizeof(uint32_t) * rows_total));
        M.idx_cap = rows_total;
    }
    allow_lds();
    const uint32_t blocks = train_blocks(M.nets);
    hipLaunchKernelGGL(k_csum, dim3(SUM_BLOCKS, M.nets), dim3(256), 0, s, weights, t, M.csum_part, M.idx);
    hipLaunchKernelGGL(k_train, dim3(blocks, M.nets), dim3(TR_WAVES * 64), TRAIN_LDS, s, M.pack, M.bias, inputs, dirs,
                       weights, t, M.csum_part, M.idx, M.part, M.loss_part);
    hipLaunchKernelGGL(k_adam_pack, dim3(ADAM_BLOCKS, M.nets), dim3(RED_QUADS * RED_SPLITS), 0, s, M.w32, M.m, M.v, M.pack,
                       M.bias, M.grad, M.loss, M.part, M.loss_part, blocks, t, opt.lr, opt.b1, opt.b2, opt.eps,
                       apply ? 1u : 2u);
    GCHECK(hipGetLastError());
}

void train_step(model& M, const float* inputs, const float* dirs, const float* weights, uint32_t n,
                const adam_opts& opt, bool apply, hipStream_t s, sample_map map) {
    if (n == 0) return;
    std::vector<train_rows> rows(M.nets);
    rows[0].n = n;
    rows[0].map = map;
    train_step_batched(M, inputs, dirs, weights, rows, opt, apply, s);
}

}  // namespace guide
}  // namespace rbw
