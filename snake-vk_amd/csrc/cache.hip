// Neural radiance cache: hash-grid encoding, fused MLP forward / backward on
// MFMA (v_mfma_f32_16x16x32_f16, FP16 in, FP32 accumulate), Adam, and the
// cache render mode (paths terminate into the cache at a fixed depth while a
// subset of paths trains it online).  No reference counterpart.
//
// Fragment layout of v_mfma_f32_16x16x32_f16 (checked on the device): lane l
// holds A[l%16][8(l/16) .. 8(l/16)+7], B[8(l/16) .. +7][l%16] and
// D[4(l/16) .. 4(l/16)+3][l%16].
//
// Determinism: MLP weight gradients are written as per-wave partials and summed
// in a fixed order; hash-grid gradients are accumulated with 64-bit fixed-point
// integer atomics (order independent).  Same seed -> same bits.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "cache_nn.h"
#include "render_modes.h"

namespace snkpt {

struct CacheState {
    bool ready = false;
    GridGeom geom{};
    DeviceBuffer adam;        // AdamDev: the optimiser's step counter and the step's factors
    DeviceBuffer grid32, grid16, grid_m, grid_v, grid_grad;  // grid_grad: int64 fixed point
    DeviceBuffer mlp32, mlp16, mlp16t, mlp_m, mlp_v, mlp_gradf, grid_gradf;
    // exponential moving averages of the weights (fp32 and the fp16 copies queries read)
    DeviceBuffer grid_e32, grid_e16, mlp_e32, mlp_e16;
    DeviceBuffer partial, loss_part, x0, dx0, y, mlp_red, grid_priv;
    uint32_t cap = 0;         // samples the inference work buffers hold
    uint32_t tcap = 0;        // samples the training work buffers hold
    uint32_t last_n = 0;
    uint32_t t_max = 0;       // most training samples the last frame's paths may emit
    // render integration
    DeviceBuffer q_samples, q_meta, train_vtx, train_meta, t_samples, counts;
    DeviceBuffer tcounts, tstats;  // the training paths' own counters and ray statistics
    uint32_t rcap = 0;
    // side stream on which a frame's training paths are traced while the render pass runs
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    ~CacheState() {
        if (side) (void)hipStreamDestroy(side);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (join_ev) (void)hipEventDestroy(join_ev);
    }
};
void cache_state_free(CacheState* c) { delete c; }

namespace {

constexpr int kRow = 72;  // LDS row stride in halves (64 + 8: 144 B rows, 16 B aligned)
constexpr int kWTRow4 = 32;  // W4^T row length (16 outputs padded to the MFMA K of 32)
constexpr int kMlpTParams = kHidden * kLayerParams + kWidth * kWTRow4;  // 18432
constexpr double kFixScale = 4294967296.0;  // 2^32 fixed point for grid gradients

// The optimiser's state that changes from step to step lives on the device, so
// that the replays of a captured frame keep counting steps: k_adam_advance
// moves the counter and sets the step's bias-corrected rate and weight-average
// decay before the Adam kernels read them.
struct AdamDev {
    uint32_t t;     // steps taken
    uint32_t live;  // this step has samples (0: the Adam kernels leave everything as it is)
    float lr_t;     // lr * sqrt(1 - b2^t) / (1 - b1^t)
    float ema;      // weight-average decay of this step (0: the average is set to the weights)
    uint32_t gmax;  // bits of the step's largest |dL/dfeature| on the hashed levels (sets their fixed point)
    uint32_t pad[3];
};

__global__ void k_adam_reset(AdamDev* a) {
    if (threadIdx.x == 0) *a = AdamDev{0u, 0u, 0.0f, 0.0f, 0u, {0u, 0u, 0u}};
}

// Fixed point of the hashed levels' gradients: both features of a corner go
// into ONE 64-bit atomic as two 32-bit halves (f0 * 2^32 + f1; the integer sum
// of such words is H * 2^32 + L with H, L the sums of the halves, which
// decode exactly while each stays in int32).  The scale puts the step's largest
// |dL/dfeature| at 2^22, so each contribution has |v| <= 2^22 and an entry
// takes 512 same-signed largest contributions before a half could overflow (a
// hashed entry gets a few per step).  Resolution: 2^-22 of the largest.
__host__ __device__ __forceinline__ float packed_scale(uint32_t gmax_bits) {
    const float gm = __uint_as_float(gmax_bits);
    if (!(gm > 0.0f)) return 0.0f;
    int ex;
    (void)frexpf(gm, &ex);  // gm < 2^ex
    return ldexpf(1.0f, min(max(22 - ex, -126), 126));
}

__global__ void k_adam_advance(AdamDev* a, uint32_t n, const uint32_t* __restrict__ n_dev, float lr, float ema,
                               float b1, float b2, int apply) {
    if (threadIdx.x != 0) return;
    if (n_dev) n = min(n, *n_dev);
    AdamDev d = *a;
    d.live = n > 0u ? 1u : 0u;
    if (d.live && apply) d.t += 1u;
    const uint32_t t = d.t ? d.t : 1u;
    d.lr_t = (float)((double)lr * sqrt(1.0 - pow((double)b2, (double)t)) / (1.0 - pow((double)b1, (double)t)));
    // decay ramp (1 + t) / (10 + t) so early averages are not dominated by the initial weights
    d.ema = fminf(ema, (1.0f + (float)t) / (10.0f + (float)t));
    d.gmax = 0u;
    *a = d;
}


__device__ __forceinline__ f32x4_t mfma16(half8_t a, half8_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------- encoding ----
typedef _Float16 half2v_t __attribute__((ext_vector_type(2)));

// Position normalised to the cache box, clamped to [0, 1].
__device__ __forceinline__ void norm_pos(const float* s, const GridGeom& G, float p[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = fminf(fmaxf((s[a] - G.lo[a]) * G.inv_ext, 0.0f), 1.0f);
}

// Cell of p at resolution r and the trilinear fractions; the cell is clamped
// so that corner + 1 <= r (p may be exactly 1).
struct Cell {
    uint32_t ix, iy, iz;
    float wx, wy, wz;
};
__device__ __forceinline__ Cell cell_of(const float p[3], float r) {
    float x = p[0] * r, y = p[1] * r, z = p[2] * r;
    float fx = fminf(floorf(x), r - 1.0f), fy = fminf(floorf(y), r - 1.0f), fz = fminf(floorf(z), r - 1.0f);
    return Cell{(uint32_t)fx, (uint32_t)fy, (uint32_t)fz, x - fx, y - fy, z - fz};
}
__device__ __forceinline__ float corner_weight(const Cell& c, int k) {
    return ((k & 1) ? c.wx : 1.0f - c.wx) * (((k >> 1) & 1) ? c.wy : 1.0f - c.wy) * ((k >> 2) ? c.wz : 1.0f - c.wz);
}
__device__ __forceinline__ uint32_t corner_index(const Cell& c, int k, const GridGeom& G, int l) {
    return grid_index(c.ix + (k & 1), c.iy + ((k >> 1) & 1), c.iz + (k >> 2), G.res[l], G.dense[l]);
}

// The two features of level l: trilinear blend of the 8 corner entries.
__device__ __forceinline__ float2 encode_level(const float p[3], const GridGeom& G, int l,
                                               const __half2* __restrict__ grid) {
    const Cell c = cell_of(p, (float)G.res[l]);
    const __half2* g = grid + (size_t)l * kGridT;
    float2 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = __half22float2(g[corner_index(c, k, G, l)]);
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float w = corner_weight(c, k);
        f0 += w * v[k].x;
        f1 += w * v[k].y;
    }
    return make_float2(f0, f1);
}

// Input feature 32 + k (k < 32) of a sample: direction and normal mapped to
// [0, 1], roughness, albedo, sin/cos(pi x) of each direction and normal
// component, zero padding.
__device__ __forceinline__ float extra_feature(const float* s, int k) {
    const float* d = s + 3;
    const float* n = s + 6;
    if (k < 3) return 0.5f * (d[k] + 1.0f);
    if (k < 6) return 0.5f * (n[k - 3] + 1.0f);
    if (k == 6) return s[9];
    if (k < 10) return s[10 + (k - 7)];
    if (k < 16) {
        float a = kPi * d[(k - 10) >> 1];
        return ((k - 10) & 1) ? cosf(a) : sinf(a);
    }
    if (k < 22) {
        float a = kPi * n[(k - 16) >> 1];
        return ((k - 16) & 1) ? cosf(a) : sinf(a);
    }
    return 0.0f;
}

// One thread per (sample, level): wave l of a block encodes level l of 64
// consecutive samples and writes input features 32 + 2l, 33 + 2l as well, so
// the 16 waves of a block fill the 128 B input rows of their 64 samples.
constexpr int kEncBlock = 64 * kGridLevels;
// Rows [64 blk, 64 blk + 64) of x0, for blocks of kEncBlock threads.
__device__ __forceinline__ void encode_block(const float* __restrict__ samples, uint32_t n, const GridGeom& G,
                                             const __half2* __restrict__ grid, _Float16* __restrict__ x0) {
    // the 64 input rows of the block are assembled in LDS and written out as
    // whole 128 B rows (4 B stores at a 128 B stride would be 64 partial lines)
    __shared__ half2v_t tile[64][33];
    const int l = threadIdx.x >> 6, r = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * 64 + r;
    if (i < n) {
        const float* s = samples + (size_t)kSampleFloats * i;
        float p[3];
        norm_pos(s, G, p);
        float2 f = encode_level(p, G, l, grid);
        half2v_t v, e;
        v[0] = (_Float16)f.x;
        v[1] = (_Float16)f.y;
        e[0] = (_Float16)extra_feature(s, 2 * l);
        e[1] = (_Float16)extra_feature(s, 2 * l + 1);
        tile[r][l] = v;
        tile[r][16 + l] = e;
    }
    __syncthreads();
    // 64 rows x 32 half2 = 2048 words: two per thread, consecutive threads on consecutive words
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int wdx = threadIdx.x + k * kEncBlock, row = wdx >> 5, col = wdx & 31;
        const uint32_t gi = blockIdx.x * 64 + row;
        if (gi < n) *(half2v_t*)(x0 + (size_t)64 * gi + 2 * col) = tile[row][col];
    }
}

__global__ __launch_bounds__(kEncBlock) void k_encode(const float* __restrict__ samples, uint32_t n,
                                                     const uint32_t* __restrict__ n_dev, GridGeom G,
                                                     const __half2* __restrict__ grid, _Float16* __restrict__ x0) {
    if (n_dev) n = min(n, *n_dev);
    if (blockIdx.x * 64 >= n) return;  // the whole block is past the end
    encode_block(samples, n, G, grid, x0);
}
inline uint32_t enc_blocks(uint32_t n) { return n == 0 ? 1u : (n + 63) / 64; }

// Network of the 64-row-aligned segment holding `row`: segments [off[k], off[k+1])
// of the rows sorted by scene, each a multiple of 64 rows.
__device__ __forceinline__ uint32_t segment_of(const uint32_t* __restrict__ off, uint32_t nseg, uint32_t row) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= row)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Encoding of queries grouped by network: rows [off[k], off[k+1]) use network
// k's grid; off[nseg] is the number of rows (a multiple of 64, at most n).
__global__ __launch_bounds__(kEncBlock) void k_encode_many(const float* __restrict__ samples, uint32_t n,
                                                          const uint32_t* __restrict__ off, uint32_t nseg,
                                                          const CacheRefD* __restrict__ refs,
                                                          _Float16* __restrict__ x0) {
    n = min(n, off[nseg]);
    if (blockIdx.x * 64 >= n) return;
    const CacheRefD& R = refs[segment_of(off, nseg, blockIdx.x * 64)];
    encode_block(samples, n, R.geom, R.grid, x0);
}

// ------------------------------------------------------------ MLP core ----
// act: [64][kRow] halves in LDS.  W: [out][64] row-major fp16 in global memory.
__device__ __forceinline__ half8_t a_frag(const _Float16* act, int m, int kb, int lane) {
    return *(const half8_t*)(act + (m * 16 + (lane & 15)) * kRow + kb * 32 + 8 * (lane >> 4));
}
__device__ __forceinline__ half8_t w_frag(const _Float16* __restrict__ W, int n, int kb, int lane, int row_len) {
    return *(const half8_t*)(W + (n * 16 + (lane & 15)) * row_len + kb * 32 + 8 * (lane >> 4));
}

// Output layer: 64 samples x 16 (3 used) outputs, no activation.  acc[m][e] is
// sample m*16 + 4(lane/16) + e, channel lane%16.
__device__ __forceinline__ void layer_out(const _Float16* act_in, const _Float16* __restrict__ W4, f32x4_t acc[4],
                                          int lane) {
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        half8_t b = w_frag(W4, 0, kb, lane, kWidth);
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = mfma16(a_frag(act_in, m, kb, lane), b, acc[m]);
    }
}

__device__ __forceinline__ void load_x0_tile(const _Float16* __restrict__ x0, uint32_t base, uint32_t n,
                                             _Float16* act, int lane) {
    // 64 rows x 64 halves; lane copies row `lane`
    uint32_t q = base + lane;
    const half8_t* src = (const half8_t*)(x0 + (size_t)64 * q);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        half8_t v = q < n ? src[j] : half8_t{0, 0, 0, 0, 0, 0, 0, 0};
        *(half8_t*)(act + lane * kRow + 8 * j) = v;
    }
}

constexpr int kInferWaves = 4;

// ReLU of FP16 values as a signed 16-bit max with 0 (a negative half has its
// sign bit set, so it reads as a negative integer), two halves per instruction:
// equal to rounding max(x, 0) to FP16 (-0 becomes +0).
typedef _Float16 half4v_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ half4v_t relu_h4(half4v_t h) {
    typedef short short4_t __attribute__((ext_vector_type(4)));
    short4_t v = __builtin_bit_cast(short4_t, h);
    v = v > (short)0 ? v : (short)0;
    return __builtin_bit_cast(half4v_t, v);
}

// Hidden layer: out = relu(in * W^T), 64 samples x 64 outputs, computed as
// out^T = W * in^T (A = weight rows, B = activation rows): a lane's four
// accumulators are four consecutive outputs of one sample and leave as one 8 B
// LDS store (16 per lane and layer, not 64 two-byte stores).  All of the wave's
// input fragments are read before the first MFMA, which the stores depend on,
// so out may be in (one tile per wave).
__device__ __forceinline__ void layer_hidden(const _Float16* in, const _Float16* __restrict__ W, _Float16* out,
                                             int lane) {
    f32x4_t acc[4][4];  // [output tile n][sample tile m]
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[n][m] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    half8_t x[2][4];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int m = 0; m < 4; ++m) x[kb][m] = a_frag(in, m, kb, lane);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const half8_t wn = w_frag(W, n, kb, lane, kWidth);
#pragma unroll
            for (int m = 0; m < 4; ++m) acc[n][m] = mfma16(wn, x[kb][m], acc[n][m]);
        }
    }
    typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const half4_t o = relu_h4(__builtin_convertvector(acc[n][m], half4_t));
            const int sample = m * 16 + (lane & 15), col = n * 16 + 4 * (lane >> 4);
            *(half4_t*)(out + sample * kRow + col) = o;
        }
}

// 64 rows per wave from `base` through the network W (rows at or past n read as zero).
__device__ __forceinline__ void infer_block(const _Float16* __restrict__ x0, uint32_t n, uint32_t base,
                                            const _Float16* __restrict__ W, float4* __restrict__ y) {
    __shared__ __attribute__((aligned(16))) _Float16 buf[kInferWaves][64 * kRow];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    _Float16* a0 = buf[wv];
    load_x0_tile(x0, base, n, a0, lane);
    __syncthreads();
#pragma unroll
    for (int l = 0; l < kHidden; ++l) {
        layer_hidden(a0, W + l * kLayerParams, a0, lane);
        __syncthreads();
    }
    // output layer, transposed as well: lanes 0..15 hold outputs 0..3 of samples
    // m * 16 + lane in acc[m][0..3]
    f32x4_t acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const half8_t wn = w_frag(W + 4 * kLayerParams, 0, kb, lane, kWidth);
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = mfma16(wn, a_frag(a0, m, kb, lane), acc[m]);
    }
    if (lane < 16) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t q = base + m * 16 + lane;
            if (q < n) y[q] = make_float4(acc[m][0], acc[m][1], acc[m][2], 0.0f);
        }
    }
}

// Inference: y[q] = MLP(x0[q]) for 64 samples per wave.
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(const _Float16* __restrict__ x0, uint32_t n,
                                                                const uint32_t* __restrict__ n_dev,
                                                                const _Float16* __restrict__ W,
                                                                float4* __restrict__ y) {
    if (n_dev) n = min(n, *n_dev);
    if (blockIdx.x * kInferWaves * 64 >= n) return;  // the whole block is past the end
    // waves past n run on zero rows
    infer_block(x0, n, (blockIdx.x * kInferWaves + (threadIdx.x >> 6)) * 64, W, y);
}

// Inference of queries grouped by network (segments as for k_encode_many): each
// wave's 64 rows lie in one segment and go through that network.
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer_many(const _Float16* __restrict__ x0, uint32_t n,
                                                                     const uint32_t* __restrict__ off, uint32_t nseg,
                                                                     const CacheRefD* __restrict__ refs,
                                                                     float4* __restrict__ y) {
    n = min(n, off[nseg]);
    if (blockIdx.x * kInferWaves * 64 >= n) return;
    const uint32_t base = (blockIdx.x * kInferWaves + (threadIdx.x >> 6)) * 64;
    // a wave past the rows runs on zero rows through network 0 and stores nothing
    const uint32_t k = base < n ? segment_of(off, nseg, base) : 0u;
    infer_block(x0, n, base, refs[k].mlp, y);
}

// dX tile (64 samples x 64 inputs) = dZ (64 x K) * W (K x 64), via W^T rows.
// Optionally multiplied by the ReLU mask of `act_mask` (> 0) and stored as fp16.
template <int KB, int NT>
__device__ __forceinline__ void back_dx(const _Float16* g_in, const _Float16* __restrict__ WT, int wt_row,
                                        f32x4_t acc[4][NT], int lane) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        half8_t b[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) b[n] = w_frag(WT, n, kb, lane, wt_row);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            half8_t a = a_frag(g_in, m, kb, lane);
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = mfma16(a, b[n], acc[m][n]);
        }
    }
}

// dW (MT*16 outputs x 64 inputs) = dZ^T (outputs x 64 samples) * X (64 samples x 64 inputs).
template <int MT>
__device__ __forceinline__ void back_dw(const _Float16* g, const _Float16* x, float* __restrict__ out, int lane) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        f32x4_t acc[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[n] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            half8_t a;
            const int o = m * 16 + (lane & 15), q0 = kb * 32 + 8 * (lane >> 4);
#pragma unroll
            for (int e = 0; e < 8; ++e) a[e] = g[(q0 + e) * kRow + o];
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                half8_t b;
                const int i = n * 16 + (lane & 15);
#pragma unroll
                for (int e = 0; e < 8; ++e) b[e] = x[(q0 + e) * kRow + i];
                acc[n] = mfma16(a, b, acc[n]);
            }
        }
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int o = m * 16 + 4 * (lane >> 4) + e, i = n * 16 + (lane & 15);
                out[o * kWidth + i] = acc[n][e];
            }
    }
}

// Training step, one wave per 64 samples: forward with stored activations,
// relative-L2 loss, backward to the weights (partials per wave) and to the 32
// grid features of each sample.
// The batch holds min(n, *n_dev) samples (n alone without n_dev); the loss is
// the mean over them of the three channels' relative squared errors, scaled by
// kLossScale for the FP16 backward.
__global__ __launch_bounds__(64) void k_mlp_train(const _Float16* __restrict__ x0, const float* __restrict__ samples,
                                                  uint32_t n, const uint32_t* __restrict__ n_dev,
                                                  const _Float16* __restrict__ W, const _Float16* __restrict__ WT,
                                                  float* __restrict__ partial, float* __restrict__ dx0,
                                                  float* __restrict__ loss_part, uint32_t hashed_levels,
                                                  AdamDev* __restrict__ ad) {
    __shared__ __attribute__((aligned(16))) _Float16 act[5][64 * kRow];
    __shared__ __attribute__((aligned(16))) _Float16 gb[2][64 * kRow];
    const int lane = threadIdx.x;
    const uint32_t base = blockIdx.x * 64;
    if (n_dev) n = min(n, *n_dev);
    if (base >= n) return;  // the whole wave is past the batch
    const float grad_scale = kLossScale / (3.0f * (float)n);
    load_x0_tile(x0, base, n, act[0], lane);
    __syncthreads();
    for (int l = 0; l < kHidden; ++l) {
        layer_hidden(act[l], W + l * kLayerParams, act[l + 1], lane);
        __syncthreads();
    }
    f32x4_t yo[4];
    layer_out(act[4], W + 4 * kLayerParams, yo, lane);
    // loss and dZ4 (scaled) into gb[0]: columns 0..31 (3 used)
    const int c = lane & 15;
    float lsum = 0.0f;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int row = m * 16 + 4 * (lane >> 4) + e;
            uint32_t q = base + row;
            float d = 0.0f;
            if (c < kOut && q < n) {
                float yv = yo[m][e];
                float t = samples[(size_t)kSampleFloats * q + 13 + c];
                float den = yv * yv + 0.01f;
                float diff = yv - t;
                lsum += diff * diff / den;
                d = 2.0f * diff / den * grad_scale;
            }
            gb[0][row * kRow + c] = (_Float16)d;
            gb[0][row * kRow + 16 + c] = (_Float16)0.0f;
        }
    for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) loss_part[blockIdx.x] = lsum;
    __syncthreads();
    float* P = partial + (size_t)blockIdx.x * kMlpParams;
    // layer 4: dW4 = dZ4^T X4 (16 x 64); dX4 = dZ4 * W4 (K = 32 padded)
    back_dw<1>(gb[0], act[4], P + 4 * kLayerParams, lane);
    int gi = 0;
    {
        f32x4_t acc[4][4];
        back_dx<1, 4>(gb[0], WT + 4 * kLayerParams, kWTRow4, acc, lane);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    int row = m * 16 + 4 * (lane >> 4) + e, col = nn * 16 + (lane & 15);
                    float v = act[4][row * kRow + col] > (_Float16)0.0f ? acc[m][nn][e] : 0.0f;
                    gb[1][row * kRow + col] = (_Float16)v;
                }
        gi = 1;
    }
    __syncthreads();
    for (int l = kHidden - 1; l >= 0; --l) {
        // gb[gi] holds dZ_l (64 x 64); dW_l = dZ_l^T X_l
        back_dw<4>(gb[gi], act[l], P + l * kLayerParams, lane);
        if (l > 0) {
            f32x4_t acc[4][4];
            back_dx<2, 4>(gb[gi], WT + l * kLayerParams, kWidth, acc, lane);
            __syncthreads();
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int nn = 0; nn < 4; ++nn)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        int row = m * 16 + 4 * (lane >> 4) + e, col = nn * 16 + (lane & 15);
                        float v = act[l][row * kRow + col] > (_Float16)0.0f ? acc[m][nn][e] : 0.0f;
                        gb[gi ^ 1][row * kRow + col] = (_Float16)v;
                    }
            gi ^= 1;
            __syncthreads();
        } else {
            // dX0 for the 32 grid features
            float gmax = 0.0f;
            f32x4_t acc[4][2];
            back_dx<2, 2>(gb[gi], WT + 0 * kLayerParams, kWidth, acc, lane);
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int nn = 0; nn < 2; ++nn)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        int row = m * 16 + 4 * (lane >> 4) + e, col = nn * 16 + (lane & 15);
                        uint32_t q = base + row;
                        if (q < n) {
                            dx0[(size_t)32 * q + col] = acc[m][nn][e];
                            if ((hashed_levels >> (col >> 1)) & 1u) gmax = fmaxf(gmax, fabsf(acc[m][nn][e]));
                        }
                    }
            // largest |dL/dfeature| on the hashed levels: the fixed point of their gradient words
            for (int off = 32; off >= 1; off >>= 1) gmax = fmaxf(gmax, __shfl_xor(gmax, off, 64));
            if (lane == 0 && gmax > 0.0f) atomicMax(&ad->gmax, __float_as_uint(gmax));
        }
    }
}

// Gradient privatisation of the coarsest levels: their few entries (17^3,
// 23^3, 31^3 dense cells) take most same-address atomics, so blocks spread
// them over kPrivCopies copies (copy = block index mod kPrivCopies) that a
// second pass sums back in a fixed order.  Integer sums: the result does not
// depend on the split.
constexpr int kPrivLevels = 3;
constexpr int kPrivCopies = 8;
__host__ __device__ constexpr uint32_t priv_cells(int l) {
    return l == 0 ? 17u * 17u * 17u : l == 1 ? 23u * 23u * 23u : 31u * 31u * 31u;
}
__host__ __device__ constexpr uint32_t priv_base(int l) {
    return l == 0 ? 0u : l == 1 ? priv_cells(0) : priv_cells(0) + priv_cells(1);
}
constexpr uint32_t kPrivEntries = priv_base(kPrivLevels - 1) + priv_cells(kPrivLevels - 1);  // cells per copy

// Scatter dL/dfeature to the 8 corners of each level (64-bit fixed point, so
// the sums do not depend on the order of the atomics).  One thread per sample
// walks the levels in turn: at any moment the threads of the grid spread
// their atomics over several levels.
__global__ void k_grid_bwd(const float* __restrict__ samples, const float* __restrict__ dx0, uint32_t n,
                           const uint32_t* __restrict__ n_dev, GridGeom G, unsigned long long* __restrict__ grad,
                           unsigned long long* __restrict__ priv, const AdamDev* __restrict__ ad) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = min(n, *n_dev);
    if (i >= n) return;
    unsigned long long* mine = priv + (size_t)(blockIdx.x % kPrivCopies) * kPrivEntries * 2;
    const float ps = packed_scale(ad->gmax);
    float p[3];
    norm_pos(samples + (size_t)kSampleFloats * i, G, p);
    for (int l = 0; l < kGridLevels; ++l) {
        const float g0 = dx0[(size_t)32 * i + 2 * l], g1 = dx0[(size_t)32 * i + 2 * l + 1];
        if (g0 == 0.0f && g1 == 0.0f) continue;
        const Cell c = cell_of(p, (float)G.res[l]);
        if (G.dense[l]) {
            // dense levels: many samples share an entry; one 64-bit word per feature
            for (int k = 0; k < 8; ++k) {
                float w = corner_weight(c, k);
                const uint32_t idx = corner_index(c, k, G, l);
                unsigned long long* dst = l < kPrivLevels ? mine + (size_t)(priv_base(l) + idx) * 2
                                                          : grad + ((size_t)l * kGridT + idx) * 2;
                long long v0 = (long long)llrint((double)(w * g0) * kFixScale);
                long long v1 = (long long)llrint((double)(w * g1) * kFixScale);
                if (v0) atomicAdd(&dst[0], (unsigned long long)v0);
                if (v1) atomicAdd(&dst[1], (unsigned long long)v1);
            }
        } else {
            for (int k = 0; k < 8; ++k) {
                float w = corner_weight(c, k);
                const uint32_t idx = corner_index(c, k, G, l);
                const long long h = llrintf(w * g0 * ps), lo = llrintf(w * g1 * ps);
                const long long v = h * 4294967296ll + lo;
                if (v) atomicAdd(&grad[((size_t)l * kGridT + idx) * 2], (unsigned long long)v);
            }
        }
    }
}

// Folds the private copies into the gradient (copies in a fixed order) and clears them.
__global__ void k_grid_priv_fold(unsigned long long* __restrict__ priv, unsigned long long* __restrict__ grad) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;  // entry * 2 + feature
    if (e >= kPrivEntries * 2) return;
    const uint32_t cell = e >> 1, f = e & 1u;
    const int l = cell < priv_base(1) ? 0 : cell < priv_base(2) ? 1 : 2;
    unsigned long long s = 0;
#pragma unroll
    for (int c = 0; c < kPrivCopies; ++c) {
        unsigned long long* q = priv + (size_t)c * kPrivEntries * 2 + e;
        s += *q;
        *q = 0ull;
    }
    const size_t g = ((size_t)l * kGridT + (cell - priv_base(l))) * 2 + f;
    grad[g] += s;
}

struct AdamP {
    float b1, b2, eps, inv_scale;
    int apply;
};

// Split of the per-wave gradient partials of n samples into reduction groups.
constexpr uint32_t kRedGroups = 64;
__host__ __device__ __forceinline__ void red_shape(uint32_t n, uint32_t& chunks, uint32_t& groups, uint32_t& per) {
    chunks = (n + 63) / 64;
    groups = chunks < kRedGroups ? chunks : kRedGroups;
    per = groups ? (chunks + groups - 1) / groups : 0u;
}

// Deterministic two-stage sum of the per-wave gradient partials: group y of
// the grid sums chunks [y * per, (y + 1) * per) in order, then k_mlp_adam sums
// the groups in order.
__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t n, const uint32_t* __restrict__ n_dev,
                             float* __restrict__ red) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    if (n_dev) n = min(n, *n_dev);
    uint32_t nchunks, groups, per;
    red_shape(n, nchunks, groups, per);
    if (blockIdx.y >= groups) return;
    uint32_t c0 = blockIdx.y * per, c1 = min(nchunks, c0 + per);
    float g = 0.0f;
    for (uint32_t c = c0; c < c1; ++c) g += partial[(size_t)c * kMlpParams + p];
    red[(size_t)blockIdx.y * kMlpParams + p] = g;
}

// Adam on the reduced gradient.  Keeps fp32 masters, the fp16 weights and the
// fp16 transposed weights in step.
__global__ void k_mlp_adam(const float* __restrict__ red, uint32_t n, const uint32_t* __restrict__ n_dev,
                           float* __restrict__ w32, float* __restrict__ m, float* __restrict__ v,
                           _Float16* __restrict__ w16, _Float16* __restrict__ wt16, float* __restrict__ gout,
                           float* __restrict__ e32, _Float16* __restrict__ e16, AdamP A,
                           const AdamDev* __restrict__ ad) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams || !ad->live) return;
    if (n_dev) n = min(n, *n_dev);
    uint32_t nchunks, ngroups, per;
    red_shape(n, nchunks, ngroups, per);
    float g = 0.0f;
    for (uint32_t c = 0; c < ngroups; ++c) g += red[(size_t)c * kMlpParams + p];
    g *= A.inv_scale;
    gout[p] = g;
    if (!A.apply) return;
    const float lr_t = ad->lr_t, ema = ad->ema;
    float mm = A.b1 * m[p] + (1.0f - A.b1) * g;
    float vv = A.b2 * v[p] + (1.0f - A.b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float w = w32[p] - lr_t * mm / (sqrtf(vv) + A.eps);
    w32[p] = w;
    w16[p] = (_Float16)w;
    const float e = ema * e32[p] + (1.0f - ema) * w;
    e32[p] = e;
    e16[p] = (_Float16)e;
    int layer = p / kLayerParams, r = p % kLayerParams;
    int o = r / kWidth, i = r % kWidth;
    if (layer < kHidden)
        wt16[layer * kLayerParams + i * kWidth + o] = (_Float16)w;
    else if (o < kOutPad)
        wt16[kHidden * kLayerParams + i * kWTRow4 + o] = (_Float16)w;
}

// Adam on the grid, one thread per entry (both features): decodes the
// entry's fixed-point gradient (a 64-bit word per feature on dense levels, one
// packed word on hashed levels), clears it, and steps the entry.  Entries
// without gradient this step are left as they are (lazy / sparse Adam: their
// moments are not decayed), so a step reads 8-16 B for most untouched entries.
__global__ void k_grid_adam(unsigned long long* __restrict__ grad, float2* __restrict__ w32, float2* __restrict__ m,
                            float2* __restrict__ v, __half2* __restrict__ w16, float2* __restrict__ gout,
                            float2* __restrict__ e32, __half2* __restrict__ e16, GridGeom G, AdamP A,
                            const AdamDev* __restrict__ ad) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)kGridLevels * kGridT || !ad->live) return;
    const uint32_t l = (uint32_t)(e >> kGridLog2T);
    float g0, g1;
    if (G.dense[l]) {
        // a dense level uses its first (res + 1)^3 slots only
        const uint32_t side = G.res[l] + 1u;
        if ((uint32_t)(e & (kGridT - 1u)) >= side * side * side) {
            if (gout) gout[e] = make_float2(0.0f, 0.0f);
            return;
        }
        const long long a = (long long)grad[2 * e], b = (long long)grad[2 * e + 1];
        if ((a | b) == 0 && !gout) return;
        if (a) grad[2 * e] = 0ull;
        if (b) grad[2 * e + 1] = 0ull;
        g0 = (float)((double)a / kFixScale);
        g1 = (float)((double)b / kFixScale);
    } else {
        const long long w = (long long)grad[2 * e];
        if (w == 0 && !gout) return;
        if (w) grad[2 * e] = 0ull;
        const long long lo = (long long)(int32_t)(uint32_t)((unsigned long long)w & 0xFFFFFFFFull);
        const long long hi = (w - lo) >> 32;
        const float inv = 1.0f / packed_scale(ad->gmax);  // a power of two: exact
        g0 = (float)hi * inv;
        g1 = (float)lo * inv;
    }
    g0 *= A.inv_scale;
    g1 *= A.inv_scale;
    if (gout) gout[e] = make_float2(g0, g1);
    if (!A.apply || (g0 == 0.0f && g1 == 0.0f)) return;
    const float lr_t = ad->lr_t, ema = ad->ema;
    float2 mm = m[e], vv = v[e], w = w32[e], ea = e32[e];
    mm.x = A.b1 * mm.x + (1.0f - A.b1) * g0;
    mm.y = A.b1 * mm.y + (1.0f - A.b1) * g1;
    vv.x = A.b2 * vv.x + (1.0f - A.b2) * g0 * g0;
    vv.y = A.b2 * vv.y + (1.0f - A.b2) * g1 * g1;
    m[e] = mm;
    v[e] = vv;
    w.x = w.x - lr_t * mm.x / (sqrtf(vv.x) + A.eps);
    w.y = w.y - lr_t * mm.y / (sqrtf(vv.y) + A.eps);
    w32[e] = w;
    w16[e] = __floats2half2_rn(w.x, w.y);
    ea.x = ema * ea.x + (1.0f - ema) * w.x;
    ea.y = ema * ea.y + (1.0f - ema) * w.y;
    e32[e] = ea;
    e16[e] = __floats2half2_rn(ea.x, ea.y);
}

__global__ void k_grid_init(float* __restrict__ w32, __half* __restrict__ w16, float* __restrict__ m,
                            float* __restrict__ v, unsigned long long* __restrict__ grad, size_t n, uint64_t seed) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    Rng r = rng_seed(seed, p);
    float w = (r.uniform() * 2.0f - 1.0f) * 1e-4f;
    w32[p] = w;
    w16[p] = __float2half(w);
    m[p] = 0.0f;
    v[p] = 0.0f;
    grad[p] = 0ull;
}

__global__ void k_sync_grid16(const float* __restrict__ w32, __half* __restrict__ w16, size_t n) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) w16[p] = __float2half(w32[p]);
}

// fp16 and transposed fp16 copies of the fp32 MLP masters (the padding columns
// of W4^T are cleared by the host once and never written).
__global__ void k_sync_mlp16(const float* __restrict__ w32, _Float16* __restrict__ w16, _Float16* __restrict__ wt16) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float w = w32[p];
    w16[p] = (_Float16)w;
    int layer = p / kLayerParams, r = p % kLayerParams;
    int o = r / kWidth, i = r % kWidth;
    if (layer < kHidden)
        wt16[layer * kLayerParams + i * kWidth + o] = (_Float16)w;
    else
        wt16[kHidden * kLayerParams + i * kWTRow4 + o] = (_Float16)w;
}

}  // namespace

#define HIPC(expr)                                                \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr);    \
    } while (0)

static void make_geom(GridGeom& G, const float lo[3], const float hi[3]) {
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) {
        G.lo[a] = lo[a];
        ext = std::max(ext, hi[a] - lo[a]);
    }
    G.inv_ext = ext > 0.0f ? 1.0f / ext : 1.0f;
    const double b = std::exp((std::log(2048.0) - std::log(16.0)) / (kGridLevels - 1));
    for (int l = 0; l < kGridLevels; ++l) {
        uint32_t r = (uint32_t)std::floor(16.0 * std::pow(b, (double)l) + 1e-6);
        G.res[l] = r;
        uint64_t cells = (uint64_t)(r + 1) * (r + 1) * (r + 1);
        G.dense[l] = cells <= kGridT ? 1u : 0u;
    }
}

// Xavier-uniform MLP initialisation from a seeded PCG32 stream (host side).
static void init_mlp(std::vector<float>& w, uint32_t seed) {
    w.assign(kMlpParams, 0.0f);
    uint64_t state = 0, inc = ((uint64_t)seed << 1) | 1u;
    auto next = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    };
    next();
    state += 0x5EEDull;
    next();
    // Glorot-uniform hidden layers; the output layer 100x smaller, so a fresh
    // cache answers ~0 (the missing indirect light) rather than O(1) noise
    for (int l = 0; l <= kHidden; ++l) {
        int outs = l < kHidden ? kWidth : kOut;
        float lim = std::sqrt(6.0f / (float)(kWidth + (l < kHidden ? kWidth : kOut))) * (l < kHidden ? 1.0f : 0.01f);
        for (int o = 0; o < outs; ++o)
            for (int i = 0; i < kWidth; ++i)
                w[l * kLayerParams + o * kWidth + i] = ((float)(next() >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * lim;
    }
}

// weight averages := weights
static hipError_t sync_ema(CacheState& C, hipStream_t st) {
    hipError_t e;
    if ((e = hipMemcpyAsync(C.grid_e32.ptr, C.grid32.ptr, kGridParams * 4, hipMemcpyDeviceToDevice, st))) return e;
    if ((e = hipMemcpyAsync(C.grid_e16.ptr, C.grid16.ptr, kGridParams * 2, hipMemcpyDeviceToDevice, st))) return e;
    if ((e = hipMemcpyAsync(C.mlp_e32.ptr, C.mlp32.ptr, kMlpParams * 4, hipMemcpyDeviceToDevice, st))) return e;
    return hipMemcpyAsync(C.mlp_e16.ptr, C.mlp16.ptr, kMlpParams * 2, hipMemcpyDeviceToDevice, st);
}

snkpt_status cache_reset(snkpt_ctx* ctx, uint32_t seed, const float* lo, const float* hi, hipStream_t st) {
    if (!ctx->cache) ctx->cache = new CacheState();
    CacheState& C = *ctx->cache;
    float l3[3], h3[3];
    if (lo && hi) {
        std::memcpy(l3, lo, 12);
        std::memcpy(h3, hi, 12);
    } else {
        if (!ctx->accel.built) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_reset: no bounds and no scene");
        std::memcpy(l3, ctx->accel.lo, 12);
        std::memcpy(h3, ctx->accel.hi, 12);
    }
    make_geom(C.geom, l3, h3);
    for (int l = 0; l < kPrivLevels; ++l) {
        const uint64_t cells = (uint64_t)(C.geom.res[l] + 1) * (C.geom.res[l] + 1) * (C.geom.res[l] + 1);
        if (!C.geom.dense[l] || cells != priv_cells(l))
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_reset: unexpected coarse-level resolution");
    }
    HIPC(C.grid32.reserve(kGridParams * 4));
    HIPC(C.grid16.reserve(kGridParams * 2));
    HIPC(C.grid_m.reserve(kGridParams * 4));
    HIPC(C.grid_v.reserve(kGridParams * 4));
    HIPC(C.grid_grad.reserve(kGridParams * 8));
    HIPC(C.mlp32.reserve(kMlpParams * 4));
    HIPC(C.mlp16.reserve(kMlpParams * 2));
    HIPC(C.mlp16t.reserve(kMlpTParams * 2));
    HIPC(C.mlp_m.reserve(kMlpParams * 4));
    HIPC(C.mlp_v.reserve(kMlpParams * 4));
    HIPC(C.mlp_gradf.reserve(kMlpParams * 4));
    HIPC(C.mlp_red.reserve((size_t)kRedGroups * kMlpParams * 4));
    HIPC(C.grid_priv.reserve((size_t)kPrivCopies * kPrivEntries * 2 * 8));
    HIPC(hipMemsetAsync(C.grid_priv.ptr, 0, (size_t)kPrivCopies * kPrivEntries * 2 * 8, st));
    k_grid_init<<<(uint32_t)((kGridParams + 255) / 256), 256, 0, st>>>(
        C.grid32.as<float>(), C.grid16.as<__half>(), C.grid_m.as<float>(), C.grid_v.as<float>(),
        C.grid_grad.as<unsigned long long>(), kGridParams, 0x6772696400000000ull ^ seed);
    std::vector<float> w;
    init_mlp(w, seed);
    std::vector<float> zeros(kMlpParams, 0.0f);
    HIPC(hipMemcpyAsync(C.mlp32.ptr, w.data(), kMlpParams * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(C.mlp_m.ptr, zeros.data(), kMlpParams * 4, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(C.mlp_v.ptr, zeros.data(), kMlpParams * 4, hipMemcpyHostToDevice, st));
    {
        std::vector<_Float16> zt(kMlpTParams, (_Float16)0.0f);
        HIPC(hipMemcpyAsync(C.mlp16t.ptr, zt.data(), kMlpTParams * 2, hipMemcpyHostToDevice, st));
        k_sync_mlp16<<<(kMlpParams + 255) / 256, 256, 0, st>>>(C.mlp32.as<float>(), C.mlp16.as<_Float16>(),
                                                               C.mlp16t.as<_Float16>());
    }
    HIPC(C.grid_e32.reserve(kGridParams * 4));
    HIPC(C.grid_e16.reserve(kGridParams * 2));
    HIPC(C.mlp_e32.reserve(kMlpParams * 4));
    HIPC(C.mlp_e16.reserve(kMlpParams * 2));
    HIPC(sync_ema(C, st));
    HIPC(C.adam.reserve(sizeof(AdamDev)));
    k_adam_reset<<<1, 64, 0, st>>>(C.adam.as<AdamDev>());
    HIPC(hipStreamSynchronize(st));
    C.ready = true;
    return SNKPT_OK;
}

// Work buffers for n samples: encodings and outputs for inference; with `train`
// also the backward buffers (per-wave weight-gradient partials: 68 KiB per 64
// samples, so they are sized by the training batch only).
static snkpt_status ensure_work(snkpt_ctx* ctx, CacheState& C, uint32_t n, bool train) {
    const uint32_t cap = ((n + 63) / 64) * 64;
    if (cap > C.cap) {
        HIPC(C.x0.reserve((size_t)cap * 128));
        HIPC(C.y.reserve((size_t)cap * 16));
        C.cap = cap;
    }
    if (train && cap > C.tcap) {
        HIPC(C.dx0.reserve((size_t)cap * 128));
        HIPC(C.partial.reserve((size_t)(cap / 64) * kMlpParams * 4));
        HIPC(C.loss_part.reserve((size_t)(cap / 64) * 4));
        C.tcap = cap;
    }
    return SNKPT_OK;
}

snkpt_status cache_infer(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out, hipStream_t st) {
    CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_infer: call snkpt_cache_reset first");
    CacheState& C = *Cp;
    if (n == 0) return SNKPT_OK;
    snkpt_status e = ensure_work(ctx, C, n, false);
    if (e != SNKPT_OK) return e;
    k_encode<<<enc_blocks(n), kEncBlock, 0, st>>>(samples, n, nullptr, C.geom, C.grid16.as<__half2>(),
                                               C.x0.as<_Float16>());
    const uint32_t waves = (n + 63) / 64;
    k_mlp_infer<<<(waves + kInferWaves - 1) / kInferWaves, 64 * kInferWaves, 0, st>>>(
        C.x0.as<_Float16>(), n, nullptr, C.mlp16.as<_Float16>(), (float4*)out);
    C.last_n = n;
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(ctx, le, "cache_infer launch");
    return SNKPT_OK;
}

// ---------------------------------------------- queries without allocation ----
// Workspace of cache_infer_noalloc: the encoded query rows (64 FP16 network
// inputs = 128 B per query) for at most kNoAllocQueries queries; 8 MiB held on
// every GPU the library runs on until the process ends.  Every call on a GPU
// shares that GPU's copy, so calls that overlap on one GPU (any context, any
// stream) race.  Nothing in it has to start at zero: k_encode writes every row
// below n before k_mlp_infer reads rows below n only.
static __device__ __align__(256) _Float16 scratch[(size_t)kNoAllocQueries * 64];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
_Float16* g_scratch[kMaxScratchDevices] = {};
}  // namespace

// Looks up this device's copy of `scratch` (the array's name is only a host
// stand-in); called by snkpt_create with `device` current.
snkpt_status cache_scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return SNKPT_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return SNKPT_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SNKPT_ERR_HIP;
    g_scratch[device] = (_Float16*)p;
    return SNKPT_OK;
}

// cache_infer with the encoded rows in `scratch`: no allocation, same kernels,
// bit-identical outputs.
snkpt_status cache_infer_noalloc(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out, hipStream_t st) {
    CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_infer_noalloc: call snkpt_cache_reset first");
    if (n > kNoAllocQueries)
        return fail(ctx, SNKPT_ERR_CAPACITY, "cache_infer_noalloc: more than 65,536 queries per call");
    if (n == 0) return SNKPT_OK;
    _Float16* x0 = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxScratchDevices) x0 = g_scratch[ctx->device];
    }
    if (!x0) return fail(ctx, SNKPT_ERR_HIP, "cache_infer_noalloc: workspace address not set up for this device");
    CacheState& C = *Cp;
    k_encode<<<enc_blocks(n), kEncBlock, 0, st>>>(samples, n, nullptr, C.geom, C.grid16.as<__half2>(), x0);
    const uint32_t waves = (n + 63) / 64;
    k_mlp_infer<<<(waves + kInferWaves - 1) / kInferWaves, 64 * kInferWaves, 0, st>>>(x0, n, nullptr,
                                                                                      C.mlp16.as<_Float16>(),
                                                                                      (float4*)out);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(ctx, le, "cache_infer_noalloc launch");
    return SNKPT_OK;
}

snkpt_status cache_train(snkpt_ctx* ctx, const float* samples, uint32_t n, float lr, int apply, float* loss_out,
                         hipStream_t st) {
    return cache_train_run(ctx, samples, n, nullptr, lr, apply, 0.0f, loss_out, st);
}

// One optimisation step on min(n, *n_dev) samples (n without n_dev): launch
// sizes follow n, the kernels read the batch size on the device, so a frame's
// step needs no readback and can be captured.  `ema` is the decay of the
// weight averages the render queries read (0 sets them to the new weights).
snkpt_status cache_train_run(snkpt_ctx* ctx, const float* samples, uint32_t n, const uint32_t* n_dev, float lr,
                             int apply, float ema, float* loss_out, hipStream_t st) {
    CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_train: call snkpt_cache_reset first");
    CacheState& C = *Cp;
    if (n == 0) return SNKPT_OK;
    if (loss_out && n_dev) return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_train: the loss needs a host batch size");
    snkpt_status e = ensure_work(ctx, C, n, true);
    if (e != SNKPT_OK) return e;
    const float b1 = 0.9f, b2 = 0.99f;
    AdamDev* ad = C.adam.as<AdamDev>();
    k_adam_advance<<<1, 64, 0, st>>>(ad, n, n_dev, lr, ema, b1, b2, apply);
    const uint32_t chunks = (n + 63) / 64;
    k_encode<<<enc_blocks(n), kEncBlock, 0, st>>>(samples, n, n_dev, C.geom, C.grid16.as<__half2>(),
                                               C.x0.as<_Float16>());
    uint32_t hashed = 0;
    for (int l = 0; l < kGridLevels; ++l)
        if (!C.geom.dense[l]) hashed |= 1u << l;
    k_mlp_train<<<chunks, 64, 0, st>>>(C.x0.as<_Float16>(), samples, n, n_dev, C.mlp16.as<_Float16>(),
                                       C.mlp16t.as<_Float16>(), C.partial.as<float>(), C.dx0.as<float>(),
                                       C.loss_part.as<float>(), hashed, ad);
    k_grid_bwd<<<grid_for(n, 128), 128, 0, st>>>(samples, C.dx0.as<float>(), n, n_dev, C.geom,
                                                 C.grid_grad.as<unsigned long long>(),
                                                 C.grid_priv.as<unsigned long long>(), ad);
    k_grid_priv_fold<<<grid_for(kPrivEntries * 2, 256), 256, 0, st>>>(C.grid_priv.as<unsigned long long>(),
                                                                      C.grid_grad.as<unsigned long long>());
    AdamP A;
    A.b1 = b1;
    A.b2 = b2;
    A.eps = 1e-15f;
    A.inv_scale = 1.0f / kLossScale;
    A.apply = apply;
    uint32_t rchunks, groups, per;
    red_shape(n, rchunks, groups, per);
    k_mlp_reduce<<<dim3((kMlpParams + 255) / 256, groups), 256, 0, st>>>(C.partial.as<float>(), n, n_dev,
                                                                         C.mlp_red.as<float>());
    k_mlp_adam<<<(kMlpParams + 255) / 256, 256, 0, st>>>(C.mlp_red.as<float>(), n, n_dev, C.mlp32.as<float>(),
                                                        C.mlp_m.as<float>(), C.mlp_v.as<float>(),
                                                        C.mlp16.as<_Float16>(), C.mlp16t.as<_Float16>(),
                                                        C.mlp_gradf.as<float>(), C.mlp_e32.as<float>(),
                                                        C.mlp_e16.as<_Float16>(), A, ad);
    // grid: the fixed-point gradient is consumed (and cleared) every call; without
    // Adam it is kept as floats so it can be read back
    float2* gout = nullptr;
    if (!apply) {
        HIPC(C.grid_gradf.reserve(kGridParams * 4));
        gout = C.grid_gradf.as<float2>();
    }
    const size_t entries = (size_t)kGridLevels * kGridT;
    k_grid_adam<<<(uint32_t)((entries + 255) / 256), 256, 0, st>>>(
        C.grid_grad.as<unsigned long long>(), C.grid32.as<float2>(), C.grid_m.as<float2>(), C.grid_v.as<float2>(),
        C.grid16.as<__half2>(), gout, C.grid_e32.as<float2>(), C.grid_e16.as<__half2>(), C.geom, A, ad);
    C.last_n = n;
    if (loss_out) {
        std::vector<float> lp(chunks);
        HIPC(hipMemcpyAsync(lp.data(), C.loss_part.ptr, chunks * 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        double s = 0;
        for (float v : lp) s += v;
        *loss_out = (float)(s / (3.0 * n));
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(ctx, le, "cache_train launch");
    return SNKPT_OK;
}

snkpt_status cache_params(snkpt_ctx* ctx, uint32_t which, void* dst, size_t bytes, size_t* out_bytes, hipStream_t st) {
    CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_params: call snkpt_cache_reset first");
    CacheState& C = *Cp;
    size_t sz = 0;
    switch (which) {
        case 0: sz = kMlpParams * 4; break;
        case 1: sz = kGridParams * 4; break;
        case 2: sz = kMlpParams * 4; break;
        case 3: sz = kGridParams * 4; break;
        case 4: sz = (size_t)C.last_n * 128; break;
        case 5:
        case 6:
        case 7: {
            // query / training records of the last cache-mode pass; counts live on the device
            // (queries: counts[0]; training samples: tcounts[1])
            uint32_t cnt[2] = {0, 0};
            const DeviceBuffer& cb = which == 7 ? C.tcounts : C.counts;
            if (cb.ptr) {
                HIPC(hipMemcpyAsync(cnt, cb.ptr, 8, hipMemcpyDeviceToHost, st));
                HIPC(hipStreamSynchronize(st));
            }
            sz = which == 7 ? (size_t)cnt[1] * 64 : (size_t)cnt[0] * (which == 5 ? 64 : 16);
            break;
        }
        default: return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_params: unknown array");
    }
    if (out_bytes) *out_bytes = sz;
    if (!dst) return SNKPT_OK;
    if (bytes < sz) return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_params: destination too small");
    if (which == 3 && !C.grid_gradf.ptr) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_params: no gradient-only step yet");
    const void* src = which == 0   ? C.mlp32.ptr
                      : which == 1 ? C.grid32.ptr
                      : which == 2 ? C.mlp_gradf.ptr
                      : which == 3 ? C.grid_gradf.ptr
                      : which == 4 ? C.x0.ptr
                      : which == 5 ? C.q_samples.ptr
                      : which == 6 ? C.q_meta.ptr
                                   : C.t_samples.ptr;
    if (sz == 0) return SNKPT_OK;
    HIPC(hipMemcpyAsync(dst, src, sz, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return SNKPT_OK;
}

snkpt_status cache_set_params(snkpt_ctx* ctx, uint32_t which, const void* src, size_t bytes, hipStream_t st) {
    CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_set_params: call snkpt_cache_reset first");
    CacheState& C = *Cp;
    if (which == 0) {
        if (bytes != kMlpParams * 4) return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_set_params: MLP size");
        HIPC(hipMemcpyAsync(C.mlp32.ptr, src, bytes, hipMemcpyHostToDevice, st));
        k_sync_mlp16<<<(kMlpParams + 255) / 256, 256, 0, st>>>(C.mlp32.as<float>(), C.mlp16.as<_Float16>(),
                                                               C.mlp16t.as<_Float16>());
    } else if (which == 1) {
        if (bytes != kGridParams * 4) return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_set_params: grid size");
        HIPC(hipMemcpyAsync(C.grid32.ptr, src, bytes, hipMemcpyHostToDevice, st));
        k_sync_grid16<<<(uint32_t)((kGridParams + 255) / 256), 256, 0, st>>>(C.grid32.as<float>(), C.grid16.as<__half>(),
                                                                              kGridParams);
    } else {
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "cache_set_params: only 0 (MLP) and 1 (grid) can be set");
    }
    HIPC(sync_ema(C, st));
    HIPC(hipStreamSynchronize(st));
    return SNKPT_OK;
}

namespace {

// rad[path] += throughput * max(cache output, 0) for every query of this pass.
__global__ void k_cache_apply(const float4* __restrict__ meta, const float4* __restrict__ y,
                              const uint32_t* __restrict__ count, float4* __restrict__ rad,
                              unsigned long long* __restrict__ stat) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *stat += *count;
    if (i >= *count) return;
    float4 m = meta[i], v = y[i];
    uint32_t path = __float_as_uint(m.w);
    float4 r = rad[path];
    rad[path] = make_float4(r.x + m.x * fmaxf(v.x, 0.0f), r.y + m.y * fmaxf(v.y, 0.0f), r.z + m.z * fmaxf(v.z, 0.0f),
                            r.w);
}

// training counters: samples emitted, recorded vertices
__global__ void k_train_clear(uint32_t* c) {
    if (threadIdx.x == 0) {
        c[1] = 0u;
        c[2] = 0u;
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

__global__ void k_fill4(float4* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Training samples from the recorded vertices of the training paths.  Vertex
// v of a path with throughput beta_v and radiance snapshot S_v (radiance of
// the path before v's own contributions) gets the target
//     (L - S_v) / beta_v                 = an unbiased estimate of L_o(x_v, wo_v),
// or, with `indirect`, (L - S_{v+1}) / beta_v (0 when the path has no
// recorded vertex v + 1): v's NEE and the MIS-weighted emission its
// continuation hits are direct light, everything later is indirect.
// Every vertex is a candidate and is kept with one probability p for the
// whole frame (p = min(1, budget / all recorded vertices); the samples past
// the budget, the last few in path order, are dropped), decided by a hash
// of (seed, path, vertex): selection never depends on how a path continues,
// which would bias the targets (picking one vertex per path favours short
// paths).  Samples are emitted in (path, vertex) order through a scan, so the
// batch, and the gradient sums over it, do not depend on timing.
__device__ __forceinline__ bool train_sample(const float* __restrict__ rec, uint32_t max_rec, uint32_t j, uint32_t k,
                                             uint32_t v, float4 L, uint32_t indirect, float smp[16]) {
    const float* e = rec + ((size_t)j * max_rec + v) * kRecFloats;
    const float b[3] = {e[16], e[17], e[18]};
    float lt[3] = {L.x - e[13], L.y - e[14], L.z - e[15]};
    if (indirect) {
        if (v + 1 < k) {
            const float* e1 = e + kRecFloats;
            lt[0] = L.x - e1[13], lt[1] = L.y - e1[14], lt[2] = L.z - e1[15];
        } else {
            lt[0] = lt[1] = lt[2] = 0.0f;
        }
    }
    bool ok = b[0] > 1e-6f && b[1] > 1e-6f && b[2] > 1e-6f;
#pragma unroll
    for (int c = 0; c < 13; ++c) smp[c] = e[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float t = fmaxf(lt[c], 0.0f) / fmaxf(b[c], 1e-6f);
        smp[13 + c] = fminf(t, 1000.0f);
        ok = ok && __builtin_isfinite(t);
    }
    return ok;
}

__device__ __forceinline__ bool keep_vertex(uint64_t seed, uint32_t j, uint32_t v, uint64_t thr) {
    uint64_t x = seed ^ (((uint64_t)j << 8) | v) * 0x9E3779B97F4A7C15ull;
    x ^= x >> 31;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 29;
    return (x >> 32) < thr;
}

// keep threshold in 2^-32 units from the frame's recorded-vertex total
__device__ __forceinline__ uint64_t keep_threshold(const uint32_t* __restrict__ cnt, uint32_t budget) {
    uint32_t total = cnt[2];
    if (total <= budget) return 1ull << 32;
    return (uint64_t)(((double)budget / (double)total) * 4294967296.0);
}

// vsel: the one vertex index that is a candidate (the query vertex), or ~0u for all
__global__ void k_train_total(const uint32_t* __restrict__ nrec, uint32_t n, uint32_t vsel,
                              uint32_t* __restrict__ cnt) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k = j < n ? nrec[j] : 0u;
    if (vsel != ~0u) k = k > vsel ? 1u : 0u;
    // wave sum, one atomic per wave
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o, 64);
    if ((threadIdx.x & 63) == 0 && k) atomicAdd(&cnt[2], k);
}

__global__ void k_train_count(const float* __restrict__ rec, const uint32_t* __restrict__ nrec, uint32_t max_rec,
                              const float4* __restrict__ rad, uint32_t n, uint64_t seed,
                              const TrainSeedD* __restrict__ tseed, uint32_t indirect, uint32_t vsel,
                              const uint32_t* __restrict__ cnt, uint32_t budget, uint32_t* __restrict__ count) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (tseed) seed = tseed->kseed;
    const uint64_t thr = keep_threshold(cnt, budget);
    const uint32_t k = nrec[j];
    const float4 L = rad[j];
    uint32_t c = 0;
    float smp[16];
    for (uint32_t v = 0; v < k; ++v)
        if ((vsel == ~0u || v == vsel) && keep_vertex(seed, j, v, thr) &&
            train_sample(rec, max_rec, j, k, v, L, indirect, smp))
            ++c;
    count[j] = c;
}

__global__ void k_train_emit(const float* __restrict__ rec, const uint32_t* __restrict__ nrec, uint32_t max_rec,
                             const float4* __restrict__ rad, uint32_t n, uint64_t seed,
                             const TrainSeedD* __restrict__ tseed, uint32_t indirect, uint32_t vsel,
                             const uint32_t* __restrict__ cnt, uint32_t budget, const uint32_t* __restrict__ off,
                             float* __restrict__ out, unsigned long long* __restrict__ stat) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) atomicAdd(stat, (unsigned long long)min(cnt[1], budget));
    if (j >= n) return;
    if (tseed) seed = tseed->kseed;
    const uint64_t thr = keep_threshold(cnt, budget);
    const uint32_t k = nrec[j];
    const float4 L = rad[j];
    uint32_t slot = off[j];
    float smp[16];
    for (uint32_t v = 0; v < k && slot < budget; ++v) {
        if ((vsel != ~0u && v != vsel) || !keep_vertex(seed, j, v, thr) ||
            !train_sample(rec, max_rec, j, k, v, L, indirect, smp))
            continue;
        float4* o = (float4*)(out + (size_t)slot * 16);
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] = make_float4(smp[4 * c], smp[4 * c + 1], smp[4 * c + 2], smp[4 * c + 3]);
        ++slot;
    }
}

}  // namespace

snkpt_status cache_ensure(snkpt_ctx* ctx, const snkpt_settings* s, hipStream_t st) {
    if (!ctx->cache || !ctx->cache->ready) return cache_reset(ctx, s->seed, nullptr, nullptr, st);
    return SNKPT_OK;
}

// Encoding + MLP on n queries (device memory) into y; `averaged`: the weight averages.
snkpt_status cache_query(snkpt_ctx* ctx, const float* samples, uint32_t n, float4* y, bool averaged,
                         hipStream_t st) {
    CacheState& C = *ctx->cache;
    if (n == 0) return SNKPT_OK;
    snkpt_status e = ensure_work(ctx, C, n, false);
    if (e != SNKPT_OK) return e;
    k_encode<<<enc_blocks(n), kEncBlock, 0, st>>>(samples, n, nullptr, C.geom,
                                                  (averaged ? C.grid_e16 : C.grid16).as<__half2>(),
                                                  C.x0.as<_Float16>());
    const uint32_t waves = (n + 63) / 64;
    k_mlp_infer<<<(waves + kInferWaves - 1) / kInferWaves, 64 * kInferWaves, 0, st>>>(
        C.x0.as<_Float16>(), n, nullptr, (averaged ? C.mlp_e16 : C.mlp16).as<_Float16>(), y);
    return SNKPT_OK;
}

__global__ void k_add_u64(unsigned long long* dst, const unsigned long long* src, uint32_t n) {
    if (threadIdx.x < n) dst[threadIdx.x] += src[threadIdx.x];
}
__global__ void k_clear_u64(unsigned long long* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Training half of a cache frame: n_paths full paths through a strided pixel
// subset of a w x h view of camera cd record their vertices; the thinned
// vertices become one optimisation step (Adam, weight averages of decay cache_ema).

// Buffers of a frame's training half for NT paths of `depth` vertices.
// train_meta: nrec | per-path sample counts | offsets (u32 each) | scan sums | radiance (float4)
static size_t train_u32_bytes(uint32_t NT) { return ((size_t)NT * 4 + 255) & ~(size_t)255; }
static size_t train_sums_bytes(uint32_t NT) { return (((size_t)(NT + 4095) / 4096) * 4 + 255) & ~(size_t)255; }
static snkpt_status train_reserve(snkpt_ctx* ctx, CacheState& C, uint32_t NT, uint32_t depth) {
    HIPC(C.tcounts.reserve(64));
    HIPC(C.tstats.reserve(64));
    if (NT == 0) return SNKPT_OK;
    HIPC(C.train_vtx.reserve((size_t)NT * depth * kRecFloats * 4));
    HIPC(C.train_meta.reserve(3 * train_u32_bytes(NT) + train_sums_bytes(NT) + (size_t)NT * 16));
    HIPC(C.t_samples.reserve((size_t)NT * depth * 64));
    return SNKPT_OK;
}

snkpt_status cache_prepare_train(snkpt_ctx* ctx, const snkpt_settings* s, uint32_t w, uint32_t h, uint32_t n_paths,
                                 hipStream_t st) {
    snkpt_status e = cache_ensure(ctx, s, st);
    if (e != SNKPT_OK) return e;
    CacheState& C = *ctx->cache;
    const uint32_t NT = std::min<uint32_t>(n_paths, w * h);
    if ((e = train_reserve(ctx, C, NT, s->max_depth ? s->max_depth : 1u)) != SNKPT_OK) return e;
    return NT ? ensure_work(ctx, C, NT, true) : SNKPT_OK;
}

snkpt_status cache_ref(snkpt_ctx* ctx, bool averaged, CacheRefD* out) {
    CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache_ref: no cache");
    out->geom = Cp->geom;
    out->grid = (averaged ? Cp->grid_e16 : Cp->grid16).as<__half2>();
    out->mlp = (averaged ? Cp->mlp_e16 : Cp->mlp16).as<_Float16>();
    return SNKPT_OK;
}

void cache_query_many(const CacheRefD* refs, const uint32_t* off, uint32_t nseg, const float* samples,
                      uint32_t rows_max, _Float16* x0, float4* y, hipStream_t st) {
    k_encode_many<<<enc_blocks(rows_max), kEncBlock, 0, st>>>(samples, rows_max, off, nseg, refs, x0);
    const uint32_t waves = (rows_max + 63) / 64;
    k_mlp_infer_many<<<(waves + kInferWaves - 1) / kInferWaves, 64 * kInferWaves, 0, st>>>(x0, rows_max, off, nseg,
                                                                                         refs, y);
}

// Training paths of a frame and their samples (no network involved, no host
// synchronisation): full-length unbiased paths through a strided pixel subset,
// recorded vertices thinned and emitted in path order into t_samples; the
// sample count lands in tcounts[1].  With `tseed` the paths' random streams and
// first pixel are read from the device (initstate only sets the stride's phase
// otherwise).
snkpt_status cache_train_paths(snkpt_ctx* ctx, const CamDev& cd, uint32_t w, uint32_t h, const snkpt_settings* s,
                               uint64_t initstate, uint32_t n_paths, unsigned long long* stats, hipStream_t st,
                               const TrainSeedD* tseed) {
    CacheState& C = *ctx->cache;
    const uint32_t n = w * h;
    const uint32_t depth = s->max_depth ? s->max_depth : 1u;
    const uint32_t NT = std::min<uint32_t>(n_paths, n);
    snkpt_status re = train_reserve(ctx, C, NT, depth);
    if (re != SNKPT_OK) return re;
    uint32_t* qc = C.tcounts.as<uint32_t>();
    k_train_clear<<<1, 64, 0, st>>>(qc);
    if (NT == 0) return SNKPT_OK;
    const size_t u32_bytes = train_u32_bytes(NT);
    const size_t sums_bytes = train_sums_bytes(NT);
    const size_t trad_off = 3 * u32_bytes + sums_bytes;
    const uint32_t indirect = (s->flags & SNKPT_FLAG_CACHE_INDIRECT) ? 1u : 0u;
    uint32_t* nrec = C.train_meta.as<uint32_t>();
    float4* trad = (float4*)(C.train_meta.as<uint8_t>() + trad_off);
    const uint32_t stride = std::max<uint32_t>(n / NT, 1u);
    const uint32_t start = (uint32_t)((initstate >> 17) % stride);
    k_fill4<<<grid_for(NT, 256), 256, 0, st>>>(trad, NT);
    ShadeHooks T;
    T.kind = 2;
    T.rec = C.train_vtx.as<float>();
    T.nrec = nrec;
    T.max_rec = depth;
    T.rad = trad;
    T.tseed = tseed;
    // the training paths are few (latency-bound per bounce as a wavefront): one thread runs each whole path
    snkpt_status e = whole_paths(ctx, cd, w, h, s, initstate ^ kTrainPathsTag, NT, start, stride, &T, trad, stats, st);
    if (e != SNKPT_OK) return e;
    uint8_t* tm = C.train_meta.as<uint8_t>();
    uint32_t* pcount = (uint32_t*)(tm + u32_bytes);
    uint32_t* off = (uint32_t*)(tm + 2 * u32_bytes);
    uint32_t* sums = (uint32_t*)(tm + 3 * u32_bytes);
    const uint64_t kseed = initstate ^ kTrainTargetTag;
    const uint32_t budget = NT;  // training samples per frame
    // with SNKPT_FLAG_CACHE_TRAIN_AT_QUERY only the vertex the cache answers is a candidate
    const uint32_t vsel = (s->flags & SNKPT_FLAG_CACHE_TRAIN_AT_QUERY)
                              ? (s->cache_query_depth ? s->cache_query_depth : 2u) - 1u
                              : ~0u;
    k_train_total<<<grid_for(NT, 256), 256, 0, st>>>(nrec, NT, vsel, qc);
    k_train_count<<<grid_for(NT, 256), 256, 0, st>>>(C.train_vtx.as<float>(), nrec, depth, trad, NT, kseed, tseed,
                                                      indirect, vsel, qc, budget, pcount);
    scan_exclusive(pcount, off, NT, sums, qc + 1, st);
    k_train_emit<<<grid_for(NT, 256), 256, 0, st>>>(C.train_vtx.as<float>(), nrec, depth, trad, NT, kseed, tseed,
                                                     indirect, vsel, qc, budget, off, C.t_samples.as<float>(), stats + 4);
    C.t_max = budget;
    return SNKPT_OK;
}

// Optimiser steps per frame on the frame's training samples (settings field; 0 = 1, at most 16).
static uint32_t train_steps(const snkpt_settings* s) {
    return s->cache_train_steps == 0 ? 1u : std::min(s->cache_train_steps, 16u);
}

// The frame's optimiser steps on the samples cache_train_paths emitted: their
// count stays on the device (tcounts[1]; at most t_max of them are emitted).
snkpt_status cache_train_step(snkpt_ctx* ctx, const snkpt_settings* s, hipStream_t st) {
    CacheState& C = *ctx->cache;
    for (uint32_t k = 0; k < train_steps(s); ++k) {
        snkpt_status e = cache_train_run(ctx, C.t_samples.as<float>(), C.t_max, C.tcounts.as<uint32_t>() + 1,
                                         s->cache_lr > 0.0f ? s->cache_lr : 1e-2f, 1,
                                         std::min(std::max(s->cache_ema, 0.0f), 0.9999f), nullptr, st);
        if (e != SNKPT_OK) return e;
    }
    return SNKPT_OK;
}

// Optimiser steps of several contexts whose training paths were traced
// (cache_train_paths), each context's steps in order on `st`; the paths' ray
// statistics (tstats) are added into `stats` first.  No readback.
snkpt_status cache_train_steps(snkpt_ctx* const* ctxs, uint32_t n, const snkpt_settings* s,
                               unsigned long long* stats, hipStream_t st) {
    for (uint32_t k = 0; k < n; ++k) {
        CacheState& C = *ctxs[k]->cache;
        if (stats) k_add_u64<<<1, 64, 0, st>>>(stats, C.tstats.as<unsigned long long>(), 8);
    }
    for (uint32_t k = 0; k < n; ++k) {
        snkpt_status e = cache_train_step(ctxs[k], s, st);
        if (e != SNKPT_OK) return e;
    }
    return SNKPT_OK;
}

// Training paths of a context forked onto `side` after `st`'s work so far;
// their statistics go to the context's own tstats (cleared first).
snkpt_status cache_train_paths_forked(snkpt_ctx* ctx, const CamDev& cd, uint32_t w, uint32_t h,
                                      const snkpt_settings* s, uint64_t initstate, uint32_t n_paths,
                                      hipStream_t side, const TrainSeedD* tseed) {
    CacheState& C = *ctx->cache;
    HIPC(C.tstats.reserve(64));
    k_clear_u64<<<1, 64, 0, side>>>(C.tstats.as<unsigned long long>(), 8);
    return cache_train_paths(ctx, cd, w, h, s, initstate, n_paths, C.tstats.as<unsigned long long>(), side, tseed);
}

// True when a cache-mode frame of w x h with these settings would allocate
// nothing: the cache exists and every buffer the frame uses (queries, work
// buffers, training paths, samples, side stream) already has its size, as after
// an uncaptured frame with the same size, depth and training paths.
bool cache_capture_ready(snkpt_ctx* ctx, const snkpt_settings* s, uint32_t w, uint32_t h) {
    const CacheState* Cp = ctx->cache;
    if (!Cp || !Cp->ready) return false;
    const CacheState& C = *Cp;
    const uint32_t n = w * h;
    const uint32_t cap = ((n + 63) / 64) * 64;
    if (C.rcap < n || C.cap < cap || C.counts.bytes < 64) return false;
    if (s->flags & SNKPT_FLAG_NO_TRAIN) return true;
    const uint32_t depth = s->max_depth ? s->max_depth : 1u;
    const uint32_t NT = std::min<uint32_t>(s->cache_train_paths ? s->cache_train_paths : 65536u, n);
    const size_t u32_bytes = ((size_t)NT * 4 + 255) & ~(size_t)255;
    const size_t sums_bytes = (((size_t)(NT + 4095) / 4096) * 4 + 255) & ~(size_t)255;
    return C.side != nullptr && C.tstats.bytes >= 64 && C.tcounts.bytes >= 64 &&
           C.train_vtx.bytes >= (size_t)NT * depth * kRecFloats * 4 &&
           C.train_meta.bytes >= 3 * u32_bytes + sums_bytes + (size_t)NT * 16 &&
           C.t_samples.bytes >= (size_t)NT * depth * 64 && C.tcap >= ((NT + 63) / 64) * 64 && C.cap >= NT;
}

snkpt_status cache_pass(snkpt_ctx* ctx, const CamDev& cd, const snkpt_settings* s, uint64_t initstate, uint32_t pass,
                        hipStream_t st) {
    RenderState& R = *ctx->render;
    const uint32_t n = R.w * R.h;
    snkpt_status e = cache_ensure(ctx, s, st);
    if (e != SNKPT_OK) return e;
    CacheState& C = *ctx->cache;
    const uint32_t depth = s->max_depth ? s->max_depth : 1u;
    // the Q-th path vertex (1 = primary hit) is answered by the cache
    const uint32_t Q = s->cache_query_depth ? s->cache_query_depth : 2u;
    if (n > C.rcap) {
        HIPC(C.q_samples.reserve((size_t)n * 64));
        HIPC(C.q_meta.reserve((size_t)n * 16));
        C.rcap = n;
    }
    HIPC(C.counts.reserve(64));
    if ((e = ensure_work(ctx, C, n, false)) != SNKPT_OK) return e;
    uint32_t* qc = C.counts.as<uint32_t>();
    unsigned long long* stats = R.stats.as<unsigned long long>();
    // 0. the training paths need no network: fork them onto the side stream so
    //    they are traced while the render pass runs (own counters and statistics)
    const bool train = !(s->flags & SNKPT_FLAG_NO_TRAIN);
    if (train) {
        if (!C.side) {
            HIPC(hipStreamCreateWithFlags(&C.side, hipStreamNonBlocking));
            HIPC(hipEventCreateWithFlags(&C.fork_ev, hipEventDisableTiming));
            HIPC(hipEventCreateWithFlags(&C.join_ev, hipEventDisableTiming));
        }
        HIPC(C.tstats.reserve(64));
        HIPC(hipEventRecord(C.fork_ev, st));
        HIPC(hipStreamWaitEvent(C.side, C.fork_ev, 0));
        k_clear_u64<<<1, 64, 0, C.side>>>(C.tstats.as<unsigned long long>(), 8);
        if ((e = cache_train_paths(ctx, cd, R.w, R.h, s, initstate,
                                   s->cache_train_paths ? s->cache_train_paths : 65536u,
                                   C.tstats.as<unsigned long long>(), C.side, nullptr)) != SNKPT_OK)
            return e;
        HIPC(hipEventRecord(C.join_ev, C.side));
    }
    // 1. render paths, terminated into the cache at vertex Q
    k_zero_u32<<<1, 64, 0, st>>>(qc, 4);
    ShadeHooks H;
    H.kind = 1;
    H.query_at = Q - 1;  // bounce index of that vertex
    H.indirect = (s->flags & SNKPT_FLAG_CACHE_INDIRECT) ? 1u : 0u;
    H.mis = (H.indirect && (s->flags & SNKPT_FLAG_CACHE_MIS)) ? 1u : 0u;
    H.q_samples = C.q_samples.as<float>();
    H.q_meta = C.q_meta.as<float4>();
    H.q_count = qc;
    // with depth >= Q the vertex before the query keeps its continuation (MIS
    // against the cache vertex's emission); with depth < Q this is plain PT.
    // The MIS probe of an indirect query vertex is one more bounce.
    const uint32_t bounces = (H.mis && depth >= Q) ? Q + 1 : std::min(depth, Q);
    if ((e = primary_bounces(ctx, cd, s, initstate, st, &H, bounces)) != SNKPT_OK) return e;
    // 2. query the cache
    const bool avg = s->cache_ema > 0.0f;  // query the weight averages
    k_encode<<<enc_blocks(n), kEncBlock, 0, st>>>(C.q_samples.as<float>(), n, qc, C.geom,
                                                  (avg ? C.grid_e16 : C.grid16).as<__half2>(), C.x0.as<_Float16>());
    const uint32_t waves = (n + 63) / 64;
    k_mlp_infer<<<(waves + kInferWaves - 1) / kInferWaves, 64 * kInferWaves, 0, st>>>(
        C.x0.as<_Float16>(), n, qc, (avg ? C.mlp_e16 : C.mlp16).as<_Float16>(), C.y.as<float4>());
    k_cache_apply<<<grid_for(n, 256), 256, 0, st>>>(C.q_meta.as<float4>(), C.y.as<float4>(), qc, R.rad.as<float4>(),
                                                      stats + 3);
    (void)pass;
    if (!train) return SNKPT_OK;
    // 3. join the training paths, then one optimiser step on their samples
    HIPC(hipStreamWaitEvent(st, C.join_ev, 0));
    k_add_u64<<<1, 64, 0, st>>>(stats, C.tstats.as<unsigned long long>(), 8);
    return cache_train_step(ctx, s, st);
}

}  // namespace snkpt
