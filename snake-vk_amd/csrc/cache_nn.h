// Neural radiance cache: sizes, input encoding and MFMA fragment helpers.
//
// No reference counterpart (the reference has no learned component).  Method:
// neural radiance caching (Mueller et al. 2021) with the multiresolution hash
// encoding of Mueller et al. 2022:
//   hash grid   16 levels, 2 features per entry, 2^19 entries per level,
//               resolutions 16 .. 2048 in a geometric progression; levels whose
//               (N+1)^3 fits are indexed densely, the rest by
//               (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod 2^19
//   inputs      64 = 32 grid features + direction, normal, roughness, albedo
//               and sin/cos(pi x) of direction and normal, zero padded
//   MLP         64 -> 64 -> 64 -> 64 -> 64 -> 3 (4 hidden ReLU layers, no biases),
//               FP16 weights and activations on MFMA with FP32 accumulation
//   loss        relative L2: (y - t)^2 / (sg(y)^2 + 0.01), mean over samples and channels
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace snkpt {

constexpr int kGridLevels = 16;
constexpr int kGridFeat = 2;
constexpr uint32_t kGridLog2T = 19;
constexpr uint32_t kGridT = 1u << kGridLog2T;
constexpr int kWidth = 64;
constexpr int kHidden = 4;
constexpr int kOut = 3;
constexpr int kOutPad = 16;
// MLP parameter layout (fp32 master and fp16 copies): W0..W3 [64][64], W4 [16][64]
constexpr int kLayerParams = kWidth * kWidth;                          // 4096
constexpr int kMlpParams = kHidden * kLayerParams + kOutPad * kWidth;  // 17408
constexpr size_t kGridParams = (size_t)kGridLevels * kGridT * kGridFeat;  // 16.7M
constexpr int kSampleFloats = 16;  // pos3 dir3 nrm3 rough albedo3 target3
constexpr float kLossScale = 1024.0f;

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

struct GridGeom {
    float lo[3];
    float inv_ext;            // 1 / largest extent of the scene box
    uint32_t res[kGridLevels];
    uint32_t dense[kGridLevels];  // 1: dense indexing at this level
};

// One context's network as the multi-network query kernels read it: grid
// geometry, FP16 grid features and FP16 MLP weights (the live weights or their
// averages).
struct CacheRefD {
    GridGeom geom;
    const __half2* grid;
    const _Float16* mlp;
};

__host__ __device__ __forceinline__ uint32_t grid_index(uint32_t x, uint32_t y, uint32_t z, uint32_t res,
                                                        uint32_t dense) {
    if (dense) {
        uint32_t n = res + 1;
        return x + y * n + z * n * n;
    }
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (kGridT - 1u);
}

}  // namespace snkpt
