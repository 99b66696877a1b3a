// GPU LBVH -> BVH8 builder.
//
// Replaces the driver builds of BLAS::GenerateFromMeshData / TLAS::BuildFromInstances
// (Core/src/resources/TLAS.cpp:10-200, TlasSystem.cpp:27-45).  Instead of a
// two-level BLAS/TLAS the instances are flattened into one world-space
// triangle set (instance and primitive ids stay in the leaves), then:
//   1. k_flatten       object -> world triangles (48 B leaf records, input order)
//   2. k_bounds        centroid and scene bounds (order-independent min/max)
//   3. k_morton        30-bit Morton code of each triangle's box centre
//   4. radix sort      LDS-staged stable LSD sort, 4 x 8-bit digits
//   5. k_karras        binary radix tree (Karras 2012), duplicate keys ordered by index
//   6. k_bottom_level  boxes, SAH cost and leaf-cluster flags (clusters of <= 3 triangles),
//                      one tree level per launch
//   7. collapse        level-synchronous greedy collapse to BVH8 (largest-area child
//                      expanded first), octant slot order, 8-bit quantised child boxes
// Every integer output and every FP32 box is reproduced bit for bit by the CPU
// twin; allocation is by prefix sums, never by atomics, so the layout is
// deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "snkpt_dev.h"
#include "snkpt_host.h"

namespace snkpt {
namespace {

constexpr uint32_t kLeafTag = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
// SAH costs of a BVH node visit and of a leaf triangle test.  A triangle costs
// two node visits: leaves then hold mostly one triangle, and the traversal
// (one node or one triangle per step) spends fewer steps per ray.  On
// the 1.29M incoherent rays of the 1.02M-triangle scene (nodes / triangles
// visited per ray, Grays/s): 0.3 -> 12.63 / 5.18, 3.33; 1.0 -> 13.10 / 3.00,
// 3.53; 2.0 -> 13.41 / 2.26, 3.59; 3.0 -> 13.42 / 2.26, 3.62; hits unchanged.
constexpr float kCostNode = 1.0f;
constexpr float kCostTri = 2.0f;
constexpr uint32_t kMaxLeafTris = 3;

// ---------------------------------------------------------------- scan -----
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanBlock * kScanItems;  // 4096

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < nw; ++w) {
            uint32_t t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[nw] = run;
    }
    __syncthreads();
    uint32_t r = x - v + lds[wave];
    if (total) *total = lds[nw];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                            uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems, s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_exclusive_scan(s, lds, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// one block: exclusive scan of nb <= kScanTile values (in place when in == out),
// grand total to *total
__global__ __launch_bounds__(kScanBlock) void k_scan_top(const uint32_t* in, uint32_t* out, uint32_t nb,
                                                         uint32_t* __restrict__ total) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    uint32_t base = threadIdx.x * kScanItems, v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < nb ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, lds, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < nb) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const uint32_t* __restrict__ in,
                                                          uint32_t* __restrict__ out, uint32_t n,
                                                          const uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems, v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t run = block_exclusive_scan(s, lds, nullptr) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}


// ---------------------------------------------------------- radix sort -----
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr uint32_t kSortTile = kSortBlock * kSortItems;  // 4096 keys per block

__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                           uint32_t shift, uint32_t* __restrict__ hist,
                                                           uint32_t nb) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        uint32_t i = base + k * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter: keys of a tile are ranked in index order (item k of thread t
// is element k*256 + t), staged sorted in LDS, then written out in runs.
__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout,
                                                              uint32_t* __restrict__ vout, uint32_t n,
                                                              uint32_t shift, const uint32_t* __restrict__ hist,
                                                              const uint32_t* __restrict__ hist_scan, uint32_t nb) {
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_wcnt[kSortBlock / 64][256];
    __shared__ uint32_t lds[kSortBlock / 64 + 1];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t b = blockIdx.x, base = b * kSortTile;
    const uint32_t tile_n = n - base < kSortTile ? n - base : kSortTile;
    uint32_t my_count = hist[t * nb + b];
    s_start[t] = block_exclusive_scan(my_count, lds, nullptr);
    s_run[t] = 0;
#pragma unroll
    for (int w = 0; w < kSortBlock / 64; ++w) s_wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < kSortItems; ++k) {
        uint32_t i = base + k * kSortBlock + t;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
        uint32_t d = (key >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            uint64_t bm = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bm : ~bm;
        }
        uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) s_wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_start[d] + s_run[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += s_wcnt[w][d];
            s_keys[pos] = key;
            s_vals[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kSortBlock / 64; ++w) {
            add += s_wcnt[w][t];
            s_wcnt[w][t] = 0;
        }
        s_run[t] += add;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        uint32_t i = k * kSortBlock + t;
        if (i < tile_n) {
            uint32_t key = s_keys[i], d = (key >> shift) & 0xFFu;
            uint32_t g = hist_scan[d * nb + b] + (i - s_start[d]);
            kout[g] = key;
            vout[g] = s_vals[i];
        }
    }
}

// ------------------------------------------------------------- flatten -----
struct InstDev {
    uint32_t transform_idx, index_offset, vertex_offset, material_idx, num_indices;
};

__global__ void k_flatten(uint32_t n, const uint32_t* __restrict__ inst_off, uint32_t n_inst,
                          const InstDev* __restrict__ inst, const float* __restrict__ mat4s,
                          const uint32_t* __restrict__ idx, const float* __restrict__ pos,
                          float4* __restrict__ out) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    // largest instance i with inst_off[i] <= g
    uint32_t lo = 0, hi = n_inst;  // invariant: inst_off[lo] <= g < inst_off[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (inst_off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    InstDev in = inst[lo];
    uint32_t prim = g - inst_off[lo];
    const float* m = mat4s + 16 * (size_t)in.transform_idx;
    float w[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t v = idx[in.index_offset + 3 * prim + k] + in.vertex_offset;
        float x = pos[3 * (size_t)v], y = pos[3 * (size_t)v + 1], z = pos[3 * (size_t)v + 2];
        w[3 * k + 0] = ((m[0] * x + m[4] * y) + m[8] * z) + m[12];
        w[3 * k + 1] = ((m[1] * x + m[5] * y) + m[9] * z) + m[13];
        w[3 * k + 2] = ((m[2] * x + m[6] * y) + m[10] * z) + m[14];
    }
    float4* o = out + 3 * (size_t)g;
    o[0] = make_float4(w[0], w[1], w[2], w[3]);
    o[1] = make_float4(w[4], w[5], w[6], w[7]);
    o[2] = make_float4(w[8], __uint_as_float(lo), __uint_as_float(prim), __uint_as_float(g));
}

// --------------------------------------------------------------- bounds ----
__device__ __forceinline__ uint32_t fenc(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ void tri_box(const float4* __restrict__ t, uint32_t g, float lo[3], float hi[3]) {
    float4 r0 = t[3 * (size_t)g], r1 = t[3 * (size_t)g + 1], r2 = t[3 * (size_t)g + 2];
    float v[9] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
        hi[a] = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
    }
}

// bounds[0..2] centroid min, [3..5] centroid max, [6..8] scene min, [9..11] scene max (encoded)
__global__ void k_bounds_init(uint32_t* b) {
    if (threadIdx.x < 12) b[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

__global__ __launch_bounds__(256) void k_bounds(const float4* __restrict__ tris, uint32_t n,
                                                uint32_t* __restrict__ bounds) {
    __shared__ float red[12][4];
    // grid-stride: few blocks, so few same-address atomics at the end
    float v[12];
    for (int a = 0; a < 3; ++a) {
        v[a] = v[6 + a] = INFINITY;
        v[3 + a] = v[9 + a] = -INFINITY;
    }
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        float lo[3], hi[3];
        tri_box(tris, g, lo, hi);
        for (int a = 0; a < 3; ++a) {
            float c = (lo[a] + hi[a]) * 0.5f;
            v[a] = fminf(v[a], c);
            v[3 + a] = fmaxf(v[3 + a], c);
            v[6 + a] = fminf(v[6 + a], lo[a]);
            v[9 + a] = fmaxf(v[9 + a], hi[a]);
        }
    }
    for (int q = 0; q < 12; ++q) {
        bool is_min = (q % 6) < 3;
        float x = v[q];
        for (int off = 32; off >= 1; off >>= 1) {
            float y = __shfl_xor(x, off, 64);
            x = is_min ? fminf(x, y) : fmaxf(x, y);
        }
        v[q] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 12; ++q) red[q][wave] = v[q];
    __syncthreads();
    if (threadIdx.x < 12) {
        int q = threadIdx.x;
        bool is_min = (q % 6) < 3;
        float x = red[q][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x = is_min ? fminf(x, red[q][w]) : fmaxf(x, red[q][w]);
        if (is_min)
            atomicMin(&bounds[q], fenc(x));
        else
            atomicMax(&bounds[q], fenc(x));
    }
}

// --------------------------------------------------------------- morton ----
__device__ __forceinline__ uint32_t expand_bits10(uint32_t x) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void k_morton(const float4* __restrict__ tris, uint32_t n, const uint32_t* __restrict__ bounds,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float lo[3], hi[3];
    tri_box(tris, g, lo, hi);
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float clo = fdec(bounds[a]), chi = fdec(bounds[3 + a]);
        float ext = chi - clo;
        float inv = ext > 0.0f ? 1024.0f / ext : 0.0f;
        float c = (lo[a] + hi[a]) * 0.5f;
        float f = fminf(fmaxf((c - clo) * inv, 0.0f), 1023.0f);
        q[a] = (uint32_t)f;
    }
    keys[g] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    ids[g] = g;
}

// --------------------------------------------------------------- karras ----
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

__global__ void k_karras(const uint32_t* __restrict__ keys, int n, uint32_t* __restrict__ child,
                         uint32_t* __restrict__ first, uint32_t* __restrict__ count,
                         uint32_t* __restrict__ parent_int, uint32_t* __restrict__ parent_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(keys, n, i, j);
    int s = 0, t, div = 2;
    do {
        t = (l + div - 1) / div;
        if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        div <<= 1;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t left = (lo == gamma) ? (kLeafTag | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t right = (hi == gamma + 1) ? (kLeafTag | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    first[i] = (uint32_t)lo;
    count[i] = (uint32_t)(hi - lo + 1);
    if (left & kLeafTag)
        parent_leaf[left & ~kLeafTag] = (uint32_t)i;
    else
        parent_int[left] = (uint32_t)i;
    if (right & kLeafTag)
        parent_leaf[right & ~kLeafTag] = (uint32_t)i;
    else
        parent_int[right] = (uint32_t)i;
    if (i == 0) parent_int[0] = kNone;
}

// ----------------------------------------------------------- bottom-up -----
__device__ __forceinline__ float box_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

// aabb layout: internal node i at 6*i, leaf j at 6*(n_int + j); lo xyz then hi xyz
__global__ void k_leaf_boxes(const float4* __restrict__ tris, const uint32_t* __restrict__ ids, uint32_t n,
                             uint32_t n_int, float* __restrict__ aabb) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float lo[3], hi[3];
    tri_box(tris, ids[j], lo, hi);
    float* o = aabb + 6 * (size_t)(n_int + j);
    o[0] = lo[0];
    o[1] = lo[1];
    o[2] = lo[2];
    o[3] = hi[0];
    o[4] = hi[1];
    o[5] = hi[2];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct BinUp {
    uint32_t n, n_int;
    const uint32_t* child;
    const uint32_t* count;
    const uint32_t* parent_int;
    const uint32_t* parent_leaf;
    uint32_t* flags;
    float* aabb;
    float* cost;
    uint32_t* height;
    uint32_t* leafify;
    float* dpc;     // C(x, i), i = 1..8, of the BVH8 collapse
    uint32_t* dpk;  // its split choices, 3 bits per i
};

// SAH cost of a binary subtree spread over i slots of a BVH8 node (Ylitie,
// Karras & Laine 2017, sec. 3.1), for the leaves the SAH pass fixed:
//   leaf (a triangle or a leaf cluster of N):  C(x, i) = c_tri N A(x)
//   internal x:  D(x, j) = min_{1 <= k < j} C(l, k) + C(r, j - k)   (x opened into j slots)
//                C(x, 1) = c_node A(x) + D(x, 8)                    (x becomes a BVH8 node)
//                C(x, i) = min(C(x, i - 1), D(x, i))  for i >= 2
// k ascending, strict comparisons: the CPU twin makes the same choices.
// dpk holds the k of D(x, 8) for i = 1 and, for i >= 2, the k of D(x, i) or 0
// when C(x, i) = C(x, i - 1).
__device__ __forceinline__ void collapse_costs(const float Cl[8], const float Cr[8], float area, float C[8],
                                               uint32_t& K) {
    float D[9];
    uint32_t Kd[9];
#pragma unroll
    for (int j = 2; j <= 8; ++j) {
        D[j] = __builtin_inff();
        Kd[j] = 1u;
#pragma unroll
        for (int k = 1; k < j; ++k) {
            const float v = Cl[k - 1] + Cr[j - k - 1];
            if (v < D[j]) {
                D[j] = v;
                Kd[j] = (uint32_t)k;
            }
        }
    }
    C[0] = kCostNode * area + D[8];
    K = Kd[8];
#pragma unroll
    for (int i = 2; i <= 8; ++i) {
        if (D[i] < C[i - 2]) {
            C[i - 1] = D[i];
            K |= Kd[i] << (3 * (i - 1));
        } else {
            C[i - 1] = C[i - 2];
        }
    }
}

// Box, SAH cost, height and leaf decision of internal node `node` from its two
// children (both already final).
__device__ __forceinline__ void bin_node(const BinUp& U, uint32_t node) {
    float bl[6], br[6], cl, cr;
    uint32_t hl, hr;
    uint32_t c0 = U.child[2 * node], c1 = U.child[2 * node + 1];
    {
        uint32_t k = (c0 & kLeafTag) ? U.n_int + (c0 & ~kLeafTag) : c0;
        for (int q = 0; q < 6; ++q) bl[q] = U.aabb[6 * (size_t)k + q];
        cl = (c0 & kLeafTag) ? kCostTri * box_area(bl) : U.cost[c0];
        hl = (c0 & kLeafTag) ? 1u : U.height[c0];
    }
    {
        uint32_t k = (c1 & kLeafTag) ? U.n_int + (c1 & ~kLeafTag) : c1;
        for (int q = 0; q < 6; ++q) br[q] = U.aabb[6 * (size_t)k + q];
        cr = (c1 & kLeafTag) ? kCostTri * box_area(br) : U.cost[c1];
        hr = (c1 & kLeafTag) ? 1u : U.height[c1];
    }
    float b[6];
    for (int a = 0; a < 3; ++a) {
        b[a] = fminf(bl[a], br[a]);
        b[3 + a] = fmaxf(bl[3 + a], br[3 + a]);
    }
    float area = box_area(b);
    float c_split = kCostNode * area + (cl + cr);
    uint32_t cnt = U.count[node];
    uint32_t lf = 0;
    float c = c_split;
    if (cnt <= kMaxLeafTris) {
        float c_leaf = (kCostTri * (float)cnt) * area;
        if (c_leaf <= c_split) {
            lf = 1;
            c = c_leaf;
        }
    }
    for (int q = 0; q < 6; ++q) U.aabb[6 * (size_t)node + q] = b[q];
    U.cost[node] = c;
    U.height[node] = (hl > hr ? hl : hr) + 1u;
    U.leafify[node] = lf;
    // collapse costs: a triangle leaf's are its SAH cost cl / cr, an internal
    // child's were written when it was made
    float Cl[8], Cr[8], C[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        Cl[i] = (c0 & kLeafTag) ? cl : U.dpc[8 * (size_t)c0 + i];
        Cr[i] = (c1 & kLeafTag) ? cr : U.dpc[8 * (size_t)c1 + i];
    }
    uint32_t K = 0;
    if (lf) {
#pragma unroll
        for (int i = 0; i < 8; ++i) C[i] = c;
    } else {
        collapse_costs(Cl, Cr, area, C, K);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) U.dpc[8 * (size_t)node + i] = C[i];
    U.dpk[node] = K;
}

// One level of the bottom-up pass per launch.  Each live node (every leaf in
// the first launch, afterwards the internal nodes the previous launch made)
// arrives at its parent; the second arrival computes the parent, which is live
// in the next launch.  Both children of a node are therefore written by
// earlier launches, and the kernel boundary makes them visible to the thread
// that reads them: no device-scope fences (each one writes back and
// invalidates a whole L2) on this many-XCD part.  Counts are read on the
// device, so the fixed sequence of launches needs no host round trip; the
// order of the live list does not affect any value.
constexpr int kUpBlock = 256;
constexpr uint32_t kMaxBinLevels = 64;  // radix tree over 64-bit (key, index) codes: height <= 65

template <bool FIRST>
__global__ __launch_bounds__(kUpBlock) void k_bottom_level(BinUp U, const uint32_t* __restrict__ live_in,
                                                           const uint32_t* __restrict__ cnt_in,
                                                           uint32_t* __restrict__ live_out,
                                                           uint32_t* __restrict__ cnt_out) {
    __shared__ uint32_t lds[kUpBlock / 64 + 1];
    const uint32_t total = FIRST ? U.n : *cnt_in;
    for (uint32_t base = blockIdx.x * kUpBlock; base < total; base += gridDim.x * kUpBlock) {
        const uint32_t i = base + threadIdx.x;
        bool made = false;
        uint32_t parent = kNone;
        if (i < total) {
            parent = FIRST ? U.parent_leaf[i] : U.parent_int[live_in[i]];
            if (parent != kNone && atomicAdd(&U.flags[parent], 1u) != 0u) {
                bin_node(U, parent);
                made = true;
            }
        }
        const uint32_t slot = block_append<kUpBlock>(made, cnt_out, lds);
        if (made) live_out[slot] = parent;
    }
}

// ------------------------------------------------------------- collapse ----
struct BinView {
    uint32_t n, n_int;
    const uint32_t* child;
    const uint32_t* first;
    const uint32_t* count;
    const float* aabb;
    const uint32_t* leafify;
    const uint32_t* dpk;
    __device__ const float* box(uint32_t tagged) const {
        uint32_t k = (tagged & kLeafTag) ? n_int + (tagged & ~kLeafTag) : tagged;
        return aabb + 6 * (size_t)k;
    }
    __device__ bool expandable(uint32_t tagged) const { return !(tagged & kLeafTag) && leafify[tagged] == 0u; }
    __device__ void range(uint32_t tagged, uint32_t& f, uint32_t& c) const {
        if (tagged & kLeafTag) {
            f = tagged & ~kLeafTag;
            c = 1;
        } else {
            f = first[tagged];
            c = count[tagged];
        }
    }
};

// Children of a BVH8 node made from binary node `b`, in slot order.  Every
// per-child array is indexed only inside fully unrolled loops, so the eight
// candidates, their areas and centres and the 8x8 octant cost table live in
// VGPRs (dynamically indexed private arrays go to scratch), and the two
// children of an expanded node are fetched with independent loads.
__device__ __forceinline__ void child_info(const BinView& bv, uint32_t t, bool& ex, float& area, float& cx,
                                           float& cy, float& cz) {
    const float* bx = bv.box(t);
    float bb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) bb[q] = bx[q];
    ex = bv.expandable(t);
    area = ex ? box_area(bb) : -1.0f;
    cx = (bb[0] + bb[3]) * 0.5f;
    cy = (bb[1] + bb[4]) * 0.5f;
    cz = (bb[2] + bb[5]) * 0.5f;
}

__global__ __launch_bounds__(128) void k_collapse_expand(BinView bv, const uint32_t* __restrict__ work,
                                                         uint32_t level_start, uint32_t level_count,
                                                         uint32_t* __restrict__ slot_child,
                                                         uint32_t* __restrict__ n_int, uint32_t* __restrict__ n_tri) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= level_count) return;
    const uint32_t b = work[level_start + i];
    uint32_t S[8];
    float A[8], X[8], Y[8], Z[8];
    uint32_t em = 0;  // bit k: S[k] is expandable
    int ns;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        S[k] = kNone;
        A[k] = X[k] = Y[k] = Z[k] = 0.0f;
    }
    {
        bool ex;
        child_info(bv, b, ex, A[0], X[0], Y[0], Z[0]);
        S[0] = b;
        ns = 1;
        if (ex) {
            // the SAH collapse: the 8 slots are split between b's children as
            // D(b, 8) chose; then the first candidate that the costs open (its
            // slot count i > 1 after dropping the i whose cost equals i - 1's)
            // is replaced by its left child and its right child appended, until
            // every candidate is a leaf, a BVH8 node of its own (i = 1) or done
            const uint32_t kb = bv.dpk[b] & 7u;
            uint32_t I[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) I[k] = 0u;
            uint32_t l = bv.child[2 * b], r = bv.child[2 * b + 1];
            bool el, er;
            child_info(bv, l, el, A[0], X[0], Y[0], Z[0]);
            child_info(bv, r, er, A[1], X[1], Y[1], Z[1]);
            S[0] = l;
            S[1] = r;
            I[0] = kb;
            I[1] = 8u - kb;
            em = (el ? 1u : 0u) | (er ? 2u : 0u);
            ns = 2;
            uint32_t done = 0;  // bit k: candidate k stays as it is
            while (true) {
                int cur = -1;
#pragma unroll
                for (int k = 7; k >= 0; --k)
                    if (k < ns && !((done >> k) & 1u)) cur = k;
                if (cur < 0) break;
                uint32_t x = 0, i = 0;
                bool xe = false;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k == cur) {
                        x = S[k];
                        i = I[k];
                        xe = (em >> k) & 1u;
                    }
                if (!xe) {
                    done |= 1u << cur;
                    continue;
                }
                const uint32_t kx = bv.dpk[x];
                while (i > 1u && ((kx >> (3u * (i - 1u))) & 7u) == 0u) --i;
                if (i == 1u) {
                    done |= 1u << cur;
                    continue;
                }
                const uint32_t kk = (kx >> (3u * (i - 1u))) & 7u;
                const uint32_t cl = bv.child[2 * x], cr = bv.child[2 * x + 1];
                bool xl, xr;
                float al, ar, lx, ly, lz, rx, ry, rz;
                child_info(bv, cl, xl, al, lx, ly, lz);
                child_info(bv, cr, xr, ar, rx, ry, rz);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (k == cur) {
                        S[k] = cl;
                        I[k] = kk;
                        A[k] = al;
                        X[k] = lx;
                        Y[k] = ly;
                        Z[k] = lz;
                        em = (em & ~(1u << k)) | (xl ? 1u << k : 0u);
                    }
                    if (k == ns) {
                        S[k] = cr;
                        I[k] = i - kk;
                        A[k] = ar;
                        X[k] = rx;
                        Y[k] = ry;
                        Z[k] = rz;
                        em = (em & ~(1u << k)) | (xr ? 1u << k : 0u);
                    }
                }
                ++ns;
            }
        }
    }
    // octant slot assignment: greedy on cost[c][s] = sum_a (+-1)(centre_c - centre_node)_a
    const float* nb = bv.box(b);
    const float nc0 = (nb[0] + nb[3]) * 0.5f, nc1 = (nb[1] + nb[4]) * 0.5f, nc2 = (nb[2] + nb[5]) * 0.5f;
    float cost[8][8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float dx = X[c] - nc0, dy = Y[c] - nc1, dz = Z[c] - nc2;
#pragma unroll
        for (int s = 0; s < 8; ++s) cost[c][s] = ((s & 1 ? dx : -dx) + (s & 2 ? dy : -dy)) + (s & 4 ? dz : -dz);
    }
    uint32_t slot_used = 0, child_used = 0;
    uint32_t slots[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) slots[s] = kNone;
    for (int r = 0; r < ns; ++r) {
        int bc = -1, bs = -1;
        float bval = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (c >= ns || (child_used & (1u << c))) continue;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                if (slot_used & (1u << s)) continue;
                if (bc < 0 || cost[c][s] > bval) {
                    bval = cost[c][s];
                    bc = c;
                    bs = s;
                }
            }
        }
        child_used |= 1u << bc;
        slot_used |= 1u << bs;
        uint32_t chosen = kNone;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c == bc) chosen = S[c];
#pragma unroll
        for (int s = 0; s < 8; ++s)
            if (s == bs) slots[s] = chosen;
    }
    uint32_t ni = 0, nt = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        slot_child[8 * (size_t)i + s] = slots[s];
        if (slots[s] == kNone) continue;
        if (bv.expandable(slots[s]))
            ++ni;
        else {
            uint32_t f, c;
            bv.range(slots[s], f, c);
            nt += c;
        }
    }
    n_int[i] = ni;
    n_tri[i] = nt;
}

__device__ __forceinline__ int quant_exp(float lo, float hi) {
    double ext = (double)hi - (double)lo;
    int e = -126;
    if (ext > 0.0) {
        int k;
        (void)frexp(ext / 255.0, &k);
        e = k;
        while (e > -126 && ldexp(255.0, e - 1) >= ext) --e;
        while (ldexp(255.0, e) < ext) ++e;
    }
    if (e < -126) e = -126;
    if (e > 127) e = 127;
    return e;
}

__device__ __forceinline__ uint32_t quant_lo(float c, float p, int e) {
    double q = floor(ldexp((double)c - (double)p, -e));
    q = q < 0.0 ? 0.0 : (q > 255.0 ? 255.0 : q);
    return (uint32_t)q;
}
__device__ __forceinline__ uint32_t quant_hi(float c, float p, int e) {
    double q = ceil(ldexp((double)c - (double)p, -e));
    q = q < 0.0 ? 0.0 : (q > 255.0 ? 255.0 : q);
    return (uint32_t)q;
}

__global__ void k_collapse_emit(BinView bv, uint32_t* __restrict__ work, uint32_t level_start,
                                uint32_t level_count, const uint32_t* __restrict__ slot_child,
                                const uint32_t* __restrict__ scan_int, const uint32_t* __restrict__ scan_tri,
                                uint32_t next_level_start, uint32_t level_tri_start,
                                const uint32_t* __restrict__ ids, const float4* __restrict__ world,
                                uint4* __restrict__ nodes, float4* __restrict__ leaf) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= level_count) return;
    uint32_t b = work[level_start + i];
    const float* nb = bv.box(b);
    float p[3] = {nb[0], nb[1], nb[2]};
    int e[3];
    for (int a = 0; a < 3; ++a) e[a] = quant_exp(nb[a], nb[3 + a]);
    uint32_t child_base = next_level_start + scan_int[i];
    uint32_t tri_base = level_tri_start + scan_tri[i];
    uint32_t imask = 0, meta[2] = {0, 0}, ql[3][2] = {{0, 0}, {0, 0}, {0, 0}}, qh[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    uint32_t ci = 0, toff = 0;
    for (int s = 0; s < 8; ++s) {
        uint32_t c = slot_child[8 * (size_t)i + s];
        if (c == kNone) continue;
        const float* cb = bv.box(c);
        int w = s >> 2, sh = 8 * (s & 3);
        for (int a = 0; a < 3; ++a) {
            ql[a][w] |= quant_lo(cb[a], p[a], e[a]) << sh;
            qh[a][w] |= quant_hi(cb[3 + a], p[a], e[a]) << sh;
        }
        uint32_t m;
        if (bv.expandable(c)) {
            imask |= 1u << s;
            m = 0x20u | (24u + (uint32_t)s);
            work[child_base + ci] = c;
            ++ci;
        } else {
            uint32_t f, cnt;
            bv.range(c, f, cnt);
            uint32_t unary = (1u << cnt) - 1u;
            m = (unary << 5) | toff;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4* src = world + 3 * (size_t)ids[f + k];
                float4* dst = leaf + 3 * (size_t)(tri_base + toff + k);
                dst[0] = src[0];
                dst[1] = src[1];
                dst[2] = src[2];
            }
            toff += cnt;
        }
        meta[w] |= m << sh;
    }
    uint4* o = nodes + (size_t)(level_start + i) * kNodeWords;
    uint32_t ew = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16) |
                  (imask << 24);
    o[0] = make_uint4(__float_as_uint(p[0]), __float_as_uint(p[1]), __float_as_uint(p[2]), ew);
    o[1] = make_uint4(child_base, tri_base, meta[0], meta[1]);
    o[2] = make_uint4(ql[0][0], ql[0][1], ql[1][0], ql[1][1]);
    o[3] = make_uint4(ql[2][0], ql[2][1], qh[0][0], qh[0][1]);
    o[4] = make_uint4(qh[1][0], qh[1][1], qh[2][0], qh[2][1]);
}

inline uint32_t blocks(uint32_t n, uint32_t b) { return n == 0 ? 1 : (n + b - 1) / b; }

}  // namespace

// Exclusive scan of n u32 (n <= 4096 * 4096); `sums` needs ceil(n/4096) entries.
// The grand total is written to *total (device) when total != nullptr.
void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, uint32_t* total,
                    hipStream_t st) {
    uint32_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {  // one tile: a single block does it
        k_scan_top<<<1, kScanBlock, 0, st>>>(in, out, n, total);
        return;
    }
    k_scan_reduce<<<nb, kScanBlock, 0, st>>>(in, n, sums);
    k_scan_top<<<1, kScanBlock, 0, st>>>(sums, sums, nb, total);
    k_scan_down<<<nb, kScanBlock, 0, st>>>(in, out, n, sums);
}

#define HIP_CHECK(expr)                                          \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr);   \
    } while (0)

snkpt_status build_accel(snkpt_ctx* ctx, hipStream_t st) {
    Accel& A = ctx->accel;
    A.built = false;
    // ---- validate every index the kernels will follow
    const uint32_t n_inst = (uint32_t)ctx->inst.size();
    const uint64_t n_verts = ctx->pos.size() / 3, n_idx = ctx->idx.size();
    const uint32_t n_mat4 = (uint32_t)(ctx->mat4s.size() / 16);
    std::vector<uint32_t> off(n_inst + 1, 0);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
        const snkpt_instance& in = ctx->inst[i];
        if (in.num_mesh_indices % 3 != 0 || in.mesh_buffer_index_offset % 3 != 0)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": index counts must be multiples of 3");
        if ((uint64_t)in.mesh_buffer_index_offset + in.num_mesh_indices > n_idx)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": index range outside the index buffer");
        if (in.transform_idx >= n_mat4)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": transform_idx out of range");
        if (in.material_idx >= ctx->mats.size())
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": material_idx out of range");
        // largest index of the instance's index range, computed once per range
        // until the meshes change (a per-frame rebuild does not rescan them)
        const uint64_t key = ((uint64_t)in.mesh_buffer_index_offset << 32) | in.num_mesh_indices;
        auto it = ctx->range_max.find(key);
        if (it == ctx->range_max.end()) {
            uint32_t mx = 0;
            for (uint32_t k = 0; k < in.num_mesh_indices; ++k) mx = std::max(mx, ctx->idx[in.mesh_buffer_index_offset + k]);
            it = ctx->range_max.emplace(key, mx).first;
        }
        if (in.num_mesh_indices && (uint64_t)it->second + in.mesh_buffer_vertex_offset >= n_verts)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": vertex index out of range");
        off[i] = (uint32_t)total;
        total += in.num_mesh_indices / 3;
    }
    off[n_inst] = (uint32_t)total;
    if (total >= (1ull << 31) || total > 4096ull * 4096ull)
        return fail(ctx, SNKPT_ERR_CAPACITY, "too many triangles for one acceleration structure");
    const uint32_t n = (uint32_t)total;
    A.num_tris = n;
    A.num_nodes = 0;
    A.depth = 0;
    A.binary_depth = 0;

    HIP_CHECK(A.inst_tri_offset.reserve((n_inst + 1) * sizeof(uint32_t)));
    HIP_CHECK(hipMemcpyAsync(A.inst_tri_offset.ptr, off.data(), (n_inst + 1) * sizeof(uint32_t),
                             hipMemcpyHostToDevice, st));
    hipEvent_t ev0, ev1;
    HIP_CHECK(hipEventCreate(&ev0));
    HIP_CHECK(hipEventCreate(&ev1));
    HIP_CHECK(hipEventRecord(ev0, st));

    const size_t nn = n ? n : 1;
    HIP_CHECK(A.world_tris.reserve(nn * 48));
    HIP_CHECK(A.leaf_tris.reserve(nn * 48));
    HIP_CHECK(A.nodes.reserve((nn + 1) * 80));
    HIP_CHECK(A.keys.reserve(nn * 4));
    HIP_CHECK(A.ids.reserve(nn * 4));
    HIP_CHECK(A.bin_child.reserve(nn * 8));
    HIP_CHECK(A.bin_aabb.reserve((2 * nn) * 24));
    HIP_CHECK(A.bin_leafify.reserve(nn * 4));
    HIP_CHECK(A.bin_dpc.reserve(nn * 32));
    HIP_CHECK(A.bin_dpk.reserve(nn * 4));

    DeviceBuffer &keys2 = A.keys2, &ids2 = A.ids2, &hist = A.hist, &hist_scan = A.hist_scan, &sums = A.sums,
                 &bounds = A.bounds, &first = A.first, &count = A.count, &pint = A.pint, &pleaf = A.pleaf,
                 &flags = A.flags, &up_live = A.up_live, &up_cnt = A.up_cnt, &cost = A.cost, &height = A.height,
                 &slot_child = A.slot_child, &nint = A.nint, &ntri = A.ntri, &sint = A.sint, &stri = A.stri,
                 &work = A.work, &totals = A.totals;
    const uint32_t nb_sort = blocks(n, kSortTile);
    HIP_CHECK(keys2.reserve(nn * 4));
    HIP_CHECK(ids2.reserve(nn * 4));
    HIP_CHECK(hist.reserve((size_t)256 * nb_sort * 4));
    HIP_CHECK(hist_scan.reserve((size_t)256 * nb_sort * 4));
    HIP_CHECK(sums.reserve((blocks((uint32_t)std::max<size_t>(256ull * nb_sort, nn), kScanTile) + 1) * 4));
    HIP_CHECK(bounds.reserve(12 * 4));
    HIP_CHECK(first.reserve(nn * 4));
    HIP_CHECK(count.reserve(nn * 4));
    HIP_CHECK(pint.reserve(nn * 4));
    HIP_CHECK(pleaf.reserve(nn * 4));
    HIP_CHECK(flags.reserve(nn * 4));
    HIP_CHECK(up_live.reserve(nn * 8));
    HIP_CHECK(up_cnt.reserve((kMaxBinLevels + 1) * 4));
    HIP_CHECK(cost.reserve(nn * 4));
    HIP_CHECK(height.reserve(nn * 4));
    HIP_CHECK(slot_child.reserve(nn * 32));
    HIP_CHECK(nint.reserve(nn * 4));
    HIP_CHECK(ntri.reserve(nn * 4));
    HIP_CHECK(sint.reserve(nn * 4));
    HIP_CHECK(stri.reserve(nn * 4));
    HIP_CHECK(work.reserve((nn + 1) * 4));
    HIP_CHECK(totals.reserve(16));

    uint32_t hb[12];
    const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
    if (n > 0) {
        k_flatten<<<blocks(n, 256), 256, 0, st>>>(n, A.inst_tri_offset.as<uint32_t>(), n_inst,
                                                    ctx->d_inst.as<InstDev>(), ctx->d_mat4.as<float>(),
                                                    ctx->d_idx.as<uint32_t>(), ctx->d_pos.as<float>(),
                                                    A.world_tris.as<float4>());
        k_bounds_init<<<1, 64, 0, st>>>(bounds.as<uint32_t>());
        k_bounds<<<std::min(blocks(n, 256), cus * 2u), 256, 0, st>>>(A.world_tris.as<float4>(), n,
                                                                   bounds.as<uint32_t>());
        k_morton<<<blocks(n, 256), 256, 0, st>>>(A.world_tris.as<float4>(), n, bounds.as<uint32_t>(),
                                                   A.keys.as<uint32_t>(), A.ids.as<uint32_t>());
        // radix sort, 4 passes, result back in A.keys / A.ids
        uint32_t *ka = A.keys.as<uint32_t>(), *va = A.ids.as<uint32_t>();
        uint32_t *kb = keys2.as<uint32_t>(), *vb = ids2.as<uint32_t>();
        for (uint32_t shift = 0; shift < 32; shift += 8) {
            k_radix_hist<<<nb_sort, kSortBlock, 0, st>>>(ka, n, shift, hist.as<uint32_t>(), nb_sort);
            scan_exclusive(hist.as<uint32_t>(), hist_scan.as<uint32_t>(), 256 * nb_sort, sums.as<uint32_t>(),
                           nullptr, st);
            k_radix_scatter<<<nb_sort, kSortBlock, 0, st>>>(ka, va, kb, vb, n, shift, hist.as<uint32_t>(),
                                                            hist_scan.as<uint32_t>(), nb_sort);
            std::swap(ka, kb);
            std::swap(va, vb);
        }
        // 4 passes: the result is back in the A.keys / A.ids buffers
        const uint32_t n_int = n - 1;
        if (n >= 2) {
            k_karras<<<blocks(n_int, 256), 256, 0, st>>>(A.keys.as<uint32_t>(), (int)n, A.bin_child.as<uint32_t>(),
                                                         first.as<uint32_t>(), count.as<uint32_t>(),
                                                         pint.as<uint32_t>(), pleaf.as<uint32_t>());
        }
        k_leaf_boxes<<<blocks(n, 256), 256, 0, st>>>(A.world_tris.as<float4>(), A.ids.as<uint32_t>(), n, n_int,
                                                     A.bin_aabb.as<float>());
        if (n >= 2) {
            k_fill_u32<<<blocks(n_int, 256), 256, 0, st>>>(flags.as<uint32_t>(), n_int, 0u);
            k_fill_u32<<<1, 256, 0, st>>>(up_cnt.as<uint32_t>(), kMaxBinLevels + 1, 0u);
            BinUp U{n, n_int, A.bin_child.as<uint32_t>(), count.as<uint32_t>(), pint.as<uint32_t>(),
                    pleaf.as<uint32_t>(), flags.as<uint32_t>(), A.bin_aabb.as<float>(), cost.as<float>(),
                    height.as<uint32_t>(), A.bin_leafify.as<uint32_t>(), A.bin_dpc.as<float>(),
                    A.bin_dpk.as<uint32_t>()};
            // the live set shrinks level by level: smaller grids (grid-stride
            // loops, so any grid is correct) keep the late, nearly empty launches cheap
            auto grid = [&](uint32_t l) { return std::max(16u, std::min(blocks(n, kUpBlock) >> l, cus * 8u)); };
            uint32_t* live[2] = {up_live.as<uint32_t>(), up_live.as<uint32_t>() + n};
            uint32_t* cnt = up_cnt.as<uint32_t>();
            k_bottom_level<true><<<grid(0), kUpBlock, 0, st>>>(U, nullptr, nullptr, live[0], cnt);
            for (uint32_t l = 1; l < kMaxBinLevels; ++l)
                k_bottom_level<false><<<grid(l), kUpBlock, 0, st>>>(U, live[(l - 1) & 1], cnt + (l - 1),
                                                                    live[l & 1], cnt + l);
        }
        // ---- collapse, one level per iteration
        BinView bv{n, n_int, A.bin_child.as<uint32_t>(), first.as<uint32_t>(), count.as<uint32_t>(),
                   A.bin_aabb.as<float>(), A.bin_leafify.as<uint32_t>(), A.bin_dpk.as<uint32_t>()};
        uint32_t root = n >= 2 ? 0u : kLeafTag;
        HIP_CHECK(hipMemcpyAsync(work.ptr, &root, 4, hipMemcpyHostToDevice, st));
        uint32_t level_start = 0, level_count = 1, tri_start = 0, depth = 0;
        while (level_count > 0) {
            k_collapse_expand<<<blocks(level_count, 128), 128, 0, st>>>(
                bv, work.as<uint32_t>(), level_start, level_count, slot_child.as<uint32_t>(), nint.as<uint32_t>(),
                ntri.as<uint32_t>());
            scan_exclusive(nint.as<uint32_t>(), sint.as<uint32_t>(), level_count, sums.as<uint32_t>(),
                           totals.as<uint32_t>(), st);
            scan_exclusive(ntri.as<uint32_t>(), stri.as<uint32_t>(), level_count, sums.as<uint32_t>(),
                           totals.as<uint32_t>() + 1, st);
            uint32_t next_start = level_start + level_count;
            k_collapse_emit<<<blocks(level_count, 128), 128, 0, st>>>(
                bv, work.as<uint32_t>(), level_start, level_count, slot_child.as<uint32_t>(), sint.as<uint32_t>(),
                stri.as<uint32_t>(), next_start, tri_start, A.ids.as<uint32_t>(), A.world_tris.as<float4>(),
                A.nodes.as<uint4>(), A.leaf_tris.as<float4>());
            uint32_t tot[2];
            HIP_CHECK(hipMemcpyAsync(tot, totals.ptr, 8, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            ++depth;
            tri_start += tot[1];
            level_start = next_start;
            level_count = tot[0];
            if (level_start + level_count > n + 1) return fail(ctx, SNKPT_ERR_CAPACITY, "BVH8 node overflow");
        }
        A.num_nodes = level_start;
        A.depth = depth;
        if (tri_start != n) return fail(ctx, SNKPT_ERR_HIP, "BVH8 collapse lost triangles");
        HIP_CHECK(hipMemcpyAsync(hb, bounds.ptr, 48, hipMemcpyDeviceToHost, st));
        if (n >= 2) {
            uint32_t h0 = 0;
            HIP_CHECK(hipMemcpyAsync(&h0, height.ptr, 4, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            A.binary_depth = h0;
            if (h0 > kMaxBinLevels + 1) return fail(ctx, SNKPT_ERR_HIP, "binary tree deeper than the bottom-up pass");
        } else {
            A.binary_depth = 1;
        }
    }
    HIP_CHECK(hipEventRecord(ev1, st));
    HIP_CHECK(hipEventSynchronize(ev1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    A.build_ms = ms;
    if (n > 0) {
        auto dec = [](uint32_t u) {
            uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
            float f;
            std::memcpy(&f, &b, 4);
            return f;
        };
        for (int a = 0; a < 3; ++a) {
            A.lo[a] = dec(hb[6 + a]);
            A.hi[a] = dec(hb[9 + a]);
        }
    } else {
        for (int a = 0; a < 3; ++a) A.lo[a] = A.hi[a] = 0.0f;
    }
    A.built = true;
    return build_emitters(ctx, st);
}

snkpt_status accel_export(snkpt_ctx* ctx, uint32_t which, void* dst, size_t bytes, size_t* out_bytes) {
    const Accel& A = ctx->accel;
    if (!A.built) return fail(ctx, SNKPT_ERR_NOT_BUILT, "snkpt_accel_export: acceleration structure not built");
    const uint32_t n = A.num_tris, n_int = n ? n - 1 : 0;
    const void* src = nullptr;
    size_t sz = 0;
    switch (which) {
        case 0: src = A.keys.ptr, sz = (size_t)n * 4; break;
        case 1: src = A.ids.ptr, sz = (size_t)n * 4; break;
        case 2: src = A.bin_child.ptr, sz = (size_t)n_int * 8; break;
        case 3: src = A.bin_aabb.ptr, sz = (size_t)(n_int + n) * 24; break;
        case 4: src = A.nodes.ptr, sz = (size_t)A.num_nodes * 80; break;
        case 5: src = A.leaf_tris.ptr, sz = (size_t)n * 48; break;
        case 6: src = A.world_tris.ptr, sz = (size_t)n * 48; break;
        case 7: src = A.bin_leafify.ptr, sz = (size_t)n_int * 4; break;
        default: return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_accel_export: unknown array");
    }
    if (out_bytes) *out_bytes = sz;
    if (!dst) return SNKPT_OK;
    if (bytes < sz) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_accel_export: destination too small");
    if (sz == 0) return SNKPT_OK;
    hipError_t e = hipMemcpy(dst, src, sz, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy");
    return SNKPT_OK;
}

}  // namespace snkpt
