// BVH8 closest-hit and any-hit tracing of a ray stream.
//
// Replaces traceRayEXT with gl_RayFlagsOpaqueEXT (RayGen.rgen:164-176,
// RayClosestHit.rchit:87-91) and with TerminateOnFirstHit|Opaque|SkipClosestHit
// (ShadowTest, RayGen.rgen:101-121).  Persistent waves refill their idle lanes
// with rays; the traversal stack lives in LDS (STACK entries of 8 B per lane,
// interleaved across the block so lanes at equal depth hit distinct banks).
// The launch shape (refill threshold, steps per refill check, resident blocks
// per CU) is picked per acceleration structure by an autotuner at build time.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "shading.h"
#include "snkpt_dev.h"
#include "snkpt_host.h"

namespace snkpt {
namespace {

constexpr int kTraceBlock = 128;

__global__ void k_trace_empty(uint32_t n, const uint32_t* __restrict__ count_dev, float* __restrict__ t_out,
                              uint32_t* __restrict__ inst_out, uint32_t* __restrict__ prim_out,
                              float* __restrict__ u_out, float* __restrict__ v_out, uint32_t* __restrict__ nvis_out,
                              uint32_t* __restrict__ tvis_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nn = count_dev ? *count_dev : n;
    if (i >= nn) return;
    t_out[i] = -1.0f;
    if (inst_out) inst_out[i] = 0xFFFFFFFFu;
    if (prim_out) prim_out[i] = 0xFFFFFFFFu;
    if (u_out) u_out[i] = 0.0f;
    if (v_out) v_out[i] = 0.0f;
    if (nvis_out) nvis_out[i] = 0;
    if (tvis_out) tvis_out[i] = 0;
}

// Persistent closest / any hit over the ray stream: each wave owns one 64-ray
// chunk per band of the rays plus an even slice of the last band, and refills
// its idle lanes from them (see refill_trace).  No atomics: a wave's cursor is
// uniform and private to it.
template <bool ANY, bool COUNT, int STACK, int STEPS>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(STACK <= 10 ? 8 : 1, 8))) void k_trace_refill(const float4* __restrict__ org_tmin,
                                                              const float4* __restrict__ dir_tmax, uint32_t n,
                                                              const uint32_t* __restrict__ count_dev,
                                                              const uint4* __restrict__ nodes,
                                                              const float4* __restrict__ tris,
                                                              float* __restrict__ t_out, uint32_t* __restrict__ inst_out,
                                                              uint32_t* __restrict__ prim_out, float* __restrict__ u_out,
                                                              float* __restrict__ v_out, uint32_t* __restrict__ nvis_out,
                                                              uint32_t* __restrict__ tvis_out, uint32_t refill) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t nn = count_dev ? *count_dev : n;
    const uint32_t nwaves = gridDim.x * (kTraceBlock / 64);
    const uint32_t wave = blockIdx.x * (kTraceBlock / 64) + (threadIdx.x >> 6);
    auto load = [&](uint32_t id, Ray& ray, const uint4*& nd, const float4*& tr) {
        float4 a = org_tmin[id], d = dir_tmax[id];
        ray = Ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
        nd = nodes;
        tr = tris;
    };
    auto done = [&](uint32_t id, const TravState& st, uint32_t nv, uint32_t tv) {
        const TraceResult& res = st.res;
        if (res.rec != 0xFFFFFFFFu) {
            float4 r2 = tris[(size_t)res.rec * kTriWords + 2];
            t_out[id] = res.t;
            if (inst_out) inst_out[id] = __float_as_uint(r2.y);
            if (prim_out) prim_out[id] = __float_as_uint(r2.z);
            if (u_out || v_out) {
                float4 a = org_tmin[id], d = dir_tmax[id];
                Ray ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
                float u, v;
                hit_uv(tris, res.rec, ray, u, v);
                if (u_out) u_out[id] = u;
                if (v_out) v_out[id] = v;
            }
        } else {
            t_out[id] = -1.0f;
            if (inst_out) inst_out[id] = 0xFFFFFFFFu;
            if (prim_out) prim_out[id] = 0xFFFFFFFFu;
            if (u_out) u_out[id] = 0.0f;
            if (v_out) v_out[id] = 0.0f;
        }
        if (COUNT) {
            if (nvis_out) nvis_out[id] = nv;
            if (tvis_out) tvis_out[id] = tv;
        }
    };
    refill_trace_tuned<ANY, COUNT, STACK, STEPS>(nn, wave, nwaves, &stk[threadIdx.x], kTraceBlock, refill, load, done);
}

// Hand-swept defaults (C2, 1.02M triangles): refill once 8 lanes are idle (4:
// 3.18 Grays/s on the 1.29M incoherent rays, 8: 3.23, 12: 3.24, 16: 3.20, 24:
// 3.15, 32: 3.09; coherent primary rays favour 8: 6.31 against 6.18 at 16), 4
// steps between refill checks, all resident blocks.
constexpr uint32_t kRefill = 8;
// Resident 128-lane blocks per CU allowed by the 160 KB of LDS: a lane's stack
// takes STACK x 8 B (10: 16 blocks, 12: 13, 16: 10, 32: 5, 64: 2).  The 10-entry
// variant is held to 64 VGPRs so that 8 waves per SIMD fit (6 waves per SIMD
// with 80 VGPRs: +3% on incoherent rays, -12% on primary rays).
inline uint32_t refill_blocks_per_cu(uint32_t need) {
    return need <= 10 ? 16u : need <= 12 ? 13u : need <= 16 ? 10u : need <= 32 ? 5u : 2u;
}
// Blocks per CU that give every SIMD the same number of waves (two waves per
// block over four SIMDs: an even count): 13 blocks put 7 waves on two SIMDs and
// 6 on the others, and ran the 12-entry variant 18 % slower than 12 blocks.
inline uint32_t balanced_blocks(uint32_t b) { return b > 1u ? (b & ~1u) : 1u; }
constexpr uint32_t kStepsDefault = 4;
TraceShape default_shape(uint32_t need) {
    return TraceShape{kRefill, kStepsDefault, balanced_blocks(refill_blocks_per_cu(need)), 0.0f};
}

template <bool ANY, bool COUNT, int STEPS>
void launch_stack_steps(uint32_t stack, uint32_t g, uint32_t refill, hipStream_t st, const float4* o,
                        const float4* d, uint32_t n, const uint32_t* cnt, const uint4* nodes, const float4* tris,
                        const snkpt_hits& h, uint32_t* nv, uint32_t* tv) {
#define K_LAUNCH(S)                                                                                          \
    k_trace_refill<ANY, COUNT, S, STEPS><<<g, kTraceBlock, 0, st>>>(o, d, n, cnt, nodes, tris, h.t, h.inst, \
                                                                     h.prim, h.u, h.v, nv, tv, refill)
    if (stack <= 10)
        K_LAUNCH(10);
    else if (stack <= 12)
        K_LAUNCH(12);
    else if (stack <= 16)
        K_LAUNCH(16);
    else if (stack <= 32)
        K_LAUNCH(32);
    else
        K_LAUNCH(64);
#undef K_LAUNCH
}

// Persistent refill launch of the tuned shape (kind 0 closest hit, 1 any hit),
// its grid capped by the (maximum) ray count.
template <bool ANY, bool COUNT>
void launch_stack(const TraceShape& sh, uint32_t stack, uint32_t grid, uint32_t cus, hipStream_t st,
                  const float4* o, const float4* d, uint32_t n, const uint32_t* cnt, const uint4* nodes,
                  const float4* tris, const snkpt_hits& h, uint32_t* nv, uint32_t* tv) {
    const uint32_t g = std::min(cus * std::max(1u, sh.blocks_per_cu), grid);
    if (sh.steps == 3)
        launch_stack_steps<ANY, COUNT, 3>(stack, g, sh.refill, st, o, d, n, cnt, nodes, tris, h, nv, tv);
    else
        launch_stack_steps<ANY, COUNT, 4>(stack, g, sh.refill, st, o, d, n, cnt, nodes, tris, h, nv, tv);
}

// Rays of the autotuner's sample, like a path tracer's bounce rays: ray i
// leaves leaf record k = i ntris / n (leaf records are in BVH order, so
// consecutive rays start on neighbouring surfaces, as the bounce rays of
// neighbouring pixels do; rays from scattered triangles favoured shapes with
// fewer resident blocks that were 3 % slower on real bounce rays), from a
// uniform point on it lifted 1e-4 of the scene extent along its normal, on a
// random side, cosine-distributed around that normal (PCG32 per ray).
__global__ void k_tune_rays(const float4* __restrict__ tris, uint32_t ntris, float4* __restrict__ org,
                            float4* __restrict__ dir, uint32_t n, float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng r = rng_seed(0x74756E65ull, i);
    const uint32_t k = min((uint32_t)(((uint64_t)i * ntris) / n), ntris - 1u);
    const float4 a = tris[(size_t)k * kTriWords], b = tris[(size_t)k * kTriWords + 1],
                 c = tris[(size_t)k * kTriWords + 2];
    const V3 v0{a.x, a.y, a.z}, v1{a.w, b.x, b.y}, v2{b.z, b.w, c.x};
    const float su = sqrtf(r.uniform()), u2 = r.uniform();
    const V3 p = v0 * (1.0f - su) + v1 * (su * (1.0f - u2)) + v2 * (su * u2);
    V3 nn = normalize(cross(v1 - v0, v2 - v0));
    if (r.uniform() < 0.5f) nn = nn * -1.0f;
    V3 t, bt;
    onb(nn, t, bt);
    const float r1 = r.uniform(), r2 = r.uniform(), rr = sqrtf(r1), ph = 6.2831853f * r2;
    const V3 d = t * (rr * cosf(ph)) + bt * (rr * sinf(ph)) + nn * sqrtf(fmaxf(0.0f, 1.0f - r1));
    const V3 o = p + nn * eps;
    org[i] = make_float4(o.x, o.y, o.z, 0.0f);
    dir[i] = make_float4(d.x, d.y, d.z, 1e30f);
}

// Tuned shapes per (device, traversal stack variant, log2 triangle count).
std::mutex g_tune_mu;
std::map<std::tuple<int, uint32_t, uint32_t>, std::pair<TraceShape, TraceShape>> g_tuned;

}  // namespace

// Picks the launch shape of the closest- and any-hit traversal for the built
// acceleration structure: refill threshold 4 / 8 / 16 idle lanes x 3 / 4 steps
// per refill check x every even block count per CU from the resident maximum
// down to half of it (C2 bounce rays, 12-entry stack: 12 blocks 5.03 Grays/s,
// 10: 4.48, 8: 4.87; 13: 4.15, 11: 4.74, 9: 4.77), each timed on 2^20
// bounce-like rays leaving the scene's triangles (best of 3 launches after one
// warm-up).  The
// shape changes which lane traces which ray, never a ray's traversal, so hits do
// not depend on it.  Results are kept per device, stack variant and scene size
// (log2 of the triangle count): later builds of similar scenes reuse them.
snkpt_status autotune_trace(snkpt_ctx* ctx, hipStream_t st) {
    Accel& A = ctx->accel;
    if (!A.built || A.shape_fixed || A.num_tris == 0) return SNKPT_OK;
    const uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
    if (need > 64) return SNKPT_OK;
    const uint32_t stack = need <= 10 ? 10u : need <= 12 ? 12u : need <= 16 ? 16u : need <= 32 ? 32u : 64u;
    uint32_t lg = 0;
    while ((1u << lg) < A.num_tris && lg < 31) ++lg;
    const auto key = std::make_tuple(ctx->device, stack, lg);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            A.shape[0] = it->second.first;
            A.shape[1] = it->second.second;
            return SNKPT_OK;
        }
    }
    const uint32_t n = 1u << 20;
    DeviceBuffer org, dir, t, ids;
    hipError_t he;
    if ((he = org.reserve((size_t)n * 16)) != hipSuccess || (he = dir.reserve((size_t)n * 16)) != hipSuccess ||
        (he = t.reserve((size_t)n * 4)) != hipSuccess || (he = ids.reserve((size_t)n * 8)) != hipSuccess)
        return hip_fail(ctx, he, "autotune: buffers");
    const float ext = std::max(std::max(A.hi[0] - A.lo[0], A.hi[1] - A.lo[1]), A.hi[2] - A.lo[2]);
    k_tune_rays<<<(n + 255) / 256, 256, 0, st>>>(A.leaf_tris.as<float4>(), A.num_tris, org.as<float4>(),
                                                 dir.as<float4>(), n, 1e-4f * ext);
    hipEvent_t e0, e1;
    if ((he = hipEventCreate(&e0)) != hipSuccess) return hip_fail(ctx, he, "autotune: event");
    if ((he = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ctx, he, "autotune: event");
    }
    const snkpt_hits h{t.as<float>(), ids.as<uint32_t>(), ids.as<uint32_t>() + n, nullptr, nullptr};
    const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
    const uint32_t bmax = refill_blocks_per_cu(need);
    TraceShape best[2];
    snkpt_status status = SNKPT_OK;
    for (int kind = 0; kind < 2 && status == SNKPT_OK; ++kind) {
        best[kind] = default_shape(need);
        best[kind].ms = 1e30f;
        for (uint32_t refill : {4u, 8u, 16u})
            for (uint32_t steps : {3u, 4u})
                for (uint32_t bpc = balanced_blocks(bmax); bpc >= std::max(1u, bmax / 2) && bpc >= 1u;
                     bpc = bpc > 2u ? bpc - 2u : 0u) {
                    const TraceShape sh{refill, steps, bpc, 0.0f};
                    float ms = 1e30f;
                    for (int rep = 0; rep < 4; ++rep) {
                        (void)hipEventRecord(e0, st);
                        if (kind)
                            launch_stack<true, false>(sh, need, (n + kTraceBlock - 1) / kTraceBlock, cus, st,
                                                      org.as<float4>(), dir.as<float4>(), n, nullptr,
                                                      A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), h, nullptr,
                                                      nullptr);
                        else
                            launch_stack<false, false>(sh, need, (n + kTraceBlock - 1) / kTraceBlock, cus, st,
                                                       org.as<float4>(), dir.as<float4>(), n, nullptr,
                                                       A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), h, nullptr,
                                                       nullptr);
                        (void)hipEventRecord(e1, st);
                        if ((he = hipEventSynchronize(e1)) != hipSuccess) {
                            status = hip_fail(ctx, he, "autotune: launch");
                            break;
                        }
                        float t_ms = 0.0f;
                        (void)hipEventElapsedTime(&t_ms, e0, e1);
                        if (rep > 0) ms = std::min(ms, t_ms);  // the first launch warms the caches
                    }
                    if (status != SNKPT_OK) break;
                    if (ms < best[kind].ms) {
                        best[kind] = sh;
                        best[kind].ms = ms;
                    }
                }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (status != SNKPT_OK) return status;
    A.shape[0] = best[0];
    A.shape[1] = best[1];
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[key] = std::make_pair(best[0], best[1]);
    return SNKPT_OK;
}

snkpt_status launch_trace(snkpt_ctx* ctx, const float4* org_tmin, const float4* dir_tmax, uint32_t n,
                          const snkpt_hits& hits, bool any_hit, uint32_t* node_visits, uint32_t* tri_visits,
                          const uint32_t* count_dev, hipStream_t st) {
    const Accel& A = ctx->accel;
    if (!A.built) return fail(ctx, SNKPT_ERR_NOT_BUILT, "trace: acceleration structure not built");
    if (n == 0) return SNKPT_OK;
    if (!org_tmin || !dir_tmax || !hits.t) return fail(ctx, SNKPT_ERR_INVALID_ARG, "trace: null ray or hit stream");
    uint32_t grid = (n + kTraceBlock - 1) / kTraceBlock;
    if (A.num_tris == 0) {
        k_trace_empty<<<(n + 255) / 256, 256, 0, st>>>(n, count_dev, hits.t, hits.inst, hits.prim, hits.u, hits.v,
                                                       node_visits, tri_visits);
    } else {
        uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
        if (need > 64) return fail(ctx, SNKPT_ERR_CAPACITY, "trace: BVH8 deeper than the 64-entry traversal stack");
        bool count = node_visits || tri_visits;
        const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
        const TraceShape sh = A.shape[any_hit ? 1 : 0].refill ? A.shape[any_hit ? 1 : 0] : default_shape(need);
        const uint4* nodes = A.nodes.as<uint4>();
        const float4* tris = A.leaf_tris.as<float4>();
        if (any_hit) {
            if (count)
                launch_stack<true, true>(sh, need, grid, cus, st, org_tmin, dir_tmax, n, count_dev, nodes, tris, hits,
                                         node_visits, tri_visits);
            else
                launch_stack<true, false>(sh, need, grid, cus, st, org_tmin, dir_tmax, n, count_dev, nodes, tris,
                                          hits, nullptr, nullptr);
        } else {
            if (count)
                launch_stack<false, true>(sh, need, grid, cus, st, org_tmin, dir_tmax, n, count_dev, nodes, tris,
                                          hits, node_visits, tri_visits);
            else
                launch_stack<false, false>(sh, need, grid, cus, st, org_tmin, dir_tmax, n, count_dev, nodes, tris,
                                           hits, nullptr, nullptr);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    return SNKPT_OK;
}

}  // namespace snkpt
