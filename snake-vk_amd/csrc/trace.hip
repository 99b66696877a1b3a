// BVH8 closest-hit and any-hit tracing of a ray stream.
//
// Replaces traceRayEXT with gl_RayFlagsOpaqueEXT (RayGen.rgen:164-176,
// RayClosestHit.rchit:87-91) and with TerminateOnFirstHit|Opaque|SkipClosestHit
// (ShadowTest, RayGen.rgen:101-121).  One lane per ray; the traversal stack
// lives in LDS (STACK entries of 8 B per lane, interleaved across the block so
// lanes at equal depth hit distinct banks).
#include <algorithm>

#include "snkpt_dev.h"
#include "snkpt_host.h"

namespace snkpt {
namespace {

constexpr int kTraceBlock = 128;

template <bool ANY, bool COUNT, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const float4* __restrict__ org_tmin,
                                                       const float4* __restrict__ dir_tmax, uint32_t n,
                                                       const uint32_t* __restrict__ count_dev,
                                                       const uint4* __restrict__ nodes,
                                                       const float4* __restrict__ tris, float* __restrict__ t_out,
                                                       uint32_t* __restrict__ inst_out,
                                                       uint32_t* __restrict__ prim_out, float* __restrict__ u_out,
                                                       float* __restrict__ v_out, uint32_t* __restrict__ nvis_out,
                                                       uint32_t* __restrict__ tvis_out) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    const uint32_t nn = count_dev ? *count_dev : n;
    if (i >= nn) return;
    float4 a = org_tmin[i], d = dir_tmax[i];
    Ray r{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
    uint32_t nv = 0, tv = 0;
    TraceResult res = traverse<ANY, COUNT, STACK>(nodes, tris, r, &stk[threadIdx.x], kTraceBlock, nv, tv);
    if (res.rec != 0xFFFFFFFFu) {
        float4 r2 = tris[(size_t)res.rec * kTriWords + 2];
        t_out[i] = res.t;
        if (inst_out) inst_out[i] = __float_as_uint(r2.y);
        if (prim_out) prim_out[i] = __float_as_uint(r2.z);
        if (u_out || v_out) {
            float u, v;
            hit_uv(tris, res.rec, r, u, v);
            if (u_out) u_out[i] = u;
            if (v_out) v_out[i] = v;
        }
    } else {
        t_out[i] = -1.0f;
        if (inst_out) inst_out[i] = 0xFFFFFFFFu;
        if (prim_out) prim_out[i] = 0xFFFFFFFFu;
        if (u_out) u_out[i] = 0.0f;
        if (v_out) v_out[i] = 0.0f;
    }
    if (COUNT) {
        if (nvis_out) nvis_out[i] = nv;
        if (tvis_out) tvis_out[i] = tv;
    }
}

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
template <bool ANY, bool COUNT, int STACK, int REFILL>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(STACK <= 10 ? 8 : 1, 8))) void k_trace_refill(const float4* __restrict__ org_tmin,
                                                              const float4* __restrict__ dir_tmax, uint32_t n,
                                                              const uint32_t* __restrict__ count_dev,
                                                              const uint4* __restrict__ nodes,
                                                              const float4* __restrict__ tris,
                                                              float* __restrict__ t_out, uint32_t* __restrict__ inst_out,
                                                              uint32_t* __restrict__ prim_out, float* __restrict__ u_out,
                                                              float* __restrict__ v_out, uint32_t* __restrict__ nvis_out,
                                                              uint32_t* __restrict__ tvis_out) {
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
    refill_trace<ANY, COUNT, STACK, REFILL>(nn, wave, nwaves, &stk[threadIdx.x], kTraceBlock, load, done);
}

// Idle lanes of a wave are refilled once this many are idle (4: 3.18 Grays/s on
// the 1.29M incoherent rays, 8: 3.23, 12: 3.24, 16: 3.20, 24: 3.15, 32: 3.09;
// coherent primary rays favour 8: 6.31 against 6.18 at 16).
constexpr int kRefill = 8;
// Resident 128-lane blocks per CU allowed by the 160 KB of LDS: a lane's stack
// takes STACK x 8 B (10: 16 blocks, 12: 13, 16: 10, 32: 5, 64: 2).  The 10-entry
// variant is held to 64 VGPRs so that 8 waves per SIMD fit (two spilled dwords;
// 6 waves per SIMD with 80 VGPRs: +3% on incoherent rays, -12% on primary rays).
inline uint32_t refill_blocks_per_cu(uint32_t need) {
    return need <= 10 ? 16u : need <= 12 ? 13u : need <= 16 ? 10u : need <= 32 ? 5u : 2u;
}

template <bool ANY, bool COUNT>
void launch_stack(uint32_t stack, uint32_t grid, uint32_t pgrid, hipStream_t st, const float4* o, const float4* d,
                  uint32_t n, const uint32_t* cnt, const uint4* nodes, const float4* tris, const snkpt_hits& h,
                  uint32_t* nv, uint32_t* tv) {
    // persistent refill launch, its grid capped by the (maximum) ray count
    const uint32_t g = std::min(pgrid, grid);
    if (stack <= 10)
        k_trace_refill<ANY, COUNT, 10, kRefill><<<g, kTraceBlock, 0, st>>>(o, d, n, cnt, nodes, tris, h.t, h.inst,
                                                                          h.prim, h.u, h.v, nv, tv);
    else if (stack <= 12)
        k_trace_refill<ANY, COUNT, 12, kRefill><<<g, kTraceBlock, 0, st>>>(o, d, n, cnt, nodes, tris, h.t, h.inst,
                                                                          h.prim, h.u, h.v, nv, tv);
    else if (stack <= 16)
        k_trace_refill<ANY, COUNT, 16, kRefill><<<g, kTraceBlock, 0, st>>>(o, d, n, cnt, nodes, tris, h.t, h.inst,
                                                                          h.prim, h.u, h.v, nv, tv);
    else if (stack <= 32)
        k_trace_refill<ANY, COUNT, 32, kRefill><<<g, kTraceBlock, 0, st>>>(o, d, n, cnt, nodes, tris, h.t, h.inst,
                                                                          h.prim, h.u, h.v, nv, tv);
    else
        k_trace_refill<ANY, COUNT, 64, kRefill><<<g, kTraceBlock, 0, st>>>(o, d, n, cnt, nodes, tris, h.t, h.inst,
                                                                          h.prim, h.u, h.v, nv, tv);
}

}  // namespace

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
        const uint32_t pgrid = cus * refill_blocks_per_cu(need);
        const uint4* nodes = A.nodes.as<uint4>();
        const float4* tris = A.leaf_tris.as<float4>();
        if (any_hit) {
            if (count)
                launch_stack<true, true>(need, grid, pgrid, st, org_tmin, dir_tmax, n, count_dev, nodes, tris, hits,
                                         node_visits, tri_visits);
            else
                launch_stack<true, false>(need, grid, pgrid, st, org_tmin, dir_tmax, n, count_dev, nodes, tris, hits,
                                          nullptr, nullptr);
        } else {
            if (count)
                launch_stack<false, true>(need, grid, pgrid, st, org_tmin, dir_tmax, n, count_dev, nodes, tris, hits,
                                          node_visits, tri_visits);
            else
                launch_stack<false, false>(need, grid, pgrid, st, org_tmin, dir_tmax, n, count_dev, nodes, tris, hits,
                                           nullptr, nullptr);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    return SNKPT_OK;
}

}  // namespace snkpt
