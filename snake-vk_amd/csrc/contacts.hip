// Short collision rays of the particle system: ParticleUpdate.rgen's
// TraceCollisionRay / ConstrainOffsetWithScene / CalculateGravity
// (ParticleUpdate.rgen:92-167), traced with gl_RayFlagsTerminateOnFirstHitEXT,
// and the contact record of ParticleUpdate.rchit (hit point, interpolated
// normal through transpose(inverse(mat3(M))), velocity of the hit surface from
// the previous frame's transform).
//
// The persistent any-hit traversal (refill_trace<ANY = true>) is reused as is;
// the contact record is computed in its per-ray completion handler, so hits
// never round-trip through HBM.  The first hit found is the same one the
// any-hit path of snkpt_trace reports (identical traversal order).
#include "render_common.h"

namespace snkpt {
namespace {

constexpr int kBlock = 128;

struct ContactScene {
    const InstD* inst;
    const float4* xf;       // per instance: 3x4 object->world rows, then 3 normal-matrix rows
    const float4* prev_xf;  // per instance: 3x4 object->world rows of the previous frame
    const uint32_t* idx;
    const float* pos;
    const float* nrm;
};

// Waves per SIMD the stacks' LDS leaves room for (8-byte entries, 128 lanes a block,
// 160 KiB a CU): 8 for 10 entries (16 blocks), 6 for 12 (13 blocks, 12 launched).  The
// register budget is held to that count: what spills is the contact record of the
// completion handler, run once per ray, not the traversal loop.
template <int STACK>
constexpr int contact_waves() { return STACK <= 10 ? 8 : STACK <= 12 ? 6 : 1; }

template <int STACK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(contact_waves<STACK>(), 8))) void k_contacts(const float4* __restrict__ org_tmin,
                                                     const float4* __restrict__ dir_tmax, uint32_t n,
                                                     const uint4* __restrict__ nodes, const float4* __restrict__ tris,
                                                     ContactScene C, float timestep, float* __restrict__ t_out,
                                                     float4* __restrict__ n_out, float4* __restrict__ p_out,
                                                     float4* __restrict__ v_out, uint32_t* __restrict__ inst_out,
                                                     uint32_t* __restrict__ prim_out, uint32_t refill) {
    __shared__ uint2 stk[STACK * kBlock];
    const uint32_t nwaves = gridDim.x * (kBlock / 64);
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    auto load = [&](uint32_t id, Ray& ray, const uint4*& nd, const float4*& tr) {
        const float4 a = org_tmin[id], d = dir_tmax[id];
        ray = Ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
        nd = nodes;
        tr = tris;
    };
    auto done = [&](uint32_t id, const TravState& st, uint32_t, uint32_t) {
        const TraceResult& res = st.res;
        if (res.rec == 0xFFFFFFFFu) {
            const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            t_out[id] = -1.0f;
            n_out[id] = z;
            p_out[id] = z;
            v_out[id] = z;
            if (inst_out) inst_out[id] = 0xFFFFFFFFu;
            if (prim_out) prim_out[id] = 0xFFFFFFFFu;
            return;
        }
        const float4 a = org_tmin[id], d = dir_tmax[id];
        const Ray ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
        const float4 r2 = tris[(size_t)res.rec * kTriWords + 2];
        const uint32_t ii = __float_as_uint(r2.y), prim = __float_as_uint(r2.z);
        float u, v;
        hit_uv(tris, res.rec, ray, u, v);
        // ParticleUpdate.rchit:40-62
        const InstD in = C.inst[ii];
        const uint32_t* ix = C.idx + in.index_offset + 3 * prim;
        const uint32_t i0 = ix[0] + in.vertex_offset, i1 = ix[1] + in.vertex_offset, i2 = ix[2] + in.vertex_offset;
        const float b0 = 1.0f - u - v;
        const V3 nl = (ld3(C.nrm + 3 * (size_t)i0) * b0 + ld3(C.nrm + 3 * (size_t)i1) * u) +
                      ld3(C.nrm + 3 * (size_t)i2) * v;
        const V3 pl = (ld3(C.pos + 3 * (size_t)i0) * b0 + ld3(C.pos + 3 * (size_t)i1) * u) +
                      ld3(C.pos + 3 * (size_t)i2) * v;
        const float4* x = C.xf + 6 * (size_t)ii;
        const float4 m0 = x[3], m1 = x[4], m2 = x[5];
        V3 nw{(m0.x * nl.x + m0.y * nl.y) + m0.z * nl.z, (m1.x * nl.x + m1.y * nl.y) + m1.z * nl.z,
              (m2.x * nl.x + m2.y * nl.y) + m2.z * nl.z};
        const float l2 = dot(nw, nw);
        nw = l2 > 0.0f ? nw * (1.0f / sqrtf(l2)) : nw;
        const float t = res.t;
        const V3 cp{ray.ox + ray.dx * t, ray.oy + ray.dy * t, ray.oz + ray.dz * t};
        const float4* q = C.prev_xf + 3 * (size_t)ii;
        const float4 q0 = q[0], q1 = q[1], q2 = q[2];
        const V3 pp{((q0.x * pl.x + q0.y * pl.y) + q0.z * pl.z) + q0.w, ((q1.x * pl.x + q1.y * pl.y) + q1.z * pl.z) + q1.w,
                    ((q2.x * pl.x + q2.y * pl.y) + q2.z * pl.z) + q2.w};
        t_out[id] = t;
        n_out[id] = make_float4(nw.x, nw.y, nw.z, 0.0f);
        p_out[id] = make_float4(cp.x, cp.y, cp.z, 0.0f);
        v_out[id] = make_float4((cp.x - pp.x) / timestep, (cp.y - pp.y) / timestep, (cp.z - pp.z) / timestep, 0.0f);
        if (inst_out) inst_out[id] = ii;
        if (prim_out) prim_out[id] = prim;
    };
    refill_trace_tuned<true, false, STACK, kStepsPerRefill>(n, wave, nwaves, &stk[threadIdx.x], kBlock, refill, load,
                                                             done);
}

__global__ void k_contacts_empty(uint32_t n, float* t_out, float4* n_out, float4* p_out, float4* v_out,
                                 uint32_t* inst_out, uint32_t* prim_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    t_out[i] = -1.0f;
    n_out[i] = z;
    p_out[i] = z;
    v_out[i] = z;
    if (inst_out) inst_out[i] = 0xFFFFFFFFu;
    if (prim_out) prim_out[i] = 0xFFFFFFFFu;
}

}  // namespace

snkpt_status trace_contacts(snkpt_ctx* ctx, const float4* org_tmin, const float4* dir_tmax, uint32_t n,
                            const snkpt_contacts& out, float timestep, hipStream_t st) {
    const Accel& A = ctx->accel;
    if (!A.built) return fail(ctx, SNKPT_ERR_NOT_BUILT, "trace_contacts: acceleration structure not built");
    if (n == 0) return SNKPT_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    float4* nrm = (float4*)out.normal;
    float4* pos = (float4*)out.position;
    float4* vel = (float4*)out.velocity;
    if (A.num_tris == 0) {
        k_contacts_empty<<<(n + 255) / 256, 256, 0, st>>>(n, out.t, nrm, pos, vel, out.inst, out.prim);
    } else {
        const uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
        if (need > 64) return fail(ctx, SNKPT_ERR_CAPACITY, "trace_contacts: BVH8 deeper than the 64-entry stack");
        const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
        // resident blocks per CU by the stacks' LDS, an even count (two waves per block over four SIMDs)
        const uint32_t per_cu = need <= 10 ? 16u : need <= 12 ? 12u : need <= 16 ? 10u : need <= 32 ? 4u : 2u;
        const uint32_t grid = std::min((n + kBlock - 1) / kBlock, cus * per_cu);
        // the any-hit refill threshold the build's autotuner picked
        const uint32_t refill = A.shape[1].refill ? A.shape[1].refill : 16u;
        const ContactScene C{ctx->d_inst.as<InstD>(), ctx->d_inst_xf.as<float4>(),
                             ctx->d_prev_xf.ptr ? ctx->d_prev_xf.as<float4>() : nullptr, ctx->d_idx.as<uint32_t>(),
                             ctx->d_pos.as<float>(), ctx->d_nrm.as<float>()};
        if (!C.prev_xf) return fail(ctx, SNKPT_ERR_NOT_BUILT, "trace_contacts: no previous-frame transforms");
        const uint4* nodes = A.nodes.as<uint4>();
        const float4* tris = A.leaf_tris.as<float4>();
#define SNKPT_CONTACTS(S)                                                                                     \
    k_contacts<S><<<grid, kBlock, 0, st>>>(org_tmin, dir_tmax, n, nodes, tris, C, timestep, out.t, nrm, pos, vel, \
                                           out.inst, out.prim, refill)
        if (need <= 10)
            SNKPT_CONTACTS(10);
        else if (need <= 12)
            SNKPT_CONTACTS(12);
        else if (need <= 16)
            SNKPT_CONTACTS(16);
        else if (need <= 32)
            SNKPT_CONTACTS(32);
        else
            SNKPT_CONTACTS(64);
#undef SNKPT_CONTACTS
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "trace_contacts launch");
    return SNKPT_OK;
}

}  // namespace snkpt
