// Device side of the wavefront path tracer shared by the single-scene and the
// batched passes: the shade step of one path vertex and the shadow step.
#pragma once

#include "render_modes.h"
#include "snkpt_dev.h"

namespace snkpt {

// Scene source of a launch with one scene: the stream id of a path's RNG is
// its path index (the pixel for camera paths).
struct SceneOne {
    SceneDev S;
    __device__ __forceinline__ const SceneDev& at(uint32_t) const { return S; }
    __device__ __forceinline__ uint32_t stream(uint32_t path) const { return path; }
};

// One path vertex of the unbiased path tracer (KIND: see ShadeHooks) for queue
// entry i.  Every thread of the block calls it (the queue appends are
// block-wide); `valid` marks threads with an entry.
template <int KIND, int BLOCK, class SS>
__device__ __forceinline__ void shade_step(const SS& ss, const PtParams& P, const Queue& in, const HitArrays& hit,
                                           const Queue& out, const ShadowQueue& sq, const PathState& ps,
                                           float4* __restrict__ rad, uint32_t* __restrict__ counts, uint32_t nxt,
                                           const ShadeHooks& H, uint32_t i, bool valid, uint32_t* lds_app) {
    bool want_ext = false, want_sh = false, want_q = false;
    float qs[13];
    V3 ext_o, ext_d, sh_o, sh_d, sh_c;
    float sh_t = 0.0f;
    uint32_t path = 0;
    float4 thr4 = make_float4(0, 0, 0, 0);
    Rng rng{0, 0};
    uint32_t bounce = 0;
    float new_pdf = 0.0f;
    if (valid) {
        path = in.path[i];
        const SceneDev& S = ss.at(path);
        thr4 = ps.thr[path];
        uint2 rs = ps.rng[path];
        rng.state = (uint64_t)rs.x | ((uint64_t)rs.y << 32);
        rng.inc = ((uint64_t)ss.stream(path) << 1u) | 1u;
        bounce = ps.bounce[path];
        float t = hit.t[i];
        if (t >= 0.0f) {
            float4 dq = in.dir[i];
            V3 dir{dq.x, dq.y, dq.z};
            V3 thr{thr4.x, thr4.y, thr4.z};
            Surf s = surface_at(S, dir, hit.inst[i], hit.prim[i], hit.u[i], hit.v[i]);
            if (s.emissive) {
                // MIS against the NEE of the previous vertex; a negative stored pdf
                // marks a vertex whose direct light came from ReSTIR (weight 0 here)
                float w = 1.0f;
                if (bounce > 0)
                    w = thr4.w < 0.0f ? 0.0f : power_heuristic(thr4.w, light_pdf(S, s.gid, s.ng, dir, t));
                V3 c = mul(thr, s.le) * w;
                float4 r = rad[path];
                rad[path] = make_float4(r.x + c.x, r.y + c.y, r.z + c.z, r.w);
            } else if (KIND == 1 && bounce == H.query_at && !H.indirect) {
                // terminate into the radiance cache: L_o(x, wo) from the network
                V3 wo = dir * -1.0f;
                want_q = true;
                float q[13] = {s.p.x, s.p.y, s.p.z, wo.x, wo.y, wo.z, s.ns.x, s.ns.y, s.ns.z, s.m.roughness,
                               s.m.albedo.x, s.m.albedo.y, s.m.albedo.z};
#pragma unroll
                for (int k = 0; k < 13; ++k) qs[k] = q[k];
            } else {
                V3 wo = dir * -1.0f;
                // indirect split: direct light here by NEE alone, the cache answers the rest
                const bool query = KIND == 1 && bounce == H.query_at;
                if (query) {
                    want_q = true;
                    float q[13] = {s.p.x, s.p.y, s.p.z, wo.x, wo.y, wo.z, s.ns.x, s.ns.y, s.ns.z, s.m.roughness,
                                   s.m.albedo.x, s.m.albedo.y, s.m.albedo.z};
#pragma unroll
                    for (int k = 0; k < 13; ++k) qs[k] = q[k];
                }
                if (KIND == 2 && bounce < H.max_rec) {
                    // training record: sample fields, radiance so far, throughput at arrival
                    float4 r = rad[path];
                    float* o = H.rec + ((size_t)path * H.max_rec + bounce) * kRecFloats;
                    float q[kRecFloats] = {s.p.x, s.p.y, s.p.z, wo.x, wo.y, wo.z, s.ns.x, s.ns.y, s.ns.z,
                                           s.m.roughness, s.m.albedo.x, s.m.albedo.y, s.m.albedo.z, r.x, r.y, r.z,
                                           thr.x, thr.y, thr.z, 0.0f};
#pragma unroll
                    for (int k = 0; k < kRecFloats; ++k) o[k] = q[k];
                    H.nrec[path] = bounce + 1;
                }
                float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
                float ul = rng.uniform(), us1 = rng.uniform(), us2 = rng.uniform(), urr = rng.uniform();
                const bool last = bounce + 1 >= P.max_depth || query;
                // next-event estimation
                LightSample ls = sample_light(S, s.p, u0, u1, u2, u3);
                if (ls.ok) {
                    float cs = dot(s.ns, ls.wi);
                    if (cs > 0.0f && dot(s.ng, ls.wi) > 0.0f) {
                        V3 f = brdf_eval(s.m, s.ns, wo, ls.wi);
                        float w = last ? 1.0f : power_heuristic(ls.pdf, brdf_pdf(s.m, s.ns, wo, ls.wi));
                        sh_c = mul(mul(thr, f), ls.le) * (cs * w / ls.pdf);
                        if (sh_c.x > 0.0f || sh_c.y > 0.0f || sh_c.z > 0.0f) {
                            want_sh = true;
                            sh_o = s.p + s.ng * s.eps;
                            sh_d = ls.wi;
                            sh_t = ls.dist * (1.0f - 1e-4f) - s.eps;
                        }
                    }
                }
                // continuation
                V3 wi;
                if (!last && brdf_sample(s.m, s.ns, wo, ul, us1, us2, wi) && dot(s.ng, wi) > 0.0f) {
                    float pdf = brdf_pdf(s.m, s.ns, wo, wi);
                    if (pdf > 0.0f) {
                        V3 f = brdf_eval(s.m, s.ns, wo, wi);
                        V3 nt = mul(thr, f) * (dot(s.ns, wi) / pdf);
                        bool alive = nt.x > 0.0f || nt.y > 0.0f || nt.z > 0.0f;
                        if (alive && P.rr_start > 0 && bounce + 1 >= P.rr_start) {
                            float q = fminf(fmaxf(fmaxf(nt.x, nt.y), nt.z), 0.95f);
                            if (urr >= q)
                                alive = false;
                            else
                                nt = nt * (1.0f / q);
                        }
                        if (alive) {
                            want_ext = true;
                            ext_o = s.p + s.ng * s.eps;
                            ext_d = wi;
                            thr4 = make_float4(nt.x, nt.y, nt.z, pdf);
                            new_pdf = pdf;
                        }
                    }
                }
            }
        }
    }
    uint32_t slot = block_append<BLOCK>(want_ext, &counts[nxt], lds_app);
    if (want_ext) {
        out.org[slot] = make_float4(ext_o.x, ext_o.y, ext_o.z, 0.0f);
        out.dir[slot] = make_float4(ext_d.x, ext_d.y, ext_d.z, 1e30f);
        out.path[slot] = path;
        ps.thr[path] = thr4;
        ps.rng[path] = make_uint2((uint32_t)rng.state, (uint32_t)(rng.state >> 32));
        ps.bounce[path] = bounce + 1;
    }
    (void)new_pdf;
    uint32_t sslot = block_append<BLOCK>(want_sh, &counts[2], lds_app);
    if (want_sh) {
        sq.org[sslot] = make_float4(sh_o.x, sh_o.y, sh_o.z, 0.0f);
        sq.dir[sslot] = make_float4(sh_d.x, sh_d.y, sh_d.z, sh_t);
        sq.contrib[sslot] = make_float4(sh_c.x, sh_c.y, sh_c.z, __uint_as_float(path));
    }
    if (KIND == 1) {
        uint32_t qslot = block_append<BLOCK>(want_q, H.q_count, lds_app);
        if (want_q) {
            float* o = H.q_samples + (size_t)qslot * 16;
#pragma unroll
            for (int k = 0; k < 13; ++k) o[k] = qs[k];
            o[13] = o[14] = o[15] = 0.0f;
            H.q_meta[qslot] = make_float4(thr4.x, thr4.y, thr4.z, __uint_as_float(path));
        }
    }
}

// Any-hit visibility of shadow-queue entry i; adds its contribution when unoccluded.
template <int STACK>
__device__ __forceinline__ void shadow_step(const uint4* __restrict__ nodes, const float4* __restrict__ tris,
                                            const ShadowQueue& sq, uint32_t i, float4* __restrict__ rad,
                                            uint2* stk, int stride) {
    float4 a = sq.org[i], d = sq.dir[i];
    Ray r{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
    uint32_t nv = 0, tv = 0;
    TraceResult res = traverse<true, false, STACK>(nodes, tris, r, stk, stride, nv, tv);
    if (res.rec == 0xFFFFFFFFu) {
        float4 c = sq.contrib[i];
        uint32_t px = __float_as_uint(c.w);
        float4 o = rad[px];
        rad[px] = make_float4(o.x + c.x, o.y + c.y, o.z + c.z, o.w);
    }
}

}  // namespace snkpt
