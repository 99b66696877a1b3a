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

// What one path vertex produced.
struct VertexOut {
    bool ext = false, sh = false, q = false;
    bool emit = false;  // c_emit is added to the path's radiance now
    V3 c_emit, ext_o, ext_d, sh_o, sh_d, sh_c;
    float sh_t = 0.0f;
    float4 thr4;        // throughput | pdf after the vertex (unchanged when the path ends here)
    float4 qthr;        // throughput arriving at a cache query vertex
    float qs[13];
};

// The unbiased path tracer at one hit (KIND: see ShadeHooks): emission under
// MIS, cache query or training record, NEE (power-heuristic MIS against BRDF
// sampling), BRDF continuation with Russian roulette.  Eight uniforms are
// drawn at every non-emissive vertex whether used or not, so the wavefront
// and the per-thread kernels consume the stream identically.  rad_now: the
// path's radiance before this vertex (read for training records only).
template <int KIND>
__device__ __forceinline__ void path_vertex_surf(const SceneDev& S, const PtParams& P, const ShadeHooks& H,
                                                 uint32_t path, uint32_t bounce, V3 dir, float t, const Surf& s,
                                                 float4 thr4, Rng& rng, float4 rad_now, VertexOut& o) {
    o.thr4 = thr4;
    const V3 thr{thr4.x, thr4.y, thr4.z};
    if (s.emissive) {
        // MIS against the NEE of the previous vertex; a negative stored pdf
        // marks a vertex whose direct light came from ReSTIR (weight 0 here)
        float w = 1.0f;
        if (bounce > 0) w = thr4.w < 0.0f ? 0.0f : power_heuristic(thr4.w, light_pdf(S, s.gid, s.ng, dir, t));
        o.c_emit = mul(thr, s.le) * w;
        o.emit = true;
        return;
    }
    // the BRDF probe of an indirect query vertex (H.mis) only looks for emission
    if (KIND == 1 && H.mis && bounce == H.query_at + 1) return;
    const V3 wo = dir * -1.0f;
    const bool query = KIND == 1 && bounce == H.query_at;
    if (query) {
        o.q = true;
        o.qthr = thr4;
        float q[13] = {s.p.x, s.p.y, s.p.z, wo.x, wo.y, wo.z, s.ns.x, s.ns.y, s.ns.z, s.m.roughness,
                       s.m.albedo.x, s.m.albedo.y, s.m.albedo.z};
#pragma unroll
        for (int k = 0; k < 13; ++k) o.qs[k] = q[k];
        // terminate into the radiance cache: L_o(x, wo) from the network
        if (!H.indirect) return;
        // indirect split: direct light here by NEE alone (or, with H.mis, NEE and a
        // traced BRDF sample under the power heuristic), the cache answers the rest
    }
    if (KIND == 2 && bounce < H.max_rec) {
        // training record: sample fields, radiance so far, throughput at arrival
        float* r = H.rec + ((size_t)path * H.max_rec + bounce) * kRecFloats;
        float q[kRecFloats] = {s.p.x, s.p.y, s.p.z, wo.x, wo.y, wo.z, s.ns.x, s.ns.y, s.ns.z,
                               s.m.roughness, s.m.albedo.x, s.m.albedo.y, s.m.albedo.z, rad_now.x, rad_now.y,
                               rad_now.z, thr.x, thr.y, thr.z, 0.0f};
#pragma unroll
        for (int k = 0; k < kRecFloats; ++k) r[k] = q[k];
        H.nrec[path] = bounce + 1;
    }
    float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    float ul = rng.uniform(), us1 = rng.uniform(), us2 = rng.uniform(), urr = rng.uniform();
    // an indirect query vertex with H.mis continues one BRDF sample as the probe
    const bool probe = KIND == 1 && query && H.mis;
    const bool last = (bounce + 1 >= P.max_depth || query) && !probe;
    // next-event estimation
    LightSample ls = sample_direct(S, s.p, u0, u1, u2, u3);
    if (ls.ok) {
        float cs = dot(s.ns, ls.wi);
        if (cs > 0.0f && dot(s.ng, ls.wi) > 0.0f) {
            V3 f = brdf_eval(s.m, s.ns, wo, ls.wi);
            float w = (last || ls.delta) ? 1.0f : power_heuristic(ls.pdf, brdf_pdf(s.m, s.ns, wo, ls.wi));
            o.sh_c = mul(mul(thr, f), ls.le) * (cs * w / ls.pdf);
            if (o.sh_c.x > 0.0f || o.sh_c.y > 0.0f || o.sh_c.z > 0.0f) {
                o.sh = true;
                shadow_segment(s.p, s.ng, s.eps, ls.wi, ls.dist, o.sh_o, o.sh_d, o.sh_t);
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
                o.ext = true;
                o.ext_o = s.p + s.ng * s.eps;
                o.ext_d = wi;
                o.thr4 = make_float4(nt.x, nt.y, nt.z, pdf);
            }
        }
    }
}

template <int KIND>
__device__ __forceinline__ void path_vertex(const SceneDev& S, const PtParams& P, const ShadeHooks& H, uint32_t path,
                                            uint32_t bounce, V3 dir, float t, uint32_t inst, uint32_t prim, float hu,
                                            float hv, float4 thr4, Rng& rng, float4 rad_now, VertexOut& o) {
    const Surf s = surface_at(S, dir, inst, prim, hu, hv);
    path_vertex_surf<KIND>(S, P, H, path, bounce, dir, t, s, thr4, rng, rad_now, o);
}

// Queue appends of one path vertex: continuation ray, shadow ray and (KIND 1)
// cache query.  Every thread of the block calls it (block-wide appends).
template <int KIND, int BLOCK>
__device__ __forceinline__ void vertex_appends(const VertexOut& o, uint32_t path, uint32_t bounce, const Rng& rng,
                                               const Queue& out, const ShadowQueue& sq, const PathState& ps,
                                               uint32_t* __restrict__ counts, uint32_t nxt, const ShadeHooks& H,
                                               uint32_t* lds_app) {
    uint32_t slot = block_append<BLOCK>(o.ext, &counts[nxt], lds_app);
    if (o.ext) {
        out.org[slot] = make_float4(o.ext_o.x, o.ext_o.y, o.ext_o.z, 0.0f);
        out.dir[slot] = make_float4(o.ext_d.x, o.ext_d.y, o.ext_d.z, 1e30f);
        out.path[slot] = path;
        ps.thr[path] = o.thr4;
        ps.rng[path] = make_uint2((uint32_t)rng.state, (uint32_t)(rng.state >> 32));
        ps.bounce[path] = bounce + 1;
    }
    uint32_t sslot = block_append<BLOCK>(o.sh, &counts[2], lds_app);
    if (o.sh) {
        sq.org[sslot] = make_float4(o.sh_o.x, o.sh_o.y, o.sh_o.z, 0.0f);
        sq.dir[sslot] = make_float4(o.sh_d.x, o.sh_d.y, o.sh_d.z, o.sh_t);
        sq.contrib[sslot] = make_float4(o.sh_c.x, o.sh_c.y, o.sh_c.z, __uint_as_float(path));
    }
    if (KIND == 1) {
        uint32_t qslot = block_append<BLOCK>(o.q, H.q_count, lds_app);
        if (o.q) {
            float* q = H.q_samples + (size_t)qslot * 16;
#pragma unroll
            for (int k = 0; k < 13; ++k) q[k] = o.qs[k];
            q[13] = q[14] = q[15] = 0.0f;
            H.q_meta[qslot] = make_float4(o.qthr.x, o.qthr.y, o.qthr.z, __uint_as_float(path));
        }
    }
}

// One path vertex of the wavefront path tracer for queue entry i.  Every
// thread of the block calls it (the queue appends are block-wide); `valid`
// marks threads with an entry.
template <int KIND, int BLOCK, class SS>
__device__ __forceinline__ void shade_step(const SS& ss, const PtParams& P, const Queue& in, const HitArrays& hit,
                                           const Queue& out, const ShadowQueue& sq, const PathState& ps,
                                           float4* __restrict__ rad, uint32_t* __restrict__ counts, uint32_t nxt,
                                           const ShadeHooks& H, uint32_t i, bool valid, uint32_t* lds_app) {
    VertexOut o;
    uint32_t path = 0, bounce = 0;
    Rng rng{0, 0};
    if (valid) {
        path = in.path[i];
        const SceneDev& S = ss.at(path);
        const float4 thr4 = ps.thr[path];
        uint2 rs = ps.rng[path];
        rng.state = (uint64_t)rs.x | ((uint64_t)rs.y << 32);
        rng.inc = ((uint64_t)ss.stream(path) << 1u) | 1u;
        bounce = ps.bounce[path];
        const float t = hit.t[i];
        if (t >= 0.0f) {
            const float4 dq = in.dir[i];
            const float4 rn = KIND == 2 ? rad[path] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            path_vertex<KIND>(S, P, H, path, bounce, V3{dq.x, dq.y, dq.z}, t, hit.inst[i], hit.prim[i], hit.u[i],
                              hit.v[i], thr4, rng, rn, o);
            if (o.emit) {
                float4 r = rad[path];
                rad[path] = make_float4(r.x + o.c_emit.x, r.y + o.c_emit.y, r.z + o.c_emit.z, r.w);
            }
        }
    }
    vertex_appends<KIND, BLOCK>(o, path, bounce, rng, out, sq, ps, counts, nxt, H, lds_app);
}

// The rest of one path, traced in place: bounces bounce0..max_depth-1 from the
// ray r of bounce bounce0, with the path's throughput and random stream at that
// bounce.  Every contribution is added to L in the wavefront's order (emission,
// then the unoccluded shadow ray, vertex by vertex).  n_ext counts the rays after
// the first, n_sh the shadow rays.
template <int KIND, int STACK>
__device__ __forceinline__ float4 path_rest(const SceneDev& S, const PtParams& P, const ShadeHooks& H, uint32_t path,
                                            uint32_t bounce0, Ray r, float4 thr4, Rng rng, uint2* stk, int sstride,
                                            float4 L, uint32_t& n_ext, uint32_t& n_sh) {
    uint32_t nv = 0, tv = 0;
    for (uint32_t bounce = bounce0; bounce < P.max_depth; ++bounce) {
        if (bounce > bounce0) ++n_ext;
        TraceResult res = traverse<false, false, STACK>(S.nodes, S.leaf, r, stk, sstride, nv, tv);
        if (res.rec == 0xFFFFFFFFu) break;
        const float4 r2 = S.leaf[(size_t)res.rec * kTriWords + 2];
        float hu, hv;
        hit_uv(S.leaf, res.rec, r, hu, hv);
        VertexOut o;
        path_vertex<KIND>(S, P, H, path, bounce, V3{r.dx, r.dy, r.dz}, res.t, __float_as_uint(r2.y),
                          __float_as_uint(r2.z), hu, hv, thr4, rng, L, o);
        if (o.emit) L = make_float4(L.x + o.c_emit.x, L.y + o.c_emit.y, L.z + o.c_emit.z, L.w);
        if (o.sh) {
            ++n_sh;
            Ray sr{o.sh_o.x, o.sh_o.y, o.sh_o.z, 0.0f, o.sh_d.x, o.sh_d.y, o.sh_d.z, o.sh_t};
            TraceResult vis = traverse<true, false, STACK>(S.nodes, S.leaf, sr, stk, sstride, nv, tv);
            if (vis.rec == 0xFFFFFFFFu) L = make_float4(L.x + o.sh_c.x, L.y + o.sh_c.y, L.z + o.sh_c.z, L.w);
        }
        if (!o.ext) break;
        r = Ray{o.ext_o.x, o.ext_o.y, o.ext_o.z, 0.0f, o.ext_d.x, o.ext_d.y, o.ext_d.z, 1e30f};
        thr4 = o.thr4;
    }
    return L;
}

// The late bounces, one thread per path: every path queued in `in` (n of them)
// runs to its end in place, from its stored throughput, random stream and
// bounce.  After Russian roulette has thinned the queue, a launch per bounce
// leaves most lanes idle; here a lane that ends its path takes the next one.
// path_rest adds the contributions in the wavefront's order, so the radiance is
// bit-identical.  The first ray of each entry was counted when it was queued;
// stats[1] / stats[2] get the rest (red: two words of LDS).
template <int STACK, int BLOCK, class SS>
__device__ __forceinline__ void tail_paths(const SS& ss, const PtParams& P, const ShadeHooks& H, const Queue& in,
                                           uint32_t n, const PathState& ps, float4* __restrict__ rad, uint2* stk,
                                           unsigned long long* __restrict__ stats, uint32_t* red) {
    if (threadIdx.x == 0) red[0] = red[1] = 0u;
    __syncthreads();
    uint32_t n_ext = 0, n_sh = 0;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const uint32_t path = in.path[i];
        const float4 o = in.org[i], d = in.dir[i];
        const uint2 rs = ps.rng[path];
        const Rng rng{(uint64_t)rs.x | ((uint64_t)rs.y << 32), ((uint64_t)ss.stream(path) << 1u) | 1u};
        rad[path] = path_rest<0, STACK>(ss.at(path), P, H, path, ps.bounce[path],
                                        Ray{o.x, o.y, o.z, o.w, d.x, d.y, d.z, d.w}, ps.thr[path], rng, stk, BLOCK,
                                        rad[path], n_ext, n_sh);
    }
    for (int o = 32; o > 0; o >>= 1) {
        n_ext += __shfl_xor(n_ext, o, 64);
        n_sh += __shfl_xor(n_sh, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&red[0], n_ext);
        atomicAdd(&red[1], n_sh);
    }
    __syncthreads();
    if (threadIdx.x == 0 && (red[0] | red[1])) {
        atomicAdd(&stats[1], (unsigned long long)red[0]);
        atomicAdd(&stats[2], (unsigned long long)red[1]);
    }
}

// A whole path per thread (no queues): camera ray of path i through pixel
// (start + i * stride) mod (w h), closest hits and shadow rays traced in place.
// Same vertex logic, random stream and order of radiance additions as the
// wavefront kernels, so the radiance is bit-identical.
// L: the path's radiance so far (the pixel's sum of earlier passes); every
// contribution is added to it in the wavefront's order.
template <int KIND, int STACK>
__device__ __forceinline__ float4 whole_path(const SceneDev& S, const CamDev& cam, uint32_t w, uint32_t h,
                                             uint32_t i, uint32_t start, uint32_t stride, uint64_t initstate,
                                             const PtParams& P, const ShadeHooks& H, uint2* stk, int sstride,
                                             float4 L, uint32_t& n_ext, uint32_t& n_sh) {
    const uint32_t px = (uint32_t)(((uint64_t)start + (uint64_t)i * stride) % ((uint64_t)w * h));
    if (KIND == 2) H.nrec[i] = 0;
    const V3 d0 = camera_dir(cam, px % w, px / w, w, h);
    const Ray r{cam.pos.x, cam.pos.y, cam.pos.z, 0.0f, d0.x, d0.y, d0.z, 1e30f};
    return path_rest<KIND, STACK>(S, P, H, i, 0u, r, make_float4(1.0f, 1.0f, 1.0f, 0.0f), rng_seed(initstate, i),
                                  stk, sstride, L, n_ext, n_sh);
}

// Lanes a persistent traversal wave waits for before refilling idle lanes.
constexpr int kRefillLanes = 16;

// Shadow-queue entry i as a ray; its contribution, added when unoccluded.
__device__ __forceinline__ Ray shadow_ray(const ShadowQueue& sq, uint32_t i) {
    float4 a = sq.org[i], d = sq.dir[i];
    return Ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
}
__device__ __forceinline__ void shadow_done(const ShadowQueue& sq, uint32_t i, const TraceResult& res,
                                            float4* __restrict__ rad) {
    if (res.rec == 0xFFFFFFFFu) {
        float4 c = sq.contrib[i];
        uint32_t px = __float_as_uint(c.w);
        float4 o = rad[px];
        rad[px] = make_float4(o.x + c.x, o.y + c.y, o.z + c.z, o.w);
    }
}

}  // namespace snkpt
