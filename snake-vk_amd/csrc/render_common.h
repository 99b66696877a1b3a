// Device views of the scene and the wavefront queues used by the pass kernels.
#pragma once

#include "shading.h"
#include "snkpt_dev.h"
#include "snkpt_host.h"

namespace snkpt {

struct InstD {
    uint32_t transform_idx, index_offset, vertex_offset, material_idx, num_indices;
};
struct MatD {  // snkpt_material as laid out in TexMatBuffers.glsl:5-20
    uint32_t tex[5];
    float roughness, metallic, ao;
    float albedo[4];
    uint32_t flags, pad[3];
};

struct SceneDev {
    const InstD* inst;
    const float4* xf;  // 6 float4 per instance: object->world rows, normal-matrix rows
    const uint32_t* tri_off;
    const MatD* mats;
    const uint32_t* idx;
    const float* nrm;
    const float4* world;  // 3 float4 per global triangle
    // 4 float4 per emitter: v0 | alias probability bits, v1 | alias index bits,
    // v2 | selection probability, emission rgb | area
    const float4* emit_rec;
    const uint32_t* gid2e;
    uint32_t n_emit;
    const uint4* nodes;
    const float4* leaf;
    // analytic lights of the light block (directional, point, spot), 4 float4 each:
    // colour rgb | type (0 dir, 1 point, 2 spot); position (direction towards a
    // directional light) | spot aperture; spot direction; attenuation c, l, e
    const float4* alights;
    uint32_t n_alights;
    float q_tri;  // probability that NEE samples the emissive triangles (else an analytic light)
};

// Surface data at a hit.
struct Surf {
    V3 p, ng, ns;  // position, geometric and shading normals (both facing the incoming ray)
    V3 nraw;       // shading normal as interpolated and transformed (not turned to the ray)
    Mat m;
    V3 le;         // emission (albedo of an emissive material, TexMatBuffers.glsl:3)
    bool emissive;
    uint32_t gid;
    float eps;     // ray offset along ng
};

__device__ __forceinline__ V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }

// RayClosestHit.rchit:88-123: barycentric interpolation of the hit triangle,
// normal through transpose(inverse(mat3(M))), material lookup.
__device__ __forceinline__ Surf surface_at(const SceneDev& S, V3 dir, uint32_t inst, uint32_t prim, float u, float v) {
    Surf s;
    InstD in = S.inst[inst];
    s.gid = S.tri_off[inst] + prim;
    const float4* w = S.world + 3 * (size_t)s.gid;
    float4 r0 = w[0], r1 = w[1], r2 = w[2];
    V3 v0{r0.x, r0.y, r0.z}, v1{r0.w, r1.x, r1.y}, v2{r1.z, r1.w, r2.x};
    float b0 = 1.0f - u - v;
    s.p = v0 * b0 + v1 * u + v2 * v;
    V3 ng = normalize(cross(v1 - v0, v2 - v0));
    const uint32_t* ix = S.idx + in.index_offset + 3 * prim;
    V3 n0 = ld3(S.nrm + 3 * (size_t)(ix[0] + in.vertex_offset));
    V3 n1 = ld3(S.nrm + 3 * (size_t)(ix[1] + in.vertex_offset));
    V3 n2 = ld3(S.nrm + 3 * (size_t)(ix[2] + in.vertex_offset));
    V3 nl = n0 * b0 + n1 * u + n2 * v;
    const float4* x = S.xf + 6 * (size_t)inst;
    float4 m0 = x[3], m1 = x[4], m2 = x[5];
    V3 ns = V3{(m0.x * nl.x + m0.y * nl.y) + m0.z * nl.z, (m1.x * nl.x + m1.y * nl.y) + m1.z * nl.z,
               (m2.x * nl.x + m2.y * nl.y) + m2.z * nl.z};
    float l2 = dot(ns, ns);
    ns = l2 > 1e-20f ? ns * (1.0f / sqrtf(l2)) : ng;
    s.nraw = ns;
    if (dot(ng, dir) > 0.0f) ng = ng * -1.0f;  // face the incoming ray
    if (dot(ns, ng) < 0.0f) ns = ns * -1.0f;
    s.ng = ng;
    s.ns = ns;
    const MatD& md = S.mats[in.material_idx];
    s.m.albedo = V3{md.albedo[0], md.albedo[1], md.albedo[2]};
    s.m.roughness = fminf(fmaxf(md.roughness, 0.02f), 1.0f);
    s.m.metallic = fminf(fmaxf(md.metallic, 0.0f), 1.0f);
    s.m.flags = md.flags;
    s.emissive = (md.flags & 1u) != 0u;
    s.le = s.m.albedo;
    float am = fmaxf(fmaxf(fabsf(s.p.x), fabsf(s.p.y)), fabsf(s.p.z));
    s.eps = 1e-4f * (1.0f + am);
    return s;
}

// Shadow ray of a light sample seen from p: it leaves the biased origin
// o = p + ng eps and is aimed at the sampled light point p + wi dist, measured
// from o (ShadowTest from biased_pos, RayGen.rgen:101-121,552), ending 1e-4
// short of it; with a large bias (G-buffer vertices) the unbiased direction
// would overshoot into the light's own surface.  Lights at infinity keep wi.
__device__ __forceinline__ void shadow_segment(V3 p, V3 ng, float eps, V3 wi, float dist, V3& o, V3& d,
                                               float& tmax) {
    o = p + ng * eps;
    if (dist >= 1e29f) {
        d = wi;
        tmax = dist * (1.0f - 1e-4f) - eps;
        return;
    }
    const V3 e = (p + wi * dist) - o;
    const float l = sqrtf(dot(e, e));
    d = l > 0.0f ? e * (1.0f / l) : wi;
    tmax = l * (1.0f - 1e-4f);
}

struct LightSample {
    V3 wi, le;
    float dist, pdf;  // solid-angle pdf; for a delta light the selection probability
    bool ok;
    bool delta;       // analytic light: BRDF sampling cannot hit it (no MIS)
};

// Record of the emitter the alias table picks for uniform index e and draw u:
// e's own record, or its alias's when u >= e's alias probability.
__device__ __forceinline__ uint32_t emitter_record(const SceneDev& S, uint32_t e, float u, float4& a, float4& b,
                                                   float4& c, float4& info) {
    const float4* r = S.emit_rec + 4 * (size_t)e;
    a = r[0];
    b = r[1];
    c = r[2];
    info = r[3];
    if (u >= a.w) {
        e = __float_as_uint(b.w);
        r = S.emit_rec + 4 * (size_t)e;
        a = r[0];
        b = r[1];
        c = r[2];
        info = r[3];
    }
    return e;
}

// Power-proportional emitter choice (alias table), uniform point on it.
__device__ __forceinline__ LightSample sample_light(const SceneDev& S, V3 p, float u0, float u1, float u2,
                                                    float u3) {
    LightSample ls;
    ls.ok = false;
    ls.delta = false;
    if (S.n_emit == 0) return ls;
    uint32_t e = (uint32_t)(u0 * (float)S.n_emit);
    if (e >= S.n_emit) e = S.n_emit - 1;
    float4 a, b, c, info;
    emitter_record(S, e, u1, a, b, c, info);
    float su = sqrtf(u2);
    float w1 = su * (1.0f - u3), w2 = su * u3, w0 = 1.0f - su;
    V3 A{a.x, a.y, a.z}, B{b.x, b.y, b.z}, C{c.x, c.y, c.z};
    V3 q = A * w0 + B * w1 + C * w2;
    V3 d = q - p;
    float d2 = dot(d, d);
    if (!(d2 > 0.0f)) return ls;
    float dist = sqrtf(d2);
    ls.wi = d * (1.0f / dist);
    V3 nl = cross(B - A, C - A);
    float nll = sqrtf(dot(nl, nl));
    if (!(nll > 0.0f)) return ls;
    float cosl = fabsf(dot(nl, ls.wi)) / nll;  // two-sided emitters (RayClosestHit.rchit:119-123)
    if (cosl <= 0.0f) return ls;
    ls.le = V3{info.x, info.y, info.z};
    ls.pdf = c.w * d2 / (info.w * cosl);
    ls.dist = dist;
    ls.ok = ls.pdf > 0.0f;
    return ls;
}

// Analytic light k of the light block (LightBuffers.glsl:
// CalcDirectionalLight, CalcPointlight, CalcSpotlight): radiance arriving from
// a delta light, attenuation c + l d + e d^2, spot falloff
// 1 - (1 - cos) / (1 - aperture) inside the aperture, nothing beyond it or where
// the attenuation exceeds 5000.  pdf = 1 / n, its probability under a uniform choice.
__device__ __forceinline__ LightSample analytic_light(const SceneDev& S, uint32_t k, V3 p) {
    LightSample ls;
    ls.ok = false;
    ls.delta = true;
    const float4* L = S.alights + 4 * (size_t)k;
    const float4 w0 = L[0], w1 = L[1];
    const uint32_t type = (uint32_t)w0.w;
    V3 col{w0.x, w0.y, w0.z};
    ls.pdf = 1.0f / (float)S.n_alights;
    if (type == 0u) {
        ls.wi = V3{w1.x, w1.y, w1.z};
        ls.dist = 1e30f;
        ls.le = col;
        ls.ok = true;
        return ls;
    }
    V3 d = V3{w1.x, w1.y, w1.z} - p;
    float d2 = dot(d, d);
    if (!(d2 > 0.0f)) return ls;
    float dist = sqrtf(d2);
    ls.wi = d * (1.0f / dist);
    const float4 w3 = L[3];
    float atten = (w3.x + w3.y * dist) + w3.z * d2;
    if (!(atten > 0.0f)) return ls;
    float f = 1.0f;
    if (type == 2u) {
        const float4 w2 = L[2];
        float sf = -dot(ls.wi, V3{w2.x, w2.y, w2.z});
        if (sf < 0.0001f || sf < w1.w || atten > 5000.0f) return ls;
        f = 1.0f - (1.0f - sf) / fmaxf(1.0f - w1.w, 1e-5f);
    }
    ls.le = col * (f / atten);
    ls.dist = dist;
    ls.ok = ls.le.x > 0.0f || ls.le.y > 0.0f || ls.le.z > 0.0f;
    return ls;
}
__device__ __forceinline__ LightSample sample_analytic(const SceneDev& S, V3 p, float u) {
    if (S.n_alights == 0) {
        LightSample ls;
        ls.ok = false;
        ls.delta = true;
        return ls;
    }
    uint32_t k = (uint32_t)(u * (float)S.n_alights);
    if (k >= S.n_alights) k = S.n_alights - 1;
    return analytic_light(S, k, p);
}

// Next-event estimation: the emissive triangles with probability q_tri (u0
// rescaled, so with q_tri = 1 this is exactly sample_light), else one analytic
// light; the returned pdf includes the choice of set.
__device__ __forceinline__ LightSample sample_direct(const SceneDev& S, V3 p, float u0, float u1, float u2,
                                                     float u3) {
    if (u0 < S.q_tri) {
        LightSample ls = sample_light(S, p, u0 / S.q_tri, u1, u2, u3);
        ls.pdf *= S.q_tri;
        return ls;
    }
    LightSample ls = sample_analytic(S, p, (u0 - S.q_tri) / (1.0f - S.q_tri));
    ls.pdf *= 1.0f - S.q_tri;
    return ls;
}

// Solid-angle pdf with which sample_direct would pick the point hit on global
// triangle gid.
__device__ __forceinline__ float light_pdf(const SceneDev& S, uint32_t gid, V3 ng_hit, V3 dir, float t) {
    if (S.n_emit == 0) return 0.0f;
    uint32_t e = S.gid2e[gid];
    if (e == 0xFFFFFFFFu) return 0.0f;
    float cosl = fabsf(dot(ng_hit, dir));
    if (cosl <= 0.0f) return 0.0f;
    const float4* r = S.emit_rec + 4 * (size_t)e;
    return S.q_tri * (r[2].w * t * t / (r[3].w * cosl));
}

// Wave-level queue append: every lane calls it (convergent); lanes with
// `want` get consecutive slots from one atomicAdd per wave.
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t* counter, unsigned long long* stat) {
    uint64_t m = __ballot(want);
    if (m == 0) return 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) {
        uint32_t n = (uint32_t)__popcll(m);
        base = atomicAdd(counter, n);
        if (stat) atomicAdd(stat, (unsigned long long)n);
    }
    base = __shfl(base, (int)leader, 64);
    return want ? base + (uint32_t)__popcll(m & lt) : 0xFFFFFFFFu;
}

}  // namespace snkpt
