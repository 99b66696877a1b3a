// Shading math shared by the path-trace, ReSTIR and cache kernels.
//
// BRDF: the reference's Cook-Torrance model (LightBuffers.glsl:65-157):
//   f0   = mix(0.04, albedo, metallic)                        (RayGen.rgen:552-553)
//   F    = f0 + (1 - f0) (1 - max(h.v, 0))^5                  FresnelSchlick
//   D    = a^2 / (PI ((n.h)^2 (a^2 - 1) + 1)^2), a = roughness^2   DistributionGGX
//   G    = G1(n.v) G1(n.l), G1(x) = x / (x (1 - k) + k), k = (r + 1)^2 / 8
//   spec = D G F / (4 max(n.v, 0) max(n.l, 0) + 1e-4)
//   f    = (1 - F)(1 - metallic) albedo / PI + spec           CalcDirectionalLight form
// with PI = 3.1415926 as in LightBuffers.glsl:29.  Materials flagged LAMBERT use
// f = albedo / PI.  Both faces shade (no cull flags, Q5): normals are flipped to
// the side the ray arrives from.
//
// Sampling (no reference counterpart; the reference's BounceDiffuse is not an
// unbiased estimator): a mixture of cosine-weighted hemisphere sampling and GGX
// normal sampling (Walter et al. 2007), light sampling of emissive triangles
// by power with uniform area sampling, combined with the power heuristic (Veach 1997).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace snkpt {

constexpr float kRefPi = 3.1415926f;  // the reference's PI (LightBuffers.glsl:29)
constexpr float kPi = 3.14159265358979323846f;

struct V3 {
    float x, y, z;
};
__device__ __host__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __host__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __host__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __host__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __host__ __forceinline__ V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __host__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __host__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize(V3 a) {
    float l = sqrtf(dot(a, a));
    return l > 0.0f ? V3{a.x / l, a.y / l, a.z / l} : a;
}
__device__ __host__ __forceinline__ float luminance(V3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }

// PCG32 (O'Neill 2014); the stream (inc) is fixed per path, only the state moves.
struct Rng {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    __device__ __forceinline__ float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};
__device__ __forceinline__ Rng rng_seed(uint64_t initstate, uint64_t initseq) {
    Rng r{0u, (initseq << 1u) | 1u};
    r.next();
    r.state += initstate;
    r.next();
    return r;
}
// initstate of sample pass `pass` of frame `frame` under `seed`
__device__ __host__ __forceinline__ uint64_t pass_state(uint32_t seed, uint32_t frame, uint32_t pass) {
    uint64_t x = ((uint64_t)seed << 32) ^ ((uint64_t)frame * 0x9E3779B97F4A7C15ULL) ^
                 ((uint64_t)pass * 0xBF58476D1CE4E5B9ULL);
    x ^= x >> 31;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 29;
    return x;
}

struct Mat {
    V3 albedo;
    float roughness, metallic;
    uint32_t flags;
};

// Orthonormal basis around n (Duff et al. 2017).
__device__ __forceinline__ void onb(V3 n, V3& t, V3& b) {
    float s = copysignf(1.0f, n.z);
    float a = -1.0f / (s + n.z);
    float bb = n.x * n.y * a;
    t = V3{1.0f + s * n.x * n.x * a, s * bb, -s * n.x};
    b = V3{bb, s + n.y * n.y * a, -n.y};
}

__device__ __forceinline__ float ggx_d(float ndh, float a2) {
    float d = ndh * ndh * (a2 - 1.0f) + 1.0f;
    return a2 / (kRefPi * d * d);
}
__device__ __forceinline__ float g1_schlick(float x, float k) { return x / (x * (1.0f - k) + k); }

// f(wo, wi) without the cosine; n is the shading normal facing wo.
__device__ __forceinline__ V3 brdf_eval(const Mat& m, V3 n, V3 wo, V3 wi) {
    if (m.flags & 2u) return m.albedo * (1.0f / kRefPi);
    V3 f0 = V3{0.04f + (m.albedo.x - 0.04f) * m.metallic, 0.04f + (m.albedo.y - 0.04f) * m.metallic,
               0.04f + (m.albedo.z - 0.04f) * m.metallic};
    V3 h = normalize(wo + wi);
    float hv = fmaxf(dot(h, wo), 0.0f);
    float c1 = fminf(fmaxf(1.0f - hv, 0.0f), 1.0f), c2 = c1 * c1;
    float fw = (c2 * c2) * c1;  // pow(x, 5.0)
    V3 F = V3{f0.x + (1.0f - f0.x) * fw, f0.y + (1.0f - f0.y) * fw, f0.z + (1.0f - f0.z) * fw};
    float a = m.roughness * m.roughness;
    float D = ggx_d(fmaxf(dot(n, h), 0.0f), a * a);
    float r = m.roughness + 1.0f;
    float k = (r * r) / 8.0f;
    float nv = fmaxf(dot(n, wo), 0.0f), nl = fmaxf(dot(n, wi), 0.0f);
    float G = g1_schlick(nv, k) * g1_schlick(nl, k);
    float denom = 4.0f * nv * nl + 0.0001f;
    V3 spec = F * (D * G / denom);
    float kdm = 1.0f - m.metallic;
    V3 diff = V3{(1.0f - F.x) * kdm * m.albedo.x, (1.0f - F.y) * kdm * m.albedo.y, (1.0f - F.z) * kdm * m.albedo.z} *
              (1.0f / kRefPi);
    return diff + spec;
}

__device__ __forceinline__ float spec_prob(const Mat& m) { return (m.flags & 2u) ? 0.0f : 0.5f + 0.5f * m.metallic; }

// Solid-angle pdf of brdf_sample producing wi.
__device__ __forceinline__ float brdf_pdf(const Mat& m, V3 n, V3 wo, V3 wi) {
    float ci = dot(n, wi);
    if (ci <= 0.0f) return 0.0f;
    float ps = spec_prob(m);
    float pdf = (1.0f - ps) * ci * (1.0f / kPi);
    if (ps > 0.0f) {
        V3 h = normalize(wo + wi);
        float a = m.roughness * m.roughness;
        float a2 = a * a;
        float nh = fmaxf(dot(n, h), 0.0f);
        float d = nh * nh * (a2 - 1.0f) + 1.0f;
        float D = a2 / (kPi * d * d);
        float wh = fabsf(dot(wo, h));
        if (wh > 0.0f) pdf += ps * D * nh / (4.0f * wh);
    }
    return pdf;
}

// Samples wi; returns false if the sample is below the surface.
__device__ __forceinline__ bool brdf_sample(const Mat& m, V3 n, V3 wo, float ul, float u1, float u2, V3& wi) {
    V3 t, b;
    onb(n, t, b);
    float ps = spec_prob(m);
    if (ul < ps) {
        float a = m.roughness * m.roughness;
        float a2 = a * a;
        float c2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
        float ct = sqrtf(fmaxf(c2, 0.0f)), st = sqrtf(fmaxf(1.0f - c2, 0.0f));
        float ph = 2.0f * kPi * u2;
        V3 h = t * (st * cosf(ph)) + b * (st * sinf(ph)) + n * ct;
        float wh = dot(wo, h);
        wi = h * (2.0f * wh) - wo;
    } else {
        float r = sqrtf(u1), ph = 2.0f * kPi * u2;
        float z = sqrtf(fmaxf(1.0f - u1, 0.0f));
        wi = t * (r * cosf(ph)) + b * (r * sinf(ph)) + n * z;
    }
    return dot(n, wi) > 0.0f;
}

__device__ __forceinline__ float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

}  // namespace snkpt
