/*
 * snkpt oracle: the unbiased path tracer, restated on the CPU.  TEST
 * INFRASTRUCTURE ONLY (checker and CPU baseline, never the product).
 *
 * Follows the reference pass where it is defined:
 *   - camera ray through the pixel centre: inverse(view) * inverse(proj) at depth 1,
 *     normalize(p - cam_pos) (Util.glsl:1-13);
 *   - hit interpolation: bary = (1-u-v, u, v), normal through
 *     transpose(inverse(mat3(M))) (RayClosestHit.rchit:88-111);
 *   - emission = material albedo when flags & 1 (TexMatBuffers.glsl:3,
 *     RayClosestHit.rchit:119-123), both faces hit and shade (Q5);
 *   - BRDF: Cook-Torrance with Schlick Fresnel, GGX D, Schlick-GGX G with
 *     k = (r+1)^2/8, spec = DGF / (4 n.v n.l + 1e-4), diffuse (1-F)(1-metallic)
 *     albedo / PI, PI = 3.1415926, f0 = mix(0.04, albedo, metallic)
 *     (LightBuffers.glsl:65-157, RayGen.rgen:552-553).
 * and replaces the reference's biased BounceDiffuse (RayGen.rgen:139-213) with
 * the estimator this build specifies: NEE to power-sampled emissive triangles +
 * cosine/GGX BRDF sampling, power-heuristic MIS, Russian roulette.  Random
 * numbers come from PCG32 (initstate from (seed, frame, pass), stream = pixel)
 * drawn in the same order as the GPU, so most pixels reproduce the GPU's
 * values; sinf/cosf of the platform library may differ in the last ulp.
 * Shadow rays leave p + ng eps and are aimed at the sampled light point from
 * there (ShadowTest from biased_pos, RayGen.rgen:101-121,552).
 * o_compat_bounce restates BounceDiffuse itself (RayGen.rgen:123-213, payload
 * RayClosestHit.rchit:87-142) for the reference-compatibility mode, pinned by
 * the closed forms of tests/test_oracle_compat.py.
 * G-buffer mode (o_pt_render_gbuffer): the primary vertex is read from the
 * reference's G-buffer images (RayGen.rgen:531-552): WorldPosFromDepth
 * (Util.glsl:1-7), texel normal, SNORM8 albedo, RGBA16F rma, R16 flags, ray bias
 * 0.01 + max(0.001 |cam - p|, 0.01); an emissive texel's radiance is its albedo.
 * Pinned by: closed forms of the BRDF terms, pdf normalisation, the furnace
 * test (tests/test_oracle_pt.py), the polygon-formula closed form for G-buffer
 * texels and an exhaustive binary16 check (tests/test_oracle_gbuffer.py), and
 * agreement with the GPU in expectation.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "snkpt_oracle.h"

#define REF_PI 3.1415926f
#define PI_F 3.14159265358979323846f

typedef struct {
    float x, y, z;
} v3;
static inline v3 V(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 mulv(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dotv(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 crossv(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 nrmz(v3 a) {
    float l = sqrtf(dotv(a, a));
    return l > 0.0f ? V(a.x / l, a.y / l, a.z / l) : a;
}
static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* ------------------------------------------------------------------ rng --- */
typedef struct {
    uint64_t state, inc;
} prng;
static inline uint32_t rnext(prng* r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}
static inline float runi(prng* r) { return (float)(rnext(r) >> 8) * (1.0f / 16777216.0f); }
static prng rseed(uint64_t initstate, uint64_t initseq) {
    prng r = {0u, (initseq << 1u) | 1u};
    rnext(&r);
    r.state += initstate;
    rnext(&r);
    return r;
}
uint64_t o_pass_state(uint32_t seed, uint32_t frame, uint32_t pass) {
    uint64_t x = ((uint64_t)seed << 32) ^ ((uint64_t)frame * 0x9E3779B97F4A7C15ULL) ^
                 ((uint64_t)pass * 0xBF58476D1CE4E5B9ULL);
    x ^= x >> 31;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 29;
    return x;
}

/* ----------------------------------------------------------------- BRDF --- */
typedef struct {
    v3 albedo;
    float roughness, metallic;
    uint32_t flags;
} omat;

static float ggx_d_ref(float ndh, float a2) {
    float d = ndh * ndh * (a2 - 1.0f) + 1.0f;
    return a2 / (REF_PI * d * d);
}
static float g1s(float x, float k) { return x / (x * (1.0f - k) + k); }

static v3 brdf_f(const omat* m, v3 n, v3 wo, v3 wi) {
    if (m->flags & 2u) return scl(m->albedo, 1.0f / REF_PI);
    v3 f0 = V(0.04f + (m->albedo.x - 0.04f) * m->metallic, 0.04f + (m->albedo.y - 0.04f) * m->metallic,
              0.04f + (m->albedo.z - 0.04f) * m->metallic);
    v3 h = nrmz(add(wo, wi));
    float hv = fmaxf(dotv(h, wo), 0.0f);
    float c1 = fminf(fmaxf(1.0f - hv, 0.0f), 1.0f), c2 = c1 * c1;
    float fw = (c2 * c2) * c1;
    v3 F = V(f0.x + (1.0f - f0.x) * fw, f0.y + (1.0f - f0.y) * fw, f0.z + (1.0f - f0.z) * fw);
    float a = m->roughness * m->roughness;
    float D = ggx_d_ref(fmaxf(dotv(n, h), 0.0f), a * a);
    float r = m->roughness + 1.0f;
    float k = (r * r) / 8.0f;
    float nv = fmaxf(dotv(n, wo), 0.0f), nl = fmaxf(dotv(n, wi), 0.0f);
    float G = g1s(nv, k) * g1s(nl, k);
    float denom = 4.0f * nv * nl + 0.0001f;
    v3 spec = scl(F, D * G / denom);
    float kdm = 1.0f - m->metallic;
    v3 diff = scl(V((1.0f - F.x) * kdm * m->albedo.x, (1.0f - F.y) * kdm * m->albedo.y,
                    (1.0f - F.z) * kdm * m->albedo.z),
                  1.0f / REF_PI);
    return add(diff, spec);
}

static float spec_p(const omat* m) { return (m->flags & 2u) ? 0.0f : 0.5f + 0.5f * m->metallic; }

static float brdf_pdf_c(const omat* m, v3 n, v3 wo, v3 wi) {
    float ci = dotv(n, wi);
    if (ci <= 0.0f) return 0.0f;
    float ps = spec_p(m);
    float pdf = (1.0f - ps) * ci * (1.0f / PI_F);
    if (ps > 0.0f) {
        v3 h = nrmz(add(wo, wi));
        float a = m->roughness * m->roughness;
        float a2 = a * a;
        float nh = fmaxf(dotv(n, h), 0.0f);
        float d = nh * nh * (a2 - 1.0f) + 1.0f;
        float D = a2 / (PI_F * d * d);
        float wh = fabsf(dotv(wo, h));
        if (wh > 0.0f) pdf += ps * D * nh / (4.0f * wh);
    }
    return pdf;
}

static void onb_c(v3 n, v3* t, v3* b) {
    float s = copysignf(1.0f, n.z);
    float a = -1.0f / (s + n.z);
    float bb = n.x * n.y * a;
    *t = V(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
    *b = V(bb, s + n.y * n.y * a, -n.y);
}

static int brdf_sample_c(const omat* m, v3 n, v3 wo, float ul, float u1, float u2, v3* wi) {
    v3 t, b;
    onb_c(n, &t, &b);
    float ps = spec_p(m);
    if (ul < ps) {
        float a = m->roughness * m->roughness;
        float a2 = a * a;
        float c2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
        float ct = sqrtf(fmaxf(c2, 0.0f)), st = sqrtf(fmaxf(1.0f - c2, 0.0f));
        float ph = 2.0f * PI_F * u2;
        v3 h = add(add(scl(t, st * cosf(ph)), scl(b, st * sinf(ph))), scl(n, ct));
        float wh = dotv(wo, h);
        *wi = sub(scl(h, 2.0f * wh), wo);
    } else {
        float r = sqrtf(u1), ph = 2.0f * PI_F * u2;
        float z = sqrtf(fmaxf(1.0f - u1, 0.0f));
        *wi = add(add(scl(t, r * cosf(ph)), scl(b, r * sinf(ph))), scl(n, z));
    }
    return dotv(n, *wi) > 0.0f;
}

static float power_h(float a, float b) {
    float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

/* exported for the closed-form tests: f (rgb) and pdf for one configuration */
void o_brdf_eval(const float albedo[3], float roughness, float metallic, uint32_t flags, const float n[3],
                 const float wo[3], const float wi[3], float out_f[3], float* out_pdf) {
    omat m = {V(albedo[0], albedo[1], albedo[2]), fminf(fmaxf(roughness, 0.02f), 1.0f), metallic, flags};
    v3 f = brdf_f(&m, V(n[0], n[1], n[2]), V(wo[0], wo[1], wo[2]), V(wi[0], wi[1], wi[2]));
    out_f[0] = f.x;
    out_f[1] = f.y;
    out_f[2] = f.z;
    *out_pdf = brdf_pdf_c(&m, V(n[0], n[1], n[2]), V(wo[0], wo[1], wo[2]), V(wi[0], wi[1], wi[2]));
}
int o_brdf_sample(const float albedo[3], float roughness, float metallic, uint32_t flags, const float n[3],
                  const float wo[3], float ul, float u1, float u2, float out_wi[3]) {
    omat m = {V(albedo[0], albedo[1], albedo[2]), fminf(fmaxf(roughness, 0.02f), 1.0f), metallic, flags};
    v3 wi;
    int ok = brdf_sample_c(&m, V(n[0], n[1], n[2]), V(wo[0], wo[1], wo[2]), ul, u1, u2, &wi);
    out_wi[0] = wi.x;
    out_wi[1] = wi.y;
    out_wi[2] = wi.z;
    return ok;
}

/* --------------------------------------------------------------- scene --- */
typedef struct {
    const float* world;
    uint32_t ntris;
    const o_bvh* bvh;
    const uint32_t* inst5;
    uint32_t n_inst;
    const uint32_t* idx;
    const float* nrm;
    const uint32_t* mats16;
    uint32_t* tri_off;     /* n_inst + 1 */
    float* nmat;           /* 9 per instance: transpose(inverse(mat3(M))) rows */
    /* emitters */
    uint32_t n_emit;
    float* e_tri;          /* 9 per emitter */
    float* e_info;         /* rgb, area */
    float* e_prob;
    uint32_t* e_alias;
    float* e_pdf;
    uint32_t* gid2e;
} oscene;

static void normal_matrix(const float* m, float out[9]) {
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = m[4 * c + r];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv_det = det != 0.0 ? 1.0 / det : 0.0;
    double cof[3][3];
    cof[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    cof[0][1] = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]);
    cof[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    cof[1][0] = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]);
    cof[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    cof[1][2] = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]);
    cof[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    cof[2][1] = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]);
    cof[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = (float)(cof[r][c] * inv_det);
}

/* Power-proportional alias table (Vose), in the order of the emissive instance list. */
static void build_emitters_c(oscene* S, const uint32_t* em_list, uint32_t n_em) {
    S->n_emit = 0;
    uint32_t cap = 0;
    for (uint32_t k = 0; k < n_em; ++k) cap += S->inst5[5 * em_list[k] + 4] / 3;
    S->e_tri = (float*)calloc(9 * (size_t)(cap + 1), sizeof(float));
    S->e_info = (float*)calloc(4 * (size_t)(cap + 1), sizeof(float));
    S->e_prob = (float*)calloc(cap + 1, sizeof(float));
    S->e_alias = (uint32_t*)calloc(cap + 1, sizeof(uint32_t));
    S->e_pdf = (float*)calloc(cap + 1, sizeof(float));
    S->gid2e = (uint32_t*)malloc(sizeof(uint32_t) * (S->ntris + 1));
    for (uint32_t g = 0; g <= S->ntris; ++g) S->gid2e[g] = 0xFFFFFFFFu;
    double* power = (double*)calloc(cap + 1, sizeof(double));
    for (uint32_t k = 0; k < n_em; ++k) {
        uint32_t i = em_list[k];
        const uint32_t* mt = S->mats16 + 16 * (size_t)S->inst5[5 * i + 3];
        float al[3] = {u2f(mt[8]), u2f(mt[9]), u2f(mt[10])};
        double lum = 0.2126 * al[0] + 0.7152 * al[1] + 0.0722 * al[2];
        for (uint32_t p = 0; p < S->inst5[5 * i + 4] / 3; ++p) {
            uint32_t g = S->tri_off[i] + p;
            const float* w = S->world + 12 * (size_t)g;
            double e1[3], e2[3];
            for (int a = 0; a < 3; ++a) {
                e1[a] = (double)w[3 + a] - w[a];
                e2[a] = (double)w[6 + a] - w[a];
            }
            double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                   cz = e1[0] * e2[1] - e1[1] * e2[0];
            double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            double pw = lum * area;
            if (!(pw > 0.0)) continue;
            uint32_t e = S->n_emit++;
            memcpy(S->e_tri + 9 * (size_t)e, w, 36);
            S->e_info[4 * e] = al[0];
            S->e_info[4 * e + 1] = al[1];
            S->e_info[4 * e + 2] = al[2];
            S->e_info[4 * e + 3] = (float)area;
            power[e] = pw;
            S->gid2e[g] = e;
        }
    }
    uint32_t ne = S->n_emit;
    double total = 0;
    for (uint32_t e = 0; e < ne; ++e) total += power[e];
    if (ne) {
        double* scaled = (double*)malloc(sizeof(double) * ne);
        uint32_t *small = (uint32_t*)malloc(sizeof(uint32_t) * ne), *large = (uint32_t*)malloc(sizeof(uint32_t) * ne);
        uint32_t ns = 0, nl = 0;
        for (uint32_t e = 0; e < ne; ++e) {
            scaled[e] = power[e] / total * ne;
            S->e_pdf[e] = (float)(power[e] / total);
            S->e_prob[e] = 1.0f;
            if (scaled[e] < 1.0)
                small[ns++] = e;
            else
                large[nl++] = e;
        }
        while (ns && nl) {
            uint32_t s = small[--ns], l = large[nl - 1];
            S->e_prob[s] = (float)scaled[s];
            S->e_alias[s] = l;
            scaled[l] = (scaled[l] + scaled[s]) - 1.0;
            if (scaled[l] < 1.0) {
                --nl;
                small[ns++] = l;
            }
        }
        while (nl) {
            uint32_t l = large[--nl];
            S->e_prob[l] = 1.0f;
            S->e_alias[l] = l;
        }
        while (ns) {
            uint32_t s = small[--ns];
            S->e_prob[s] = 1.0f;
            S->e_alias[s] = s;
        }
        free(scaled);
        free(small);
        free(large);
    }
    free(power);
}

typedef struct {
    v3 p, ng, ns;
    v3 nraw; /* shading normal as interpolated and transformed, not turned to the ray */
    omat m;
    v3 le;
    int emissive;
    uint32_t gid;
    float eps;
} osurf;

static osurf surface(const oscene* S, v3 dir, uint32_t inst, uint32_t prim, float u, float v) {
    osurf s;
    const uint32_t* in = S->inst5 + 5 * inst;
    s.gid = S->tri_off[inst] + prim;
    const float* w = S->world + 12 * (size_t)s.gid;
    v3 v0 = V(w[0], w[1], w[2]), v1 = V(w[3], w[4], w[5]), v2 = V(w[6], w[7], w[8]);
    float b0 = 1.0f - u - v;
    s.p = add(add(scl(v0, b0), scl(v1, u)), scl(v2, v));
    v3 ng = nrmz(crossv(sub(v1, v0), sub(v2, v0)));
    const uint32_t* ix = S->idx + in[1] + 3 * prim;
    const float* n0p = S->nrm + 3 * (size_t)(ix[0] + in[2]);
    const float* n1p = S->nrm + 3 * (size_t)(ix[1] + in[2]);
    const float* n2p = S->nrm + 3 * (size_t)(ix[2] + in[2]);
    v3 nl = add(add(scl(V(n0p[0], n0p[1], n0p[2]), b0), scl(V(n1p[0], n1p[1], n1p[2]), u)),
                scl(V(n2p[0], n2p[1], n2p[2]), v));
    const float* nm = S->nmat + 9 * (size_t)inst;
    v3 ns = V((nm[0] * nl.x + nm[1] * nl.y) + nm[2] * nl.z, (nm[3] * nl.x + nm[4] * nl.y) + nm[5] * nl.z,
              (nm[6] * nl.x + nm[7] * nl.y) + nm[8] * nl.z);
    float l2 = dotv(ns, ns);
    ns = l2 > 1e-20f ? scl(ns, 1.0f / sqrtf(l2)) : ng;
    s.nraw = ns;
    if (dotv(ng, dir) > 0.0f) ng = scl(ng, -1.0f);
    if (dotv(ns, ng) < 0.0f) ns = scl(ns, -1.0f);
    s.ng = ng;
    s.ns = ns;
    const uint32_t* mt = S->mats16 + 16 * (size_t)in[3];
    s.m.albedo = V(u2f(mt[8]), u2f(mt[9]), u2f(mt[10]));
    s.m.roughness = fminf(fmaxf(u2f(mt[5]), 0.02f), 1.0f);
    s.m.metallic = fminf(fmaxf(u2f(mt[6]), 0.0f), 1.0f);
    s.m.flags = mt[12];
    s.emissive = (mt[12] & 1u) != 0u;
    s.le = s.m.albedo;
    float am = fmaxf(fmaxf(fabsf(s.p.x), fabsf(s.p.y)), fabsf(s.p.z));
    s.eps = 1e-4f * (1.0f + am);
    return s;
}

typedef struct {
    v3 wi, le;
    float dist, pdf;
    int ok;
} olight;

/* Alias-table draw: u0 picks a column, u1 against its probability keeps it or takes its alias. */
static uint32_t pick_emitter(const oscene* S, float u0, float u1) {
    uint32_t e = (uint32_t)(u0 * (float)S->n_emit);
    if (e >= S->n_emit) e = S->n_emit - 1;
    if (u1 >= S->e_prob[e]) e = S->e_alias[e];
    return e;
}

static olight sample_light_c(const oscene* S, v3 p, float u0, float u1, float u2, float u3) {
    olight ls;
    memset(&ls, 0, sizeof(ls));
    if (S->n_emit == 0) return ls;
    uint32_t e = pick_emitter(S, u0, u1);
    const float* t = S->e_tri + 9 * (size_t)e;
    float su = sqrtf(u2);
    float w1 = su * (1.0f - u3), w2 = su * u3, w0 = 1.0f - su;
    v3 A = V(t[0], t[1], t[2]), B = V(t[3], t[4], t[5]), C = V(t[6], t[7], t[8]);
    v3 q = add(add(scl(A, w0), scl(B, w1)), scl(C, w2));
    v3 d = sub(q, p);
    float d2 = dotv(d, d);
    if (!(d2 > 0.0f)) return ls;
    float dist = sqrtf(d2);
    ls.wi = scl(d, 1.0f / dist);
    v3 nl = crossv(sub(B, A), sub(C, A));
    float nll = sqrtf(dotv(nl, nl));
    if (!(nll > 0.0f)) return ls;
    float cosl = fabsf(dotv(nl, ls.wi)) / nll;
    if (cosl <= 0.0f) return ls;
    const float* info = S->e_info + 4 * (size_t)e;
    ls.le = V(info[0], info[1], info[2]);
    ls.pdf = S->e_pdf[e] * d2 / (info[3] * cosl);
    ls.dist = dist;
    ls.ok = ls.pdf > 0.0f;
    return ls;
}

static float light_pdf_c(const oscene* S, uint32_t gid, v3 ng, v3 dir, float t) {
    if (S->n_emit == 0) return 0.0f;
    uint32_t e = S->gid2e[gid];
    if (e == 0xFFFFFFFFu) return 0.0f;
    float cosl = fabsf(dotv(ng, dir));
    if (cosl <= 0.0f) return 0.0f;
    return S->e_pdf[e] * t * t / (S->e_info[4 * (size_t)e + 3] * cosl);
}

/* The emitter table as o_pt_render builds it, and `n` draws of its emitter
 * choice from PCG32(seed, 0x51).  Writes, per
 * emitter in table order: its world triangle id, its selection pdf, and how
 * often pick_emitter (the sampler o_pt_render uses) drew it.  Returns the emitter count (<= cap). */
uint32_t o_emitter_sample(const float* world, uint32_t ntris, const uint32_t* inst5, uint32_t n_inst,
                          const uint32_t* mats16, const uint32_t* emissive, uint32_t n_emissive, uint64_t seed,
                          uint32_t n, uint32_t cap, uint32_t* out_gid, float* out_pdf, uint32_t* out_count) {
    oscene S;
    memset(&S, 0, sizeof(S));
    S.world = world;
    S.ntris = ntris;
    S.inst5 = inst5;
    S.n_inst = n_inst;
    S.mats16 = mats16;
    S.tri_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_inst + 1));
    S.tri_off[0] = 0;
    for (uint32_t i = 0; i < n_inst; ++i) S.tri_off[i + 1] = S.tri_off[i] + inst5[5 * i + 4] / 3;
    build_emitters_c(&S, emissive, n_emissive);
    uint32_t ne = S.n_emit < cap ? S.n_emit : cap;
    for (uint32_t e = 0; e < ne; ++e) {
        out_pdf[e] = S.e_pdf[e];
        out_count[e] = 0;
    }
    for (uint32_t g = 0; g < ntris; ++g)
        if (S.gid2e[g] < ne) out_gid[S.gid2e[g]] = g;
    prng r = rseed(seed, 0x51u);
    for (uint32_t k = 0; k < n && S.n_emit; ++k) {
        float u0 = runi(&r), u1 = runi(&r);
        uint32_t e = pick_emitter(&S, u0, u1);
        if (e < ne) out_count[e]++;
    }
    uint32_t total = S.n_emit;
    free(S.tri_off);
    free(S.e_tri);
    free(S.e_info);
    free(S.e_prob);
    free(S.e_alias);
    free(S.e_pdf);
    free(S.gid2e);
    return total;
}

/* --------------------------------------------------------------- camera --- */
static int inv4(const double m[16], double inv[16]) {
    /* Gauss-Jordan with partial pivoting */
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = c < 4 ? m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return 0;
        for (int k = 0; k < 8; ++k) {
            double t = a[c][k];
            a[c][k] = a[piv][k];
            a[piv][k] = t;
        }
        double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            double f = a[r][c];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[4 * r + c] = a[r][4 + c];
    return 1;
}

void o_camera_matrix(const float* view16, const float* proj16, float out16[16]) {
    double v[16], p[16], vi[16], pi[16];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            v[4 * r + k] = view16[4 * k + r];
            p[4 * r + k] = proj16[4 * k + r];
        }
    inv4(v, vi);
    inv4(p, pi);
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            double acc = 0;
            for (int j = 0; j < 4; ++j) acc += vi[4 * r + j] * pi[4 * j + k];
            out16[4 * r + k] = (float)acc;
        }
}

static v3 cam_dir(const float* m, v3 pos, uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    float u = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
    float v = ((float)y + 0.5f) / (float)h * 2.0f - 1.0f;
    float wp[4];
    for (int r = 0; r < 4; ++r) wp[r] = ((m[4 * r] * u + m[4 * r + 1] * v) + m[4 * r + 2]) + m[4 * r + 3];
    float iw = 1.0f / fmaxf(wp[3], 1e-6f);
    v3 p = V(wp[0] * iw, wp[1] * iw, wp[2] * iw);
    return nrmz(sub(p, pos));
}

/* ----------------------------------------------------------------- path --- */
void o_trace_one(const o_bvh* b, const float* ray, int any_hit, float* t, uint32_t* inst, uint32_t* prim,
                 float* u, float* v);

typedef struct {
    const oscene* S;
    const float* cam;
    v3 pos;
    uint32_t w, h, spp, max_depth, rr_start, seed, frame, seed_xor;
    float* out;
    uint32_t begin, end;
    const o_gbuffer* gb; /* NULL: camera rays */
} pt_job;

/* ------------------------------------------------------------- G-buffer --- */
/* IEEE binary16 -> binary32 (exact). */
float o_half_to_float(uint16_t hb) {
    uint32_t sgn = (uint32_t)(hb >> 15) << 31, e = (hb >> 10) & 0x1Fu, m = hb & 0x3FFu, bits;
    if (e == 0) {
        bits = m == 0 ? sgn : (f2u((float)m * 5.9604644775390625e-8f) | sgn); /* m * 2^-24 */
    } else if (e == 31) {
        bits = sgn | 0x7F800000u | (m << 13);
    } else {
        bits = sgn | ((e + 112u) << 23) | (m << 13);
    }
    return u2f(bits);
}
/* R8_SNORM texel: max(c / 127, -1) (Vulkan spec, fixed-point conversion). */
static float snorm8_c(int8_t c) { return fmaxf((float)c / 127.0f, -1.0f); }
static float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* The primary vertex of pixel px read from a rasterized G-buffer
 * (RayGen.rgen:531-552): WorldPosFromDepth (Util.glsl:1-7) with the texel's
 * depth, the texel normal as shading and geometric normal (facing the camera),
 * material from albedo / rma / flags, ray bias 0.01 + max(0.001 |cam - p|, 0.01). */
static int gsurface(const pt_job* J, uint32_t px, osurf* s, v3* dir, float* t) {
    const o_gbuffer* g = J->gb;
    const float depth = g->depth[px];
    if (!(depth < 1.0f)) return 0;
    const uint32_t x = px % J->w, y = px / J->w;
    const float u = ((float)x + 0.5f) / (float)J->w * 2.0f - 1.0f;
    const float v = ((float)y + 0.5f) / (float)J->h * 2.0f - 1.0f;
    const float* m = J->cam;
    float wp[4];
    for (int r = 0; r < 4; ++r) wp[r] = ((m[4 * r] * u + m[4 * r + 1] * v) + m[4 * r + 2] * depth) + m[4 * r + 3];
    const float iw = 1.0f / fmaxf(wp[3], 1e-6f);
    s->p = V(wp[0] * iw, wp[1] * iw, wp[2] * iw);
    const v3 d = sub(s->p, J->pos);
    const float l2 = dotv(d, d);
    if (!(l2 > 0.0f)) return 0;
    *t = sqrtf(l2);
    *dir = scl(d, 1.0f / *t);
    const uint16_t* nb = g->normal + 4 * (size_t)px;
    v3 n = V(o_half_to_float(nb[0]), o_half_to_float(nb[1]), o_half_to_float(nb[2]));
    const float n2 = dotv(n, n);
    if (!(n2 > 1e-20f)) return 0;
    n = scl(n, 1.0f / sqrtf(n2));
    if (dotv(n, *dir) > 0.0f) n = scl(n, -1.0f);
    s->ng = n;
    s->ns = n;
    s->nraw = n;
    const int8_t* a = g->albedo + 4 * (size_t)px;
    s->m.albedo = V(clampf(snorm8_c(a[0]), 0.0f, 1.0f), clampf(snorm8_c(a[1]), 0.0f, 1.0f),
                    clampf(snorm8_c(a[2]), 0.0f, 1.0f));
    const uint16_t* rb = g->rma + 4 * (size_t)px;
    s->m.roughness = clampf(o_half_to_float(rb[0]), 0.02f, 1.0f);
    s->m.metallic = clampf(o_half_to_float(rb[1]), 0.0f, 1.0f);
    s->m.flags = g->flags[px];
    s->emissive = (s->m.flags & 1u) != 0u;
    s->le = s->m.albedo;
    s->gid = 0xFFFFFFFFu;
    s->eps = 0.01f + fmaxf(*t * 0.001f, 0.01f);
    return 1;
}

static void* pt_worker(void* arg) {
    pt_job* J = (pt_job*)arg;
    const oscene* S = J->S;
    for (uint32_t px = J->begin; px < J->end; ++px) {
        float rad[3] = {0, 0, 0};
        for (uint32_t k = 0; k < J->spp; ++k) {
            prng rng = rseed(o_pass_state(J->seed ^ J->seed_xor, J->frame, k), px);
            v3 o = J->pos;
            v3 d = cam_dir(J->cam, J->pos, px % J->w, px / J->w, J->w, J->h);
            v3 thr = V(1, 1, 1);
            float last_pdf = 0.0f;
            for (uint32_t bounce = 0; bounce < J->max_depth; ++bounce) {
                osurf s;
                float t;
                if (bounce == 0 && J->gb) {
                    if (!gsurface(J, px, &s, &d, &t)) break;
                } else {
                    float ray[8] = {o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, 1e30f};
                    float hu, hv;
                    uint32_t hi, hp;
                    o_trace_one(S->bvh, ray, 0, &t, &hi, &hp, &hu, &hv);
                    if (!(t >= 0.0f)) break;
                    s = surface(S, d, hi, hp, hu, hv);
                }
                if (s.emissive) {
                    float w = 1.0f;
                    if (bounce > 0) w = power_h(last_pdf, light_pdf_c(S, s.gid, s.ng, d, t));
                    v3 c = scl(mulv(thr, s.le), w);
                    rad[0] = rad[0] + c.x;
                    rad[1] = rad[1] + c.y;
                    rad[2] = rad[2] + c.z;
                    break;
                }
                v3 wo = scl(d, -1.0f);
                float u0 = runi(&rng), u1 = runi(&rng), u2 = runi(&rng), u3 = runi(&rng);
                float ul = runi(&rng), us1 = runi(&rng), us2 = runi(&rng), urr = runi(&rng);
                int last = bounce + 1 >= J->max_depth;
                olight ls = sample_light_c(S, s.p, u0, u1, u2, u3);
                if (ls.ok) {
                    float cs = dotv(s.ns, ls.wi);
                    if (cs > 0.0f && dotv(s.ng, ls.wi) > 0.0f) {
                        v3 f = brdf_f(&s.m, s.ns, wo, ls.wi);
                        float w = last ? 1.0f : power_h(ls.pdf, brdf_pdf_c(&s.m, s.ns, wo, ls.wi));
                        v3 c = scl(mulv(mulv(thr, f), ls.le), cs * w / ls.pdf);
                        if (c.x > 0.0f || c.y > 0.0f || c.z > 0.0f) {
                            /* from the biased origin, aimed at the light point p + wi dist */
                            v3 so = add(s.p, scl(s.ng, s.eps));
                            v3 e = sub(add(s.p, scl(ls.wi, ls.dist)), so);
                            float l = sqrtf(dotv(e, e));
                            v3 sd = l > 0.0f ? scl(e, 1.0f / l) : ls.wi;
                            float sray[8] = {so.x, so.y, so.z, 0.0f, sd.x, sd.y, sd.z, l * (1.0f - 1e-4f)};
                            float st, su_, sv_;
                            uint32_t si, sp2;
                            o_trace_one(S->bvh, sray, 1, &st, &si, &sp2, &su_, &sv_);
                            if (!(st >= 0.0f)) {
                                rad[0] = rad[0] + c.x;
                                rad[1] = rad[1] + c.y;
                                rad[2] = rad[2] + c.z;
                            }
                        }
                    }
                }
                v3 wi;
                if (last || !brdf_sample_c(&s.m, s.ns, wo, ul, us1, us2, &wi) || !(dotv(s.ng, wi) > 0.0f)) break;
                float pdf = brdf_pdf_c(&s.m, s.ns, wo, wi);
                if (!(pdf > 0.0f)) break;
                v3 f = brdf_f(&s.m, s.ns, wo, wi);
                v3 nt = scl(mulv(thr, f), dotv(s.ns, wi) / pdf);
                int alive = nt.x > 0.0f || nt.y > 0.0f || nt.z > 0.0f;
                if (alive && J->rr_start > 0 && bounce + 1 >= J->rr_start) {
                    float q = fminf(fmaxf(fmaxf(nt.x, nt.y), nt.z), 0.95f);
                    if (urr >= q)
                        alive = 0;
                    else
                        nt = scl(nt, 1.0f / q);
                }
                if (!alive) break;
                o = add(s.p, scl(s.ng, s.eps));
                d = wi;
                thr = nt;
                last_pdf = pdf;
            }
        }
        float inv = 1.0f / (float)J->spp;
        J->out[3 * (size_t)px] = rad[0] * inv;
        J->out[3 * (size_t)px + 1] = rad[1] * inv;
        J->out[3 * (size_t)px + 2] = rad[2] * inv;
    }
    return NULL;
}

static int pt_render_impl(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5,
                          uint32_t n_inst, const uint32_t* idx, const float* nrm, const float* mat4s,
                          const uint32_t* mats16, const uint32_t* emissive, uint32_t n_emissive, const float* view16,
                          const float* proj16, const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp,
                          uint32_t max_depth, uint32_t rr_start, uint32_t seed, uint32_t frame, uint32_t seed_xor,
                          float* out_rgb, int threads, const o_gbuffer* gb) {
    oscene S;
    memset(&S, 0, sizeof(S));
    S.world = world;
    S.ntris = ntris;
    S.bvh = bvh;
    S.inst5 = inst5;
    S.n_inst = n_inst;
    S.idx = idx;
    S.nrm = nrm;
    S.mats16 = mats16;
    S.tri_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_inst + 1));
    S.tri_off[0] = 0;
    for (uint32_t i = 0; i < n_inst; ++i) S.tri_off[i + 1] = S.tri_off[i] + inst5[5 * i + 4] / 3;
    S.nmat = (float*)malloc(sizeof(float) * 9 * (n_inst ? n_inst : 1));
    for (uint32_t i = 0; i < n_inst; ++i) normal_matrix(mat4s + 16 * (size_t)inst5[5 * i], S.nmat + 9 * (size_t)i);
    build_emitters_c(&S, emissive, n_emissive);
    float cam[16];
    o_camera_matrix(view16, proj16, cam);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    pt_job jobs[256];
    uint32_t n = w * h, per = (n + (uint32_t)threads - 1) / (uint32_t)threads;
    for (int k = 0; k < threads; ++k) {
        pt_job j = {&S, cam, V(cam_pos[0], cam_pos[1], cam_pos[2]), w, h, spp ? spp : 1, max_depth ? max_depth : 1,
                    rr_start, seed, frame, seed_xor, out_rgb, 0, 0, gb};
        j.begin = (uint32_t)k * per < n ? (uint32_t)k * per : n;
        j.end = j.begin + per < n ? j.begin + per : n;
        jobs[k] = j;
        if (pthread_create(&th[k], NULL, pt_worker, &jobs[k]) != 0) {
            pt_worker(&jobs[k]);
            th[k] = 0;
        }
    }
    for (int k = 0; k < threads; ++k)
        if (th[k]) pthread_join(th[k], NULL);
    free(S.tri_off);
    free(S.nmat);
    free(S.e_tri);
    free(S.e_info);
    free(S.e_prob);
    free(S.e_alias);
    free(S.e_pdf);
    free(S.gid2e);
    return 0;
}

int o_pt_render(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5, uint32_t n_inst,
                const uint32_t* idx, const float* nrm, const float* mat4s, const uint32_t* mats16,
                const uint32_t* emissive, uint32_t n_emissive, const float* view16, const float* proj16,
                const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth, uint32_t rr_start,
                uint32_t seed, uint32_t frame, uint32_t seed_xor, float* out_rgb, int threads) {
    return pt_render_impl(world, ntris, bvh, inst5, n_inst, idx, nrm, mat4s, mats16, emissive, n_emissive, view16,
                          proj16, cam_pos, w, h, spp, max_depth, rr_start, seed, frame, seed_xor, out_rgb, threads,
                          NULL);
}

int o_pt_render_gbuffer(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5,
                        uint32_t n_inst, const uint32_t* idx, const float* nrm, const float* mat4s,
                        const uint32_t* mats16, const uint32_t* emissive, uint32_t n_emissive, const float* view16,
                        const float* proj16, const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp,
                        uint32_t max_depth, uint32_t rr_start, uint32_t seed, uint32_t frame, const o_gbuffer* gb,
                        float* out_rgb, int threads) {
    if (!gb || !gb->depth || !gb->normal || !gb->albedo || !gb->rma || !gb->flags) return -1;
    return pt_render_impl(world, ntris, bvh, inst5, n_inst, idx, nrm, mat4s, mats16, emissive, n_emissive, view16,
                          proj16, cam_pos, w, h, spp, max_depth, rr_start, seed, frame, 0u, out_rgb, threads, gb);
}

/* Primary surfaces of the camera rays, for building test G-buffers: per pixel
 * clip depth (proj_view * p, z / w), shading normal, albedo, roughness,
 * metallic, flags; depth 1 and zeros where the ray misses. */
void o_primary_surfaces(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5,
                        uint32_t n_inst, const uint32_t* idx, const float* nrm, const float* mat4s,
                        const uint32_t* mats16, const float* view16, const float* proj16, const float* proj_view16,
                        const float* cam_pos, uint32_t w, uint32_t h, float* depth, float* normal3, float* albedo3,
                        float* rm2, uint32_t* flags) {
    oscene S;
    memset(&S, 0, sizeof(S));
    S.world = world;
    S.ntris = ntris;
    S.bvh = bvh;
    S.inst5 = inst5;
    S.n_inst = n_inst;
    S.idx = idx;
    S.nrm = nrm;
    S.mats16 = mats16;
    S.tri_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_inst + 1));
    S.tri_off[0] = 0;
    for (uint32_t i = 0; i < n_inst; ++i) S.tri_off[i + 1] = S.tri_off[i] + inst5[5 * i + 4] / 3;
    S.nmat = (float*)malloc(sizeof(float) * 9 * (n_inst ? n_inst : 1));
    for (uint32_t i = 0; i < n_inst; ++i) normal_matrix(mat4s + 16 * (size_t)inst5[5 * i], S.nmat + 9 * (size_t)i);
    float cam[16];
    o_camera_matrix(view16, proj16, cam);
    const v3 pos = V(cam_pos[0], cam_pos[1], cam_pos[2]);
    for (uint32_t px = 0; px < w * h; ++px) {
        v3 d = cam_dir(cam, pos, px % w, px / w, w, h);
        float ray[8] = {pos.x, pos.y, pos.z, 0.0f, d.x, d.y, d.z, 1e30f};
        float t, hu, hv;
        uint32_t hi, hp;
        o_trace_one(bvh, ray, 0, &t, &hi, &hp, &hu, &hv);
        float* n3 = normal3 + 3 * (size_t)px;
        float* a3 = albedo3 + 3 * (size_t)px;
        if (!(t >= 0.0f)) {
            depth[px] = 1.0f;
            n3[0] = n3[1] = n3[2] = a3[0] = a3[1] = a3[2] = rm2[2 * px] = rm2[2 * px + 1] = 0.0f;
            flags[px] = 0;
            continue;
        }
        osurf s = surface(&S, d, hi, hp, hu, hv);
        double c[4];
        for (int r = 0; r < 4; ++r)
            c[r] = (double)proj_view16[r] * s.p.x + (double)proj_view16[4 + r] * s.p.y +
                   (double)proj_view16[8 + r] * s.p.z + (double)proj_view16[12 + r];
        depth[px] = (float)(c[2] / c[3]);
        n3[0] = s.ns.x;
        n3[1] = s.ns.y;
        n3[2] = s.ns.z;
        a3[0] = s.m.albedo.x;
        a3[1] = s.m.albedo.y;
        a3[2] = s.m.albedo.z;
        rm2[2 * px] = s.m.roughness;
        rm2[2 * px + 1] = s.m.metallic;
        flags[px] = s.m.flags;
    }
    free(S.tri_off);
    free(S.nmat);
}

/* ------------------------------------------------- reference compat --- */
/* BounceDiffuse (RayGen.rgen:123-213) with the payload of RayClosestHit.rchit:
 * 87-142 / RayMiss.rmiss:15-26, the term SNKPT_MODE_REFERENCE_COMPAT adds to
 * ReSTIR's direct light.  Random numbers: PCG32, state = the pass's ^
 * 0x434f4d50, stream = pixel; per outer ray r1, r2, then r1, r2 of each of its
 * two inner rays when it hit a non-emissive surface. */
#define COMPAT_PHASE 0x434f4d50ull

typedef struct {
    const oscene* S;
    const float* cam;
    v3 pos;
    uint32_t w, h, spp, seed, frame, n_dir;
    const float* dir_lights; /* n_dir x (colour rgb, direction towards the light) */
    float* out;
    uint32_t begin, end;
} compat_job;

static void ref_frame_c(v3 n, v3* nt, v3* nb) {
    if (fabsf(n.x) > fabsf(n.y))
        *nt = scl(V(n.z, 0.0f, -n.x), 1.0f / sqrtf(n.x * n.x + n.z * n.z));
    else
        *nt = scl(V(0.0f, -n.z, n.y), 1.0f / sqrtf(n.y * n.y + n.z * n.z));
    *nb = crossv(n, *nt);
}

static v3 ref_hemi_c(v3 n, v3 nt, v3 nb, float r1, float r2) {
    const float st = sqrtf(1.0f - r1 * r1), phi = 2.0f * REF_PI * r2;
    const float x = st * cosf(phi), z = st * sinf(phi);
    return V(x * nb.x + r1 * n.x + z * nt.x, x * nb.y + r1 * n.y + z * nt.y, x * nb.z + r1 * n.z + z * nt.z);
}

/* payload of one closest-hit ray; *surf = 1 for a non-emissive hit */
static v3 compat_payload(const compat_job* J, v3 o, v3 d, int* surf, v3* hp, v3* hn) {
    *surf = 0;
    float ray[8] = {o.x, o.y, o.z, 0.01f, d.x, d.y, d.z, 1000.0f};
    float t, hu, hv;
    uint32_t hi, hpr;
    o_trace_one(J->S->bvh, ray, 0, &t, &hi, &hpr, &hu, &hv);
    if (!(t >= 0.0f)) return V(0, 0, 0);
    osurf s = surface(J->S, d, hi, hpr, hu, hv);
    if (s.emissive) return s.le;
    *surf = 1;
    *hp = s.p;
    *hn = s.nraw;
    v3 c = V(0, 0, 0);
    const v3 v = scl(d, -1.0f);
    for (uint32_t k = 0; k < J->n_dir; ++k) {
        const float* L = J->dir_lights + 6 * (size_t)k;
        const v3 l = V(L[3], L[4], L[5]);
        v3 ck = scl(mulv(brdf_f(&s.m, s.nraw, v, l), V(L[0], L[1], L[2])), fmaxf(dotv(s.nraw, l), 0.0f));
        if (!(ck.x > 0.0f || ck.y > 0.0f || ck.z > 0.0f)) continue;
        const v3 so = add(s.p, scl(s.nraw, 0.01f));
        float sray[8] = {so.x, so.y, so.z, 0.001f, l.x, l.y, l.z, 10000.0f};
        float st, su, sv;
        uint32_t si, sp;
        o_trace_one(J->S->bvh, sray, 1, &st, &si, &sp, &su, &sv);
        if (!(st >= 0.0f)) c = add(c, ck);
    }
    return c;
}

static void* compat_worker(void* arg) {
    compat_job* J = (compat_job*)arg;
    for (uint32_t px = J->begin; px < J->end; ++px) {
        float acc[3] = {0, 0, 0};
        for (uint32_t k = 0; k < J->spp; ++k) {
            v3 d = cam_dir(J->cam, J->pos, px % J->w, px / J->w, J->w, J->h);
            float ray[8] = {J->pos.x, J->pos.y, J->pos.z, 0.0f, d.x, d.y, d.z, 1e30f};
            float t, hu, hv;
            uint32_t hi, hp;
            o_trace_one(J->S->bvh, ray, 0, &t, &hi, &hp, &hu, &hv);
            if (!(t >= 0.0f)) continue;
            osurf s = surface(J->S, d, hi, hp, hu, hv);
            if (s.emissive) continue;
            const v3 n = s.ns, alb = s.m.albedo;
            const v3 ro = add(s.p, scl(n, 0.01f + fmaxf(t * 0.001f, 0.01f)));
            v3 nt, nb;
            ref_frame_c(n, &nt, &nb);
            prng rng = rseed(o_pass_state(J->seed, J->frame, k) ^ COMPAT_PHASE, px);
            v3 sum = V(0, 0, 0);
            for (int r = 0; r < 8; ++r) {
                const float r1 = runi(&rng), r2 = runi(&rng);
                const v3 rd = ref_hemi_c(n, nt, nb, r1, r2);
                int surf;
                v3 hpos, hn;
                v3 col = compat_payload(J, ro, rd, &surf, &hpos, &hn);
                sum = add(sum, scl(mulv(alb, col), fmaxf(cosf(dotv(rd, n)), 0.0f) / 8.0f));
                if (!surf) continue;
                v3 it, ib;
                ref_frame_c(hn, &it, &ib);
                for (int b = 0; b < 2; ++b) {
                    const float q1 = runi(&rng), q2 = runi(&rng);
                    const v3 rdi = ref_hemi_c(hn, it, ib, q1, q2);
                    int s2;
                    v3 p2, n2;
                    v3 ci = compat_payload(J, hpos, rdi, &s2, &p2, &n2);
                    sum = add(sum, scl(mulv(alb, ci), 0.25f * fmaxf(cosf(dotv(rdi, hn)), 0.0f) / 8.0f));
                }
            }
            sum = scl(sum, 2.0f * REF_PI);
            acc[0] += sum.x;
            acc[1] += sum.y;
            acc[2] += sum.z;
        }
        const float inv = 1.0f / (float)J->spp;
        J->out[3 * (size_t)px] = acc[0] * inv;
        J->out[3 * (size_t)px + 1] = acc[1] * inv;
        J->out[3 * (size_t)px + 2] = acc[2] * inv;
    }
    return NULL;
}

int o_compat_bounce(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5, uint32_t n_inst,
                    const uint32_t* idx, const float* nrm, const float* mat4s, const uint32_t* mats16,
                    const float* dir_lights, uint32_t n_dir, const float* view16, const float* proj16,
                    const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp, uint32_t seed, uint32_t frame,
                    float* out_rgb, int threads) {
    oscene S;
    memset(&S, 0, sizeof(S));
    S.world = world;
    S.ntris = ntris;
    S.bvh = bvh;
    S.inst5 = inst5;
    S.n_inst = n_inst;
    S.idx = idx;
    S.nrm = nrm;
    S.mats16 = mats16;
    S.tri_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_inst + 1));
    S.tri_off[0] = 0;
    for (uint32_t i = 0; i < n_inst; ++i) S.tri_off[i + 1] = S.tri_off[i] + inst5[5 * i + 4] / 3;
    S.nmat = (float*)malloc(sizeof(float) * 9 * (n_inst ? n_inst : 1));
    for (uint32_t i = 0; i < n_inst; ++i) normal_matrix(mat4s + 16 * (size_t)inst5[5 * i], S.nmat + 9 * (size_t)i);
    float cam[16];
    o_camera_matrix(view16, proj16, cam);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    compat_job jobs[256];
    uint32_t n = w * h, per = (n + (uint32_t)threads - 1) / (uint32_t)threads;
    for (int k = 0; k < threads; ++k) {
        compat_job j = {&S, cam, V(cam_pos[0], cam_pos[1], cam_pos[2]), w, h, spp ? spp : 1, seed, frame,
                        dir_lights ? n_dir : 0u, dir_lights, out_rgb, 0, 0};
        j.begin = (uint32_t)k * per < n ? (uint32_t)k * per : n;
        j.end = j.begin + per < n ? j.begin + per : n;
        jobs[k] = j;
        if (pthread_create(&th[k], NULL, compat_worker, &jobs[k]) != 0) {
            compat_worker(&jobs[k]);
            th[k] = 0;
        }
    }
    for (int k = 0; k < threads; ++k)
        if (th[k]) pthread_join(th[k], NULL);
    free(S.tri_off);
    free(S.nmat);
    return 0;
}

/* ------------------------------ reference-compat direct light (Q10) --- */
/* The reference's own ReSTIR DI (RayGen.rgen:215-528) as SNKPT_MODE_REFERENCE_
 * COMPAT computes it: DoInitialWRS (32 uniform point-light candidates, weight
 * length(CalcPointlight) N_pl; 32 emissive candidates: instance uniform in the
 * emissive list, triangle uniform in it, folded uniform barycentrics, pdf
 * 1 / (N_inst N_tri area), weight length(ComputeTriangleRadiance) / pdf, which
 * has no emitter cosine; W = visibility w_sum / M / p_hat, the point light's ray
 * aimed at pos - biased_pos / dist as written at :345), WriteInitialReservoirs
 * (history with normal dot > 0.95, M capped at 20 M with w_sum scaled, the
 * fresh and history reservoirs combined with weights p_hat M W where p_hat
 * includes visibility, no MIS, M = M_fresh + M_hist; without a history the
 * fresh reservoir is discarded), ComputeAndWriteSpatialReservoirReuse (own +
 * k neighbours in a +-radius box redrawn until inside, skipped when the
 * previous frame's normal there has dot < 0.95; each combine also counts 1 into
 * M) and CalculateRadianceFromReservoirs plus the shadowed directional light.
 * Every UpdateReservoir adds 1 to M.  Random numbers: PCG32 (state =
 * o_pass_state(seed, frame, 0) ^ phase, stream = pixel); pass A (phase
 * 0x51314141): u_idx, u_wrs per point-light candidate, u_inst, u_tri, u_b1,
 * u_b2, u_wrs per emissive candidate, then u_pl, u_em per combine; pass B
 * (0x51314242): u_pl, u_em of the own combine, per neighbour (ux, uy) per try
 * (at most 16) and u_pl, u_em when it is combined.  The emissive point is
 * interpolated from the world-space triangle, whose area is used for the pdf
 * (the reference's det(M) x object-space area, equal for rigid transforms); the
 * history pixel is found by reprojection into the previous camera.  The camera
 * is static: frames 0..frames-1 are rendered and the last is returned
 * (emission of directly seen emitters + this direct light). */
#define QA_PHASE 0x51314141ull
#define QB_PHASE 0x51314242ull
#define QINV 0xFFFFFFFFu

typedef struct {
    float wsum_pl, w_pl;
    uint32_t pl, m_pl, inst, tri;
    float b1, b2, wsum_em, w_em;
    uint32_t m_em;
} qres;

typedef struct {
    int valid;
    v3 p, n, v, f0, biased;
    omat m;
    float t;
    v3 emit; /* radiance of a directly seen emitter */
} qvert;

typedef struct {
    const oscene* S;
    const float* pl; /* 12 floats per point light: colour, pad, position, pad, (c, l, e), pad */
    uint32_t n_pl;
    const uint32_t* em;
    uint32_t n_em;
    v3 dir_colour, dir_l;
    uint32_t w, h, k, temporal, have_prev;
    float radius;
    uint64_t state;
    float pv[16];
    const qvert* G;
    const qvert* Gp;
    qres* hist;
    qres* tmp;
    float* out;
    uint32_t begin, end;
    int pass;
} qjob;

static qres q_empty(void) {
    qres r = {0.0f, 0.0f, QINV, 0u, QINV, QINV, 0.0f, 0.0f, 0.0f, 0.0f, 0u};
    return r;
}

static float vlen_c(v3 a) { return sqrtf(dotv(a, a)); }

static v3 nrmz_q(v3 a) {
    float l = sqrtf(dotv(a, a));
    return l > 0.0f ? V(a.x / l, a.y / l, a.z / l) : a;
}

/* CookTorranceBRDF (LightBuffers.glsl:103-124): (kd / PI + spec) n.l, kd without albedo */
static v3 ct_ref_c(v3 lpos, v3 sp, v3 v, v3 f0, v3 n, float rough, float metal) {
    const v3 l = nrmz_q(sub(lpos, sp));
    const v3 hh = nrmz_q(add(v, l));
    const float c1 = fminf(fmaxf(1.0f - fmaxf(dotv(hh, v), 0.0f), 0.0f), 1.0f), c2 = c1 * c1;
    const float fw = (c2 * c2) * c1;
    const v3 F = V(f0.x + (1.0f - f0.x) * fw, f0.y + (1.0f - f0.y) * fw, f0.z + (1.0f - f0.z) * fw);
    const float a = rough * rough;
    const float D = ggx_d_ref(fmaxf(dotv(n, hh), 0.0f), a * a);
    const float r = rough + 1.0f, k = (r * r) / 8.0f;
    const float nv = fmaxf(dotv(n, v), 0.0f), nl = fmaxf(dotv(n, l), 0.0f);
    const float G = g1s(nv, k) * g1s(nl, k);
    const float den = 4.0f * nv * nl + 0.0001f;
    const float km = 1.0f - metal;
    const v3 spec = scl(F, D * G / den);
    const v3 kd = V((1.0f - F.x) * km, (1.0f - F.y) * km, (1.0f - F.z) * km);
    return scl(add(scl(kd, 1.0f / REF_PI), spec), nl);
}

/* CalcPointlight (LightBuffers.glsl:145-157) */
static v3 calc_point_c(const qjob* J, uint32_t k, const qvert* x) {
    const float* L = J->pl + 12 * (size_t)k;
    const v3 lp = V(L[4], L[5], L[6]);
    const float dist = vlen_c(sub(lp, x->p));
    const float atten = (L[8] + L[9] * dist) + L[10] * (dist * dist);
    const v3 brdf = ct_ref_c(lp, x->p, x->v, x->f0, x->n, x->m.roughness, x->m.metallic);
    return scl(mulv(mulv(brdf, x->m.albedo), V(L[0], L[1], L[2])), 1.0f / atten);
}

/* CalcDirectionalLight (LightBuffers.glsl:127-143) */
static v3 calc_dir_c(const qjob* J, const qvert* x) {
    const v3 l = J->dir_l, v = x->v, n = x->n, f0 = x->f0;
    const v3 hh = nrmz_q(add(v, l));
    const float c1 = fminf(fmaxf(1.0f - fmaxf(dotv(hh, v), 0.0f), 0.0f), 1.0f), c2 = c1 * c1;
    const float fw = (c2 * c2) * c1;
    const v3 F = V(f0.x + (1.0f - f0.x) * fw, f0.y + (1.0f - f0.y) * fw, f0.z + (1.0f - f0.z) * fw);
    const float a = x->m.roughness * x->m.roughness;
    const float D = ggx_d_ref(fmaxf(dotv(n, hh), 0.0f), a * a);
    const float r = x->m.roughness + 1.0f, k = (r * r) / 8.0f;
    const float nv = fmaxf(dotv(n, v), 0.0f), nl = fmaxf(dotv(n, l), 0.0f);
    const float G = g1s(nv, k) * g1s(nl, k);
    const float den = 4.0f * nv * nl + 0.0001f;
    const float km = 1.0f - x->m.metallic;
    const v3 spec = scl(F, D * G / den);
    const v3 kd = V((1.0f - F.x) * km * x->m.albedo.x, (1.0f - F.y) * km * x->m.albedo.y,
                    (1.0f - F.z) * km * x->m.albedo.z);
    return mulv(scl(add(scl(kd, 1.0f / REF_PI), spec), nl), J->dir_colour);
}

typedef struct {
    v3 p, n, le;
    float area;
} qtri;

static qtri tri_point_c(const oscene* S, uint32_t inst, uint32_t tri, float b1, float b2) {
    const uint32_t* in = S->inst5 + 5 * inst;
    const float* w = S->world + 12 * (size_t)(S->tri_off[inst] + tri);
    const v3 v0 = V(w[0], w[1], w[2]), v1 = V(w[3], w[4], w[5]), v2 = V(w[6], w[7], w[8]);
    const float b3 = 1.0f - b1 - b2;
    qtri q;
    q.p = add(add(scl(v0, b1), scl(v1, b2)), scl(v2, b3));
    const uint32_t* ix = S->idx + in[1] + 3 * tri;
    const float* n0 = S->nrm + 3 * (size_t)(ix[0] + in[2]);
    const float* n1 = S->nrm + 3 * (size_t)(ix[1] + in[2]);
    const float* n2 = S->nrm + 3 * (size_t)(ix[2] + in[2]);
    const v3 nl = add(add(scl(V(n0[0], n0[1], n0[2]), b1), scl(V(n1[0], n1[1], n1[2]), b2)),
                      scl(V(n2[0], n2[1], n2[2]), b3));
    const float* nm = S->nmat + 9 * (size_t)inst;
    q.n = nrmz_q(V((nm[0] * nl.x + nm[1] * nl.y) + nm[2] * nl.z, (nm[3] * nl.x + nm[4] * nl.y) + nm[5] * nl.z,
                   (nm[6] * nl.x + nm[7] * nl.y) + nm[8] * nl.z));
    q.area = 0.5f * vlen_c(crossv(sub(v1, v0), sub(v1, v2)));
    const uint32_t* mt = S->mats16 + 16 * (size_t)in[3];
    q.le = V(u2f(mt[8]), u2f(mt[9]), u2f(mt[10]));
    return q;
}

/* ShadowTest (:101-121) */
static float qshadow_c(const oscene* S, v3 ro, v3 rd, float maxd) {
    float ray[8] = {ro.x, ro.y, ro.z, 0.001f, rd.x, rd.y, rd.z, maxd};
    float t, u, v;
    uint32_t i, p;
    o_trace_one(S->bvh, ray, 1, &t, &i, &p, &u, &v);
    return t >= 0.0f ? 0.0f : 1.0f;
}

/* ComputeTriangleRadiance (:258-282) */
static v3 tri_radiance_c(const oscene* S, uint32_t inst, uint32_t tri, float b1, float b2, const qvert* x,
                         int shadow) {
    const qtri q = tri_point_c(S, inst, tri, b1, b2);
    const v3 brdf = ct_ref_c(q.p, x->p, x->v, x->f0, x->n, x->m.roughness, x->m.metallic);
    const v3 d = sub(x->p, q.p);
    v3 rad = scl(mulv(mulv(q.le, x->m.albedo), brdf), 1.0f / fabsf(dotv(d, d)));
    if (shadow) {
        const v3 tgt = add(q.p, scl(q.n, 0.01f));
        const float pd = vlen_c(sub(x->biased, tgt));
        rad = scl(rad, qshadow_c(S, x->biased, scl(sub(tgt, x->biased), 1.0f / pd), pd));
    }
    return rad;
}

static int q_update_c(float* wsum, uint32_t* M, float weight, float u) {
    *M += 1u;
    *wsum += weight;
    return *wsum > 0.0f && u < weight / *wsum;
}

/* CombinePixelReservoir (:361-386) */
static void q_combine_c(const qjob* J, const qvert* x, qres* r, const qres* nb, float* ph_pl, float* ph_em, prng* rng) {
    const float u_pl = runi(rng), u_em = runi(rng);
    if (nb->pl != QINV) {
        const float* L = J->pl + 12 * (size_t)nb->pl;
        const v3 lp = V(L[4], L[5], L[6]);
        const float ld = vlen_c(sub(lp, x->biased));
        const float ph = vlen_c(calc_point_c(J, nb->pl, x)) *
                         qshadow_c(J->S, x->biased, scl(sub(lp, x->biased), 1.0f / ld), ld - 0.001f);
        const float weight = ph * (float)nb->m_pl * nb->w_pl;
        if (q_update_c(&r->wsum_pl, &r->m_pl, weight, u_pl)) {
            r->pl = nb->pl;
            *ph_pl = ph;
        }
    }
    if (nb->inst != QINV) {
        const float ph = vlen_c(tri_radiance_c(J->S, nb->inst, nb->tri, nb->b1, nb->b2, x, 1));
        const float weight = ph * nb->w_em * (float)nb->m_em;
        if (q_update_c(&r->wsum_em, &r->m_em, weight, u_em)) {
            r->inst = nb->inst;
            r->tri = nb->tri;
            r->b1 = nb->b1;
            r->b2 = nb->b2;
            *ph_em = ph;
        }
    }
}

static void q_pass_a(const qjob* J, uint32_t i) {
    const qvert* x = &J->G[i];
    qres out = q_empty();
    if (x->valid) {
        prng rng = rseed(J->state ^ QA_PHASE, i);
        qres r = q_empty();
        float ph_pl = 0.0f, ph_em = 0.0f;
        for (uint32_t c = 0; c < 32u; ++c) {
            const float u = runi(&rng), uw = runi(&rng);
            uint32_t k = 0;
            float ph = 0.0f, weight = 0.0f;
            if (J->n_pl > 0) {
                k = (uint32_t)(u * (float)J->n_pl);
                if (k > J->n_pl - 1u) k = J->n_pl - 1u;
                ph = vlen_c(calc_point_c(J, k, x));
                weight = ph / (1.0f / (float)J->n_pl);
            }
            if (q_update_c(&r.wsum_pl, &r.m_pl, weight, uw)) {
                r.pl = k;
                ph_pl = ph;
            }
        }
        for (uint32_t c = 0; c < 32u; ++c) {
            const float ui = runi(&rng), ut = runi(&rng), ub1 = runi(&rng), ub2 = runi(&rng), uw = runi(&rng);
            if (J->n_em == 0) {
                r.m_em += 1u;
                continue;
            }
            uint32_t ei = (uint32_t)(ui * (float)J->n_em);
            if (ei > J->n_em - 1u) ei = J->n_em - 1u;
            const uint32_t inst = J->em[ei];
            const uint32_t ntri = J->S->inst5[5 * inst + 4] / 3u;
            if (ntri == 0) {
                r.m_em += 1u;
                continue;
            }
            uint32_t tri = (uint32_t)(ut * (float)ntri);
            if (tri > ntri - 1u) tri = ntri - 1u;
            float b1 = ub1, b2 = ub2;
            if (b1 + b2 >= 1.0f) {
                b1 = 1.0f - b1;
                b2 = 1.0f - b2;
            }
            const qtri q = tri_point_c(J->S, inst, tri, b1, b2);
            const float ph = vlen_c(tri_radiance_c(J->S, inst, tri, b1, b2, x, 0));
            const float pdf = 1.0f / (((float)J->n_em * (float)ntri) * q.area);
            const float weight = ph / pdf;
            if (q_update_c(&r.wsum_em, &r.m_em, weight, uw)) {
                r.inst = inst;
                r.tri = tri;
                r.b1 = b1;
                r.b2 = b2;
                ph_em = ph;
            }
        }
        if (r.pl != QINV) {
            const float* L = J->pl + 12 * (size_t)r.pl;
            const v3 lp = V(L[4], L[5], L[6]);
            const float d = vlen_c(sub(lp, x->biased));
            r.w_pl = qshadow_c(J->S, x->biased, sub(lp, scl(x->biased, 1.0f / d)), d) * ((r.wsum_pl / (float)r.m_pl) / ph_pl);
        }
        if (r.inst != QINV) {
            const qtri q = tri_point_c(J->S, r.inst, r.tri, r.b1, r.b2);
            const v3 tgt = add(q.p, scl(q.n, 0.01f));
            const float pd = vlen_c(sub(x->biased, tgt));
            r.w_em = qshadow_c(J->S, x->biased, scl(sub(tgt, x->biased), 1.0f / pd), pd) *
                     ((r.wsum_em / (float)r.m_em) / ph_em);
        }
        int q = -1;
        if (J->temporal) {
            float cl[4];
            for (int k = 0; k < 4; ++k)
                cl[k] = ((J->pv[k] * x->p.x + J->pv[4 + k] * x->p.y) + J->pv[8 + k] * x->p.z) + J->pv[12 + k];
            if (cl[3] > 1e-6f) {
                const float fx = (cl[0] / cl[3] * 0.5f + 0.5f) * (float)J->w;
                const float fy = (cl[1] / cl[3] * 0.5f + 0.5f) * (float)J->h;
                if (fx > 0.0f && fy > 0.0f && fx < (float)J->w && fy < (float)J->h)
                    q = (int)fy * (int)J->w + (int)fx;
            }
        }
        if (q >= 0 && (!J->Gp[q].valid || !(dotv(J->Gp[q].n, x->n) > 0.95f))) q = -1;
        if (q >= 0) {
            qres hh = J->hist[q];
            if ((float)hh.m_pl > 20.0f * (float)r.m_pl) {
                hh.wsum_pl *= 20.0f * (float)r.m_pl / (float)hh.m_pl;
                hh.m_pl = 20u * r.m_pl;
            }
            if ((float)hh.m_em > 20.0f * (float)r.m_em) {
                hh.wsum_em *= 20.0f * (float)r.m_em / (float)hh.m_em;
                hh.m_em = 20u * r.m_em;
            }
            float tp_pl = 0.0f, tp_em = 0.0f;
            q_combine_c(J, x, &out, &r, &tp_pl, &tp_em, &rng);
            q_combine_c(J, x, &out, &hh, &tp_pl, &tp_em, &rng);
            out.m_em = hh.m_em + r.m_em;
            out.m_pl = hh.m_pl + r.m_pl;
            out.w_em = tp_em > 0.0f ? (out.wsum_em / (float)out.m_em) / tp_em : 0.0f;
            out.w_pl = tp_pl > 0.0f ? (out.wsum_pl / (float)out.m_pl) / tp_pl : 0.0f;
        }
    }
    J->tmp[i] = out;
}

static void q_pass_b(const qjob* J, uint32_t i) {
    const qvert* x = &J->G[i];
    qres r = q_empty();
    float* o = J->out + 3 * (size_t)i;
    o[0] = x->emit.x;
    o[1] = x->emit.y;
    o[2] = x->emit.z;
    if (x->valid) {
        prng rng = rseed(J->state ^ QB_PHASE, i);
        float ph_pl = 0.0f, ph_em = 0.0f;
        const qres own = J->tmp[i];
        q_combine_c(J, x, &r, &own, &ph_pl, &ph_em, &rng);
        r.m_em += own.m_em;
        r.m_pl += own.m_pl;
        const float cx0 = (float)(i % J->w) / (float)J->w, cy0 = (float)(i / J->w) / (float)J->h;
        for (uint32_t s = 0; s < J->k; ++s) {
            float cx = -1.0f, cy = -1.0f;
            for (int t = 0; t < 16; ++t) {
                const float ux = runi(&rng), uy = runi(&rng);
                const float ox = (ux - 0.5f) * 2.0f * J->radius / (float)J->w;
                const float oy = (uy - 0.5f) * 2.0f * J->radius / (float)J->h;
                const float tx = cx0 + ox, ty = cy0 + oy;
                if (tx >= 0.0f && ty >= 0.0f && tx <= 1.0f && ty <= 1.0f) {
                    cx = tx;
                    cy = ty;
                    break;
                }
            }
            if (cx < 0.0f || !J->have_prev) continue;
            uint32_t nx = (uint32_t)((float)J->w * cx), ny = (uint32_t)((float)J->h * cy);
            if (nx > J->w - 1u) nx = J->w - 1u;
            if (ny > J->h - 1u) ny = J->h - 1u;
            const uint32_t q = ny * J->w + nx;
            if (!J->Gp[q].valid || dotv(J->Gp[q].n, x->n) < 0.95f) continue;
            const qres nb = J->tmp[q];
            q_combine_c(J, x, &r, &nb, &ph_pl, &ph_em, &rng);
            r.m_em += nb.m_em;
            r.m_pl += nb.m_pl;
        }
        r.w_em = ph_em > 0.0f ? r.wsum_em / ph_em / (float)r.m_em : 0.0f;
        r.w_pl = ph_pl > 0.0f ? r.wsum_pl / ph_pl / (float)r.m_pl : 0.0f;
        v3 light = calc_dir_c(J, x);
        if (light.x > 0.0f || light.y > 0.0f || light.z > 0.0f)
            light = scl(light, qshadow_c(J->S, x->biased, J->dir_l, 1000.0f));
        if (J->n_pl > 0 && r.pl != QINV) {
            const float* L = J->pl + 12 * (size_t)r.pl;
            const v3 lp = V(L[4], L[5], L[6]);
            const float ld = vlen_c(sub(lp, x->biased));
            const v3 c = scl(calc_point_c(J, r.pl, x),
                             qshadow_c(J->S, x->biased, scl(sub(lp, x->biased), 1.0f / ld), ld));
            light = add(light, scl(c, r.w_pl));
        }
        if (J->n_em > 0 && r.inst != QINV) {
            const v3 c = tri_radiance_c(J->S, r.inst, r.tri, r.b1, r.b2, x, 1);
            light = add(light, scl(c, r.w_em));
        }
        o[0] = o[0] + light.x;
        o[1] = o[1] + light.y;
        o[2] = o[2] + light.z;
    }
    J->hist[i] = r;
}

static void* q_worker(void* arg) {
    const qjob* J = (const qjob*)arg;
    for (uint32_t i = J->begin; i < J->end; ++i) {
        if (J->pass == 0)
            q_pass_a(J, i);
        else
            q_pass_b(J, i);
    }
    return NULL;
}

static void q_run(qjob* base, int threads, int pass) {
    pthread_t th[256];
    qjob jobs[256];
    uint32_t n = base->w * base->h, per = (n + (uint32_t)threads - 1) / (uint32_t)threads;
    for (int k = 0; k < threads; ++k) {
        jobs[k] = *base;
        jobs[k].pass = pass;
        jobs[k].begin = (uint32_t)k * per < n ? (uint32_t)k * per : n;
        jobs[k].end = jobs[k].begin + per < n ? jobs[k].begin + per : n;
        if (pthread_create(&th[k], NULL, q_worker, &jobs[k]) != 0) {
            q_worker(&jobs[k]);
            th[k] = 0;
        }
    }
    for (int k = 0; k < threads; ++k)
        if (th[k]) pthread_join(th[k], NULL);
}

int o_compat_direct(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5, uint32_t n_inst,
                    const uint32_t* idx, const float* nrm, const float* mat4s, const uint32_t* mats16,
                    const uint32_t* emissive, uint32_t n_emissive, const uint8_t* light_block, size_t block_bytes,
                    const float* view16, const float* proj16, const float* proj_view16, const float* cam_pos,
                    uint32_t w, uint32_t h, uint32_t frames, uint32_t temporal, uint32_t spatial, float radius,
                    uint32_t seed, float* out_rgb, int threads) {
    oscene S;
    memset(&S, 0, sizeof(S));
    S.world = world;
    S.ntris = ntris;
    S.bvh = bvh;
    S.inst5 = inst5;
    S.n_inst = n_inst;
    S.idx = idx;
    S.nrm = nrm;
    S.mats16 = mats16;
    S.tri_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_inst + 1));
    S.tri_off[0] = 0;
    for (uint32_t i = 0; i < n_inst; ++i) S.tri_off[i + 1] = S.tri_off[i] + inst5[5 * i + 4] / 3;
    S.nmat = (float*)malloc(sizeof(float) * 9 * (n_inst ? n_inst : 1));
    for (uint32_t i = 0; i < n_inst; ++i) normal_matrix(mat4s + 16 * (size_t)inst5[5 * i], S.nmat + 9 * (size_t)i);
    /* the light block as the shaders read it (LightBuffers.glsl:1-63) */
    uint32_t n_pl = 0;
    float* pl = NULL;
    qjob J;
    memset(&J, 0, sizeof(J));
    if (light_block && block_bytes >= 112) {
        float f[8];
        memcpy(f, light_block, 32);
        J.dir_colour = V(f[0], f[1], f[2]);
        J.dir_l = V(f[4], f[5], f[6]);
        memcpy(&n_pl, light_block + 96, 4);
        pl = (float*)calloc(12 * (size_t)(n_pl ? n_pl : 1), sizeof(float));
        for (uint32_t i = 0; i < n_pl && 112 + 48 * (size_t)(i + 1) <= block_bytes; ++i) {
            float e[12];
            memcpy(e, light_block + 112 + 48 * (size_t)i, 48);
            float* d = pl + 12 * (size_t)i;
            d[0] = e[0], d[1] = e[1], d[2] = e[2];
            d[4] = e[4], d[5] = e[5], d[6] = e[6];
            d[8] = e[10], d[9] = e[9], d[10] = e[8]; /* (constant, linear, exp) */
        }
    }
    const uint32_t n = w * h;
    qvert* G = (qvert*)calloc(n, sizeof(qvert));
    qvert* Gp = (qvert*)calloc(n, sizeof(qvert));
    qres* hist = (qres*)malloc(sizeof(qres) * n);
    qres* tmp = (qres*)malloc(sizeof(qres) * n);
    for (uint32_t i = 0; i < n; ++i) hist[i] = tmp[i] = q_empty();
    float cam[16];
    o_camera_matrix(view16, proj16, cam);
    const v3 pos = V(cam_pos[0], cam_pos[1], cam_pos[2]);
    /* the primary surfaces (the same every frame: the camera does not move) */
    for (uint32_t px = 0; px < n; ++px) {
        qvert* x = &G[px];
        v3 d = cam_dir(cam, pos, px % w, px / w, w, h);
        float ray[8] = {pos.x, pos.y, pos.z, 0.0f, d.x, d.y, d.z, 1e30f};
        float t, hu, hv;
        uint32_t hi, hp;
        o_trace_one(bvh, ray, 0, &t, &hi, &hp, &hu, &hv);
        if (!(t >= 0.0f)) continue;
        osurf s = surface(&S, d, hi, hp, hu, hv);
        if (s.emissive) {
            x->emit = s.le;
            continue;
        }
        x->valid = 1;
        x->p = s.p;
        x->n = s.ns;
        x->v = scl(d, -1.0f);
        x->m = s.m;
        x->t = t;
        x->f0 = V(0.04f + (s.m.albedo.x - 0.04f) * s.m.metallic, 0.04f + (s.m.albedo.y - 0.04f) * s.m.metallic,
                  0.04f + (s.m.albedo.z - 0.04f) * s.m.metallic);
        x->biased = add(s.p, scl(s.ns, 0.01f + fmaxf(t * 0.001f, 0.01f)));
    }
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    J.S = &S;
    J.pl = pl;
    J.n_pl = n_pl;
    J.em = emissive;
    J.n_em = n_emissive;
    J.w = w;
    J.h = h;
    J.k = spatial;
    J.radius = radius;
    memcpy(J.pv, proj_view16, sizeof(J.pv));
    J.G = G;
    J.Gp = Gp;
    J.hist = hist;
    J.tmp = tmp;
    J.out = out_rgb;
    for (uint32_t f = 0; f < frames; ++f) {
        J.state = o_pass_state(seed, f, 0);
        J.have_prev = f > 0;
        J.temporal = f > 0 && temporal != 0;
        q_run(&J, threads, 0);
        q_run(&J, threads, 1);
        memcpy(Gp, G, sizeof(qvert) * n);
    }
    free(G);
    free(Gp);
    free(hist);
    free(tmp);
    free(pl);
    free(S.tri_off);
    free(S.nmat);
    return 0;
}
