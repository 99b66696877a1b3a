// ReSTIR DI for the primary vertex: candidate, temporal and spatial passes.
//
// Reference: RayGen.rgen:215-528 — DoInitialWRS (32 point-light + 32 emissive
// candidates, :296-357), WriteInitialReservoirs (temporal reuse with
// reprojection, normal test and M_CAP 20, :411-468), ComputeAndWriteSpatial-
// ReservoirReuse (num_spatial_resamples neighbours in a 30 px box, :470-506),
// CalculateRadianceFromReservoirs (:508-528).  The reference combine is biased
// (no MIS / 1/Z weights, Q10) and discards a fresh reservoir without history;
// this build keeps its structure (candidate -> temporal -> spatial -> shade,
// reservoirs ping-ponged between passes as in Q9) but uses the unbiased
// combination of generalized resampled importance sampling (Bitterli et al.
// 2020; Lin et al. 2022) with balance-heuristic resampling weights:
//   candidates  RESTIR_CANDIDATES samples from the power-proportional emitter
//               distribution, RIS weight p_hat / p_src, weighted reservoir sampling
//   temporal    reproject the primary hit into the previous camera, reject on
//               normal / distance mismatch, cap the history M at m_cap * M_cur
//   spatial     k neighbours uniformly in a disk of `restir_radius` pixels
//   weights     input l contributes m_l(y_l) p_hat(y_l) W_l with
//               m_l(y) = M_l p_hat_l(y) / sum_j M_j p_hat_j(y), each p_hat_j
//               evaluated on input j's own surface; W = w_sum / p_hat(y)
// Target p_hat = luminance(f Le cos_s cos_l / d^2), unshadowed; visibility is
// traced once for the final sample.  The rest of the path (bounces >= 1)
// continues with the plain path tracer; BRDF-sampled emission seen from the
// primary vertex gets MIS weight 0 because direct light there is ReSTIR's.
#include <cstring>

#include "render_modes.h"

namespace snkpt {

struct RestirState {
    uint32_t cap = 0;
    DeviceBuffer g[2][5];  // per frame parity: p|eps (0: no surface), ns|rough, ng|metal, albedo|flags, wo|t
    DeviceBuffer r_init, r_tmp, r_fin;  // float4 (emitter bits, b1, b2, W)
    DeviceBuffer m_init, m_tmp, m_fin;  // M
    DeviceBuffer a_init, a_tmp, a_fin;  // analytic-light reservoirs: float4 (light bits, W, M, 0)
    DeviceBuffer q_tmp, q_hist;  // reference-compatible reservoirs, 3 float4 per pixel (pass A out, pass B out)
    int cur = 0;
    bool have_prev = false;
    float prev_pv[16] = {0};
    float prev_pos[3] = {0, 0, 0};
};
void restir_state_free(RestirState* r) { delete r; }

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxSpatial = 8;

struct GView {
    float4 *p, *n, *g, *a, *w;
};

struct GSurf {
    V3 p, ns, ng, wo;
    Mat m;
    float t, eps;  // eps: ray bias along ng
    bool valid;
};

__device__ __forceinline__ GSurf load_g(const GView& G, uint32_t i) {
    GSurf s;
    float4 p = G.p[i], n = G.n[i], g = G.g[i], a = G.a[i], w = G.w[i];
    s.p = V3{p.x, p.y, p.z};
    s.valid = p.w > 0.0f;
    s.eps = p.w;
    s.ns = V3{n.x, n.y, n.z};
    s.m.roughness = n.w;
    s.ng = V3{g.x, g.y, g.z};
    s.m.metallic = g.w;
    s.m.albedo = V3{a.x, a.y, a.z};
    s.m.flags = __float_as_uint(a.w);
    s.wo = V3{w.x, w.y, w.z};
    s.t = w.w;
    return s;
}

struct Res {
    uint32_t e;
    float b1, b2, W, M, wsum;
};

// Unshadowed contribution of emitter sample (e, b1, b2) at surface g, from the
// emitter's record a | b | c | info (4 float4).
__device__ __forceinline__ float target_rec(const GSurf& g, const float4& a, const float4& b, const float4& c,
                                            const float4& info, float b1, float b2, V3* c_out, V3* wi_out,
                                            float* dist_out) {
    V3 A{a.x, a.y, a.z}, B{b.x, b.y, b.z}, C{c.x, c.y, c.z};
    float b0 = 1.0f - b1 - b2;
    V3 q = A * b0 + B * b1 + C * b2;
    V3 d = q - g.p;
    float d2 = dot(d, d);
    if (!(d2 > 0.0f)) return 0.0f;
    float dist = sqrtf(d2);
    V3 wi = d * (1.0f / dist);
    float cs = dot(g.ns, wi);
    if (cs <= 0.0f || dot(g.ng, wi) <= 0.0f) return 0.0f;
    V3 nl = cross(B - A, C - A);
    float nll = sqrtf(dot(nl, nl));
    if (!(nll > 0.0f)) return 0.0f;
    float cosl = fabsf(dot(nl, wi)) / nll;
    V3 f = brdf_eval(g.m, g.ns, g.wo, wi);
    V3 cc = mul(f, V3{info.x, info.y, info.z}) * (cs * cosl / d2);
    if (c_out) *c_out = cc;
    if (wi_out) *wi_out = wi;
    if (dist_out) *dist_out = dist;
    return luminance(cc);
}
__device__ __forceinline__ float target(const SceneDev& S, const GSurf& g, uint32_t e, float b1, float b2, V3* c_out,
                                        V3* wi_out, float* dist_out) {
    const float4* r = S.emit_rec + 4 * (size_t)e;
    return target_rec(g, r[0], r[1], r[2], r[3], b1, b2, c_out, wi_out, dist_out);
}

// Unshadowed contribution of analytic light k (point, spot or directional) at
// surface g: the target of the analytic-light reservoir (the reference's
// point-light reservoir, RayGen.rgen:300-310, with this build's light kinds).
__device__ __forceinline__ float atarget(const SceneDev& S, const GSurf& g, uint32_t k, V3* c_out, V3* wi_out,
                                         float* dist_out) {
    const LightSample ls = analytic_light(S, k, g.p);
    if (!ls.ok) return 0.0f;
    const float cs = dot(g.ns, ls.wi);
    if (cs <= 0.0f || dot(g.ng, ls.wi) <= 0.0f) return 0.0f;
    const V3 cc = mul(brdf_eval(g.m, g.ns, g.wo, ls.wi), ls.le) * cs;
    if (c_out) *c_out = cc;
    if (wi_out) *wi_out = ls.wi;
    if (dist_out) *dist_out = ls.dist;
    return luminance(cc);
}

struct ResA {
    uint32_t k;
    float W, M, wsum;
};
__device__ __forceinline__ ResA load_resa(const float4* a, uint32_t i) {
    const float4 v = a[i];
    return ResA{__float_as_uint(v.x), v.y, v.z, 0.0f};
}
__device__ __forceinline__ void store_resa(float4* a, uint32_t i, const ResA& r) {
    a[i] = make_float4(__uint_as_float(r.k), r.W, r.M, 0.0f);
}
__device__ __forceinline__ bool resa_update(ResA& r, uint32_t k, float w, float u) {
    if (!(w > 0.0f)) return false;
    r.wsum += w;
    if (u * r.wsum < w) {
        r.k = k;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool res_update(Res& r, uint32_t e, float b1, float b2, float w, float u) {
    if (!(w > 0.0f)) return false;
    r.wsum += w;
    if (u * r.wsum < w) {
        r.e = e;
        r.b1 = b1;
        r.b2 = b2;
        return true;
    }
    return false;
}

__device__ __forceinline__ void store_res(float4* r, float* m, uint32_t i, const Res& x) {
    r[i] = make_float4(__uint_as_float(x.e), x.b1, x.b2, x.W);
    m[i] = x.M;
}
__device__ __forceinline__ Res load_res(const float4* r, const float* m, uint32_t i) {
    float4 a = r[i];
    Res x;
    x.e = __float_as_uint(a.x);
    x.b1 = a.y;
    x.b2 = a.z;
    x.W = a.w;
    x.M = m[i];
    x.wsum = 0.0f;
    return x;
}

constexpr uint64_t kPhaseCand = 0x52455354ull, kPhaseTemp = 0x54454d50ull, kPhaseSpat = 0x53504154ull,
                   kPhaseCont = 0x434f4e54ull;

// Primary vertex: G-buffer, directly seen emission, candidate RIS, continuation ray.
// GB: the vertex is read from the caller's rasterized G-buffer (gb, cam, w)
// instead of the traced camera ray (in, hit).
template <bool GB>
__global__ __launch_bounds__(kBlock) void k_primary(SceneDev S, GView G, Queue in, HitArrays hit, GbufDev gb,
                                                    CamDev cam, uint32_t w, Queue out,
                                                    PathState ps, float4* __restrict__ rad, float4* __restrict__ r_out,
                                                    float* __restrict__ m_out, float4* __restrict__ a_out,
                                                    uint32_t* __restrict__ counts,
                                                    uint32_t n, uint32_t n_cand, uint32_t max_depth,
                                                    uint32_t rr_start, uint64_t initstate,
                                                    unsigned long long* __restrict__ stats) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    bool want_ext = false;
    V3 ext_o, ext_d;
    float4 thr4 = make_float4(0, 0, 0, 0);
    Rng rc{0, 0};
    if (valid) {
        Res r{0xFFFFFFFFu, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        ResA ra{0xFFFFFFFFu, 0.0f, 0.0f, 0.0f};
        V3 dir;
        float t;
        Surf s;
        bool have;
        if (GB) {
            have = gbuffer_surface(gb, cam, i, w, n / w, s, dir, t);
        } else {
            const float4 dq = in.dir[i];
            dir = V3{dq.x, dq.y, dq.z};
            t = hit.t[i];
            have = t >= 0.0f;
            if (have) s = surface_at(S, dir, hit.inst[i], hit.prim[i], hit.u[i], hit.v[i]);
        }
        float4 gp = make_float4(0, 0, 0, 0);
        if (have) {
            if (s.emissive) {
                float4 o = rad[i];
                rad[i] = make_float4(o.x + s.le.x, o.y + s.le.y, o.z + s.le.z, o.w);
            } else {
                V3 wo = dir * -1.0f;
                gp = make_float4(s.p.x, s.p.y, s.p.z, s.eps);  // w > 0: valid
                G.n[i] = make_float4(s.ns.x, s.ns.y, s.ns.z, s.m.roughness);
                G.g[i] = make_float4(s.ng.x, s.ng.y, s.ng.z, s.m.metallic);
                G.a[i] = make_float4(s.m.albedo.x, s.m.albedo.y, s.m.albedo.z, __uint_as_float(s.m.flags));
                G.w[i] = make_float4(wo.x, wo.y, wo.z, t);
                GSurf g;
                g.p = s.p;
                g.ns = s.ns;
                g.ng = s.ng;
                g.wo = wo;
                g.m = s.m;
                g.eps = s.eps;
                Rng rr = rng_seed(initstate ^ kPhaseCand, i);
                float ph = 0.0f;
                if (S.n_emit > 0 && n_cand > 0) {
                    // Candidate c + 1's uniforms and the record its uniform draw lands on
                    // are fetched while candidate c is weighed (the same draws in the
                    // same order), so the record's L2 latency overlaps the work.
                    auto draw = [&](float& a0, float& a1, float& a2, float& a3, float& aw, uint32_t& ei) {
                        a0 = rr.uniform();
                        a1 = rr.uniform();
                        a2 = rr.uniform();
                        a3 = rr.uniform();
                        aw = rr.uniform();
                        ei = (uint32_t)(a0 * (float)S.n_emit);
                        if (ei >= S.n_emit) ei = S.n_emit - 1;
                    };
                    float u0, u1, u2, u3, uw;
                    uint32_t e;
                    draw(u0, u1, u2, u3, uw, e);
                    const float4* rp = S.emit_rec + 4 * (size_t)e;
                    float4 ra = rp[0], rb = rp[1], rc = rp[2], ri = rp[3];
                    for (uint32_t c = 0; c < n_cand; ++c) {
                        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f, vw = 0.0f;
                        uint32_t en = 0;
                        float4 na = ra, nb = rb, nc = rc, ni = ri;
                        if (c + 1 < n_cand) {
                            draw(v0, v1, v2, v3, vw, en);
                            const float4* np = S.emit_rec + 4 * (size_t)en;
                            na = np[0];
                            nb = np[1];
                            nc = np[2];
                            ni = np[3];
                        }
                        if (u1 >= ra.w) {  // the alias's record
                            e = __float_as_uint(rb.w);
                            const float4* ap = S.emit_rec + 4 * (size_t)e;
                            ra = ap[0];
                            rb = ap[1];
                            rc = ap[2];
                            ri = ap[3];
                        }
                        float su = sqrtf(u2);
                        float b1 = su * (1.0f - u3), b2 = su * u3;
                        float p = target_rec(g, ra, rb, rc, ri, b1, b2, nullptr, nullptr, nullptr);
                        float src = rc.w / ri.w;  // area measure
                        if (res_update(r, e, b1, b2, p / src, uw)) ph = p;
                        u0 = v0;
                        u1 = v1;
                        u2 = v2;
                        u3 = v3;
                        uw = vw;
                        e = en;
                        ra = na;
                        rb = nb;
                        rc = nc;
                        ri = ni;
                    }
                    (void)u0;
                }
                r.M = (float)n_cand;
                r.W = ph > 0.0f ? r.wsum / (r.M * ph) : 0.0f;
                // analytic lights: n_cand candidates drawn uniformly (source pdf 1 / n_alights)
                if (S.n_alights > 0 && n_cand > 0) {
                    float pa = 0.0f;
                    for (uint32_t c = 0; c < n_cand; ++c) {
                        const float u = rr.uniform(), uw = rr.uniform();
                        uint32_t k = (uint32_t)(u * (float)S.n_alights);
                        if (k >= S.n_alights) k = S.n_alights - 1;
                        const float p = atarget(S, g, k, nullptr, nullptr, nullptr);
                        if (resa_update(ra, k, p * (float)S.n_alights, uw)) pa = p;
                    }
                    ra.M = (float)n_cand;
                    ra.W = pa > 0.0f ? ra.wsum / (ra.M * pa) : 0.0f;
                }
                // continuation into bounce 1 (direct light at this vertex is ReSTIR's)
                if (max_depth > 1) {
                    rc = rng_seed(initstate ^ kPhaseCont, i);
                    float ul = rc.uniform(), us1 = rc.uniform(), us2 = rc.uniform(), urr = rc.uniform();
                    V3 wi;
                    if (brdf_sample(s.m, s.ns, wo, ul, us1, us2, wi) && dot(s.ng, wi) > 0.0f) {
                        float pdf = brdf_pdf(s.m, s.ns, wo, wi);
                        if (pdf > 0.0f) {
                            V3 nt = brdf_eval(s.m, s.ns, wo, wi) * (dot(s.ns, wi) / pdf);
                            bool alive = nt.x > 0.0f || nt.y > 0.0f || nt.z > 0.0f;
                            if (alive && rr_start > 0 && 1u >= rr_start) {
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
                                thr4 = make_float4(nt.x, nt.y, nt.z, -1.0f);
                            }
                        }
                    }
                }
            }
        }
        G.p[i] = gp;
        store_res(r_out, m_out, i, r);
        store_resa(a_out, i, ra);
    }
    __shared__ uint32_t lds_app[kBlock / 64 + 1];
    uint32_t slot = block_append<kBlock>(want_ext, &counts[1], lds_app);
    if (want_ext) {
        out.org[slot] = make_float4(ext_o.x, ext_o.y, ext_o.z, 0.0f);
        out.dir[slot] = make_float4(ext_d.x, ext_d.y, ext_d.z, 1e30f);
        out.path[slot] = i;
        ps.thr[i] = thr4;
        ps.rng[i] = make_uint2((uint32_t)rc.state, (uint32_t)(rc.state >> 32));
        ps.bounce[i] = 1;
    }
}

struct PrevCam {
    float pv[16];  // column-major proj_view of the previous frame
    V3 pos;
};

// Temporal reuse (WriteInitialReservoirs, RayGen.rgen:411-468) with 1/Z weights,
// for pixel i; true when the pixel's history was found and combined.
__device__ __forceinline__ bool temporal_pixel(const SceneDev& S, const GView& G, const GView& Gp, uint32_t i,
                                               const float4* __restrict__ r_in, const float* __restrict__ m_in,
                                               const float4* __restrict__ r_hist, const float* __restrict__ m_hist,
                                               float4* __restrict__ r_out, float* __restrict__ m_out, uint32_t w,
                                               uint32_t h, const PrevCam& pc, float m_cap, uint64_t initstate,
                                               const uint16_t* __restrict__ motion, const float4* __restrict__ a_in,
                                               const float4* __restrict__ a_hist, float4* __restrict__ a_out) {
    Res rc = load_res(r_in, m_in, i);
    GSurf g = load_g(G, i);
    const bool alights = S.n_alights > 0;
    if (!g.valid) {
        store_res(r_out, m_out, i, rc);
        if (alights) a_out[i] = a_in[i];
        return false;
    }
    int q = -1;
    if (motion) {
        // history texcoord = texcoord + motion (RayGen.rgen:424-428)
        const uint32_t mb = *(const uint32_t*)(motion + 2 * (size_t)i);
        const float hx = ((float)(i % w) + 0.5f) / (float)w + half_bits((uint16_t)(mb & 0xFFFFu));
        const float hy = ((float)(i / w) + 0.5f) / (float)h + half_bits((uint16_t)(mb >> 16));
        if (hx > 0.0f && hy > 0.0f && hx < 1.0f && hy < 1.0f) {
            const uint32_t qx = min((uint32_t)(hx * (float)w), w - 1), qy = min((uint32_t)(hy * (float)h), h - 1);
            q = (int)(qy * w + qx);
        }
    } else {
        // reproject into the previous frame
        float cl[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            cl[r] = ((pc.pv[r] * g.p.x + pc.pv[4 + r] * g.p.y) + pc.pv[8 + r] * g.p.z) + pc.pv[12 + r];
        if (cl[3] > 1e-6f) {
            float fx = (cl[0] / cl[3] * 0.5f + 0.5f) * (float)w, fy = (cl[1] / cl[3] * 0.5f + 0.5f) * (float)h;
            if (fx >= 0.0f && fy >= 0.0f && fx < (float)w && fy < (float)h) q = (int)fy * (int)w + (int)fx;
        }
    }
    GSurf gp;
    gp.valid = false;
    if (q >= 0) {
        gp = load_g(Gp, (uint32_t)q);
        // motion vectors also follow moving geometry: then only the normal test applies
        // (RayGen.rgen:436); reprojection of a static scene also checks the distance
        float dprev = sqrtf(dot(g.p - pc.pos, g.p - pc.pos));
        if (!gp.valid || dot(gp.ns, g.ns) < 0.9f || (!motion && fabsf(dprev - gp.t) > 0.05f * gp.t))
            gp.valid = false;
    }
    if (!gp.valid) {
        store_res(r_out, m_out, i, rc);
        if (alights) a_out[i] = a_in[i];
        return false;
    }
    Res rh = load_res(r_hist, m_hist, (uint32_t)q);
    if (rh.M > m_cap * rc.M) rh.M = m_cap * rc.M;
    Rng rr = rng_seed(initstate ^ kPhaseTemp, i);
    Res s{0xFFFFFFFFu, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float ph = 0.0f;
    if (rc.e != 0xFFFFFFFFu && rc.W > 0.0f) {
        float pc = target(S, g, rc.e, rc.b1, rc.b2, nullptr, nullptr, nullptr);
        float pp = target(S, gp, rc.e, rc.b1, rc.b2, nullptr, nullptr, nullptr);
        float den = rc.M * pc + rh.M * pp;
        float m = den > 0.0f ? rc.M * pc / den : 0.0f;
        if (res_update(s, rc.e, rc.b1, rc.b2, m * pc * rc.W, rr.uniform())) ph = pc;
    }
    if (rh.e != 0xFFFFFFFFu && rh.W > 0.0f && rh.e < S.n_emit) {
        float pc = target(S, g, rh.e, rh.b1, rh.b2, nullptr, nullptr, nullptr);
        float pp = target(S, gp, rh.e, rh.b1, rh.b2, nullptr, nullptr, nullptr);
        float den = rc.M * pc + rh.M * pp;
        float m = den > 0.0f ? rh.M * pp / den : 0.0f;
        if (res_update(s, rh.e, rh.b1, rh.b2, m * pc * rh.W, rr.uniform())) ph = pc;
    }
    s.M = rc.M + rh.M;
    s.W = ph > 0.0f ? s.wsum / ph : 0.0f;
    store_res(r_out, m_out, i, s);
    if (alights) {
        // the analytic-light reservoirs, combined the same way
        const ResA ac = load_resa(a_in, i);
        ResA ah = load_resa(a_hist, (uint32_t)q);
        if (ah.M > m_cap * ac.M) ah.M = m_cap * ac.M;
        ResA sa{0xFFFFFFFFu, 0.0f, 0.0f, 0.0f};
        float pa = 0.0f;
        if (ac.k != 0xFFFFFFFFu && ac.W > 0.0f) {
            const float p_c = atarget(S, g, ac.k, nullptr, nullptr, nullptr);
            const float p_p = atarget(S, gp, ac.k, nullptr, nullptr, nullptr);
            const float den = ac.M * p_c + ah.M * p_p;
            const float m = den > 0.0f ? ac.M * p_c / den : 0.0f;
            if (resa_update(sa, ac.k, m * p_c * ac.W, rr.uniform())) pa = p_c;
        }
        if (ah.k < S.n_alights && ah.W > 0.0f) {
            const float p_c = atarget(S, g, ah.k, nullptr, nullptr, nullptr);
            const float p_p = atarget(S, gp, ah.k, nullptr, nullptr, nullptr);
            const float den = ac.M * p_c + ah.M * p_p;
            const float m = den > 0.0f ? ah.M * p_p / den : 0.0f;
            if (resa_update(sa, ah.k, m * p_c * ah.W, rr.uniform())) pa = p_c;
        }
        sa.M = ac.M + ah.M;
        sa.W = pa > 0.0f ? sa.wsum / pa : 0.0f;
        store_resa(a_out, i, sa);
    }
    return true;
}

// stats[5] += pixels whose history was reused (one atomic per block)
__global__ __launch_bounds__(kBlock) void k_temporal(SceneDev S, GView G, GView Gp, const float4* __restrict__ r_in,
                                                     const float* __restrict__ m_in, const float4* __restrict__ r_hist,
                                                     const float* __restrict__ m_hist, float4* __restrict__ r_out,
                                                     float* __restrict__ m_out, uint32_t w, uint32_t h, PrevCam pc,
                                                     float m_cap, uint64_t initstate,
                                                     const uint16_t* __restrict__ motion,
                                                     unsigned long long* __restrict__ stats,
                                                     const float4* __restrict__ a_in,
                                                     const float4* __restrict__ a_hist, float4* __restrict__ a_out) {
    __shared__ uint32_t red;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) red = 0u;
    __syncthreads();
    const bool reused = i < w * h && temporal_pixel(S, G, Gp, i, r_in, m_in, r_hist, m_hist, r_out, m_out, w, h, pc,
                                                    m_cap, initstate, motion, a_in, a_hist, a_out);
    const uint64_t b = __ballot(reused);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&red, (uint32_t)__popcll(b));
    __syncthreads();
    if (threadIdx.x == 0 && stats && red) atomicAdd(&stats[5], (unsigned long long)red);
}

// Spatial reuse (ComputeAndWriteSpatialReservoirReuse, RayGen.rgen:470-506) with 1/Z weights.
// A block shades one kSpatTile x kSpatTile pixel tile, so the neighbours its
// pixels read (within `radius` pixels) overlap in L2 far more than for a run of
// one row.  Each input surface is read once: the balance-heuristic sums run with
// the surfaces in the outer loop and the inputs' samples in the inner one (the
// same additions in the same order as sample-major loops, so the same bits).
constexpr uint32_t kSpatTile = 16;
static_assert(kSpatTile * kSpatTile == kBlock, "one pixel per thread");

// The hardware deals consecutive blocks to the 8 XCDs in turn, each with its own
// L2.  Tile t of the row-major tile order runs as block (t % per) * 8 + t / per,
// so each XCD walks a contiguous run of tiles and the neighbour rows and columns
// that tiles near each other read stay in that XCD's L2.
constexpr uint32_t kXcds = 8;
inline uint32_t xcd_tile_blocks(uint32_t w, uint32_t h) {
    const uint32_t n = ((w + kSpatTile - 1) / kSpatTile) * ((h + kSpatTile - 1) / kSpatTile);
    return ((n + kXcds - 1) / kXcds) * kXcds;
}
__device__ __forceinline__ bool xcd_tile_origin(uint32_t w, uint32_t h, uint32_t& x0, uint32_t& y0) {
    const uint32_t nbx = (w + kSpatTile - 1) / kSpatTile, n = nbx * ((h + kSpatTile - 1) / kSpatTile);
    const uint32_t per = (n + kXcds - 1) / kXcds;
    const uint32_t t = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    x0 = (t % nbx) * kSpatTile;
    y0 = (t / nbx) * kSpatTile;
    return t < n;
}

__device__ __forceinline__ bool neighbour_ok(const GView& G, uint32_t q, const GSurf& g) {
    // the reuse test needs the validity, shading normal and distance only (3 of 5 words)
    const float4 p = G.p[q], n = G.n[q], wv = G.w[q];
    const V3 ns{n.x, n.y, n.z};
    return p.w > 0.0f && !(dot(ns, g.ns) < 0.9f) && !(fabsf(wv.w - g.t) > 0.1f * g.t);
}

template <uint32_t KMAX>
__global__ __launch_bounds__(kBlock) void k_spatial(SceneDev S, GView G, const float4* __restrict__ r_in,
                                                    const float* __restrict__ m_in, float4* __restrict__ r_out,
                                                    float* __restrict__ m_out, uint32_t w, uint32_t h, uint32_t k,
                                                    float radius, uint64_t initstate) {
    uint32_t x0, y0;
    if (!xcd_tile_origin(w, h, x0, y0)) return;
    const uint32_t x = x0 + (threadIdx.x % kSpatTile);
    const uint32_t y = y0 + (threadIdx.x / kSpatTile);
    if (x >= w || y >= h) return;
    const uint32_t i = y * w + x;
    Res r0 = load_res(r_in, m_in, i);
    GSurf g = load_g(G, i);
    if (!g.valid || k == 0) {
        store_res(r_out, m_out, i, r0);
        return;
    }
    Rng rr = rng_seed(initstate ^ kPhaseSpat, i);
    // inputs: this pixel's reservoir and up to k accepted neighbours
    uint32_t px[KMAX + 1];
    px[0] = i;
    int ni = 1;
    if (k > KMAX) k = KMAX;
    for (uint32_t j = 0; j < k; ++j) {
        float ur = rr.uniform(), ua = rr.uniform();
        float rad = radius * sqrtf(ur), ang = 2.0f * kPi * ua;
        int nx = (int)x + (int)lrintf(rad * cosf(ang)), ny = (int)y + (int)lrintf(rad * sinf(ang));
        if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)h || (nx == (int)x && ny == (int)y)) continue;
        uint32_t q = (uint32_t)ny * w + (uint32_t)nx;
        if (!neighbour_ok(G, q, g)) continue;
        px[ni++] = q;
    }
    // the inputs' samples and their uniforms (drawn in input order)
    float Ms[KMAX + 1], uw[KMAX + 1], Wl[KMAX + 1], b1[KMAX + 1], b2[KMAX + 1];
    uint32_t el[KMAX + 1];
    float den[KMAX + 1], p_self[KMAX + 1], p_l[KMAX + 1];
#pragma unroll
    for (int l = 0; l < (int)KMAX + 1; ++l) {
        if (l >= ni) break;
        const Res rl = load_res(r_in, m_in, px[l]);
        Ms[l] = rl.M;
        uw[l] = rr.uniform();
        el[l] = (rl.e == 0xFFFFFFFFu || !(rl.W > 0.0f)) ? 0xFFFFFFFFu : rl.e;
        b1[l] = rl.b1;
        b2[l] = rl.b2;
        Wl[l] = rl.W;
        den[l] = p_self[l] = p_l[l] = 0.0f;
    }
    // balance heuristic over the inputs' targets, each on its own surface
#pragma unroll
    for (int j = 0; j < (int)KMAX + 1; ++j) {
        if (j >= ni) break;
        const GSurf gj = j == 0 ? g : load_g(G, px[j]);
#pragma unroll
        for (int l = 0; l < (int)KMAX + 1; ++l) {
            if (l >= ni) break;
            if (el[l] == 0xFFFFFFFFu) continue;
            const float pj = target(S, gj, el[l], b1[l], b2[l], nullptr, nullptr, nullptr);
            if (j == 0) p_self[l] = pj;
            if (j == l) p_l[l] = pj;
            den[l] += Ms[j] * pj;
        }
    }
    Res s{0xFFFFFFFFu, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float ph = 0.0f;
    s.M = 0.0f;
#pragma unroll
    for (int l = 0; l < (int)KMAX + 1; ++l) {
        if (l >= ni) break;
        s.M += Ms[l];
        if (el[l] == 0xFFFFFFFFu) continue;
        const float m = den[l] > 0.0f ? Ms[l] * p_l[l] / den[l] : 0.0f;
        if (res_update(s, el[l], b1[l], b2[l], m * p_self[l] * Wl[l], uw[l])) ph = p_self[l];
    }
    s.W = ph > 0.0f ? s.wsum / ph : 0.0f;
    store_res(r_out, m_out, i, s);
}

// Spatial reuse of the analytic-light reservoirs (the reference combines its
// point-light reservoirs in the same pass, RayGen.rgen:470-506): the same
// neighbour tests and balance-heuristic weights as k_spatial, neighbours drawn
// from a stream of their own.
constexpr uint64_t kPhaseSpatA = 0x53504141ull;
__global__ __launch_bounds__(kBlock) void k_spatial_analytic(SceneDev S, GView G, const float4* __restrict__ a_in,
                                                             float4* __restrict__ a_out, uint32_t w, uint32_t h,
                                                             uint32_t k, float radius, uint64_t initstate) {
    uint32_t x0, y0;
    if (!xcd_tile_origin(w, h, x0, y0)) return;
    const uint32_t x = x0 + (threadIdx.x % kSpatTile);
    const uint32_t y = y0 + (threadIdx.x / kSpatTile);
    if (x >= w || y >= h) return;
    const uint32_t i = y * w + x;
    const GSurf g = load_g(G, i);
    if (!g.valid || k == 0) {
        a_out[i] = a_in[i];
        return;
    }
    Rng rr = rng_seed(initstate ^ kPhaseSpatA, i);
    uint32_t px[kMaxSpatial + 1];
    px[0] = i;
    int ni = 1;
    if (k > kMaxSpatial) k = kMaxSpatial;
    for (uint32_t j = 0; j < k; ++j) {
        const float ur = rr.uniform(), ua = rr.uniform();
        const float rad = radius * sqrtf(ur), ang = 2.0f * kPi * ua;
        const int nx = (int)x + (int)lrintf(rad * cosf(ang)), ny = (int)y + (int)lrintf(rad * sinf(ang));
        if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)h || (nx == (int)x && ny == (int)y)) continue;
        const uint32_t q = (uint32_t)ny * w + (uint32_t)nx;
        if (!neighbour_ok(G, q, g)) continue;
        px[ni++] = q;
    }
    float Ms[kMaxSpatial + 1], uw[kMaxSpatial + 1], Wl[kMaxSpatial + 1];
    uint32_t kl[kMaxSpatial + 1];
    float den[kMaxSpatial + 1], p_self[kMaxSpatial + 1], p_l[kMaxSpatial + 1];
#pragma unroll
    for (int l = 0; l < (int)kMaxSpatial + 1; ++l) {
        if (l >= ni) break;
        const ResA rl = load_resa(a_in, px[l]);
        Ms[l] = rl.M;
        uw[l] = rr.uniform();
        kl[l] = (rl.k < S.n_alights && rl.W > 0.0f) ? rl.k : 0xFFFFFFFFu;
        Wl[l] = rl.W;
        den[l] = p_self[l] = p_l[l] = 0.0f;
    }
#pragma unroll
    for (int j = 0; j < (int)kMaxSpatial + 1; ++j) {
        if (j >= ni) break;
        const GSurf gj = j == 0 ? g : load_g(G, px[j]);
#pragma unroll
        for (int l = 0; l < (int)kMaxSpatial + 1; ++l) {
            if (l >= ni) break;
            if (kl[l] == 0xFFFFFFFFu) continue;
            const float pj = atarget(S, gj, kl[l], nullptr, nullptr, nullptr);
            if (j == 0) p_self[l] = pj;
            if (j == l) p_l[l] = pj;
            den[l] += Ms[j] * pj;
        }
    }
    ResA s{0xFFFFFFFFu, 0.0f, 0.0f, 0.0f};
    float ph = 0.0f;
#pragma unroll
    for (int l = 0; l < (int)kMaxSpatial + 1; ++l) {
        if (l >= ni) break;
        s.M += Ms[l];
        if (kl[l] == 0xFFFFFFFFu) continue;
        const float m = den[l] > 0.0f ? Ms[l] * p_l[l] / den[l] : 0.0f;
        if (resa_update(s, kl[l], m * p_self[l] * Wl[l], uw[l])) ph = p_self[l];
    }
    s.W = ph > 0.0f ? s.wsum / ph : 0.0f;
    store_resa(a_out, i, s);
}

// CalculateRadianceFromReservoirs (RayGen.rgen:508-528): f Le G W, visibility by shadow ray.
__global__ __launch_bounds__(kBlock) void k_rshade(SceneDev S, GView G, const float4* __restrict__ r_in,
                                                   const float* __restrict__ m_in, ShadowQueue sq,
                                                   uint32_t* __restrict__ counts, uint32_t n,
                                                   unsigned long long* __restrict__ stats) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool want = false;
    V3 c, wi, o;
    float dist = 0.0f, tmax = 0.0f;
    if (i < n) {
        GSurf g = load_g(G, i);
        if (g.valid) {
            Res r = load_res(r_in, m_in, i);
            if (r.e != 0xFFFFFFFFu && r.W > 0.0f) {
                float p = target(S, g, r.e, r.b1, r.b2, &c, &wi, &dist);
                if (p > 0.0f) {
                    c = c * r.W;
                    shadow_segment(g.p, g.ng, g.eps, wi, dist, o, wi, tmax);
                    want = true;
                }
            }
        }
    }
    __shared__ uint32_t lds_app[kBlock / 64 + 1];
    uint32_t slot = block_append<kBlock>(want, &counts[2], lds_app);
    if (want) {
        sq.org[slot] = make_float4(o.x, o.y, o.z, 0.0f);
        sq.dir[slot] = make_float4(wi.x, wi.y, wi.z, tmax);
        sq.contrib[slot] = make_float4(c.x, c.y, c.z, __uint_as_float(i));
    }
}

// The analytic-light reservoir's sample at the primary hit (the reference's
// point-light half of CalculateRadianceFromReservoirs, RayGen.rgen:513-519):
// f Le cos W, visibility by shadow ray.
__global__ __launch_bounds__(kBlock) void k_rshade_analytic(SceneDev S, GView G, const float4* __restrict__ a_in,
                                                            ShadowQueue sq, uint32_t* __restrict__ counts,
                                                            uint32_t n) {
    __shared__ uint32_t lds_app[kBlock / 64 + 1];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool want = false;
    V3 c, wi, o;
    float tmax = 0.0f, dist = 0.0f;
    if (i < n) {
        const GSurf g = load_g(G, i);
        if (g.valid) {
            const ResA r = load_resa(a_in, i);
            if (r.k < S.n_alights && r.W > 0.0f && atarget(S, g, r.k, &c, &wi, &dist) > 0.0f) {
                c = c * r.W;
                shadow_segment(g.p, g.ng, g.eps, wi, dist, o, wi, tmax);
                want = true;
            }
        }
    }
    uint32_t slot = block_append<kBlock>(want, &counts[2], lds_app);
    if (want) {
        sq.org[slot] = make_float4(o.x, o.y, o.z, 0.0f);
        sq.dir[slot] = make_float4(wi.x, wi.y, wi.z, tmax);
        sq.contrib[slot] = make_float4(c.x, c.y, c.z, __uint_as_float(i));
    }
}

// shadow rays of k_rshade_analytic
__global__ void k_restir_stats2(uint32_t* c, unsigned long long* stats) {
    if (threadIdx.x == 0) {
        if (stats) stats[2] += c[2];
        c[2] = 0;
    }
}

// extension rays queued by k_primary and shadow rays of k_rshade
__global__ void k_restir_stats(const uint32_t* c, unsigned long long* stats) {
    if (threadIdx.x == 0) {
        stats[1] += c[1];
        stats[2] += c[2];
    }
}

__global__ void k_counts3(uint32_t* c, uint32_t a, uint32_t b, uint32_t d) {
    if (threadIdx.x == 0) {
        c[0] = a;
        c[1] = b;
        c[2] = d;
    }
}

// ---- reference-compatibility indirect term (SNKPT_MODE_REFERENCE_COMPAT) ----
//
// BounceDiffuse (RayGen.rgen:139-213) as the reference computes it, with PCG32
// numbers (stream = pixel, state = the pass's ^ kPhaseCompat) in place of the
// wall-clock gold noise: from biased_pos = p + n (0.01 + max(0.001 |cam - p|,
// 0.01)) (:552) eight rays rd = UniformSampleHemisphere(r1, r2) in the frame of
// CreateCoordinateSystem(n) (:123-136), tmin 0.01, tmax 1000, each adding
// albedo * payload * max(cos(dot(rd, n)), 0) / 8; a non-emissive hit at t sends
// two more from ro + rd t about the hit's own normal, adding
// albedo * payload * 0.25 * max(cos(dot(rd', n')), 0) / 8; the sum is scaled by
// 2 PI.  The payload (RayClosestHit.rchit:87-142, RayMiss.rmiss:15-26): 0 on a
// miss, the albedo of an emissive hit, else the directional lights through the
// hit's BRDF (its normal as interpolated, v = -rd) times a shadow test from
// pos + n 0.01 (tmin 0.001, tmax 10000).  Numbers per pixel, in order: r1, r2
// of each outer ray, followed by r1, r2 of its two inner rays when it spawns them.
constexpr uint64_t kPhaseCompat = 0x434f4d50ull;
constexpr int kCompatBlock = 128;

__device__ __forceinline__ void ref_frame(V3 n, V3& nt, V3& nb) {
    if (fabsf(n.x) > fabsf(n.y))
        nt = V3{n.z, 0.0f, -n.x} * (1.0f / sqrtf(n.x * n.x + n.z * n.z));
    else
        nt = V3{0.0f, -n.z, n.y} * (1.0f / sqrtf(n.y * n.y + n.z * n.z));
    nb = cross(n, nt);
}

__device__ __forceinline__ V3 ref_hemisphere(V3 n, V3 nt, V3 nb, float r1, float r2) {
    const float st = sqrtf(1.0f - r1 * r1), phi = 2.0f * kRefPi * r2;
    const float x = st * cosf(phi), z = st * sinf(phi);
    return V3{x * nb.x + r1 * n.x + z * nt.x, x * nb.y + r1 * n.y + z * nt.y, x * nb.z + r1 * n.z + z * nt.z};
}

struct CompatHit {
    bool surface;   // a non-emissive hit (it spawns inner rays)
    V3 colour;      // payload.colour
    V3 pos, nraw;   // hit position and normal (RayClosestHit.rchit:109-111)
};

template <int STACK>
__device__ CompatHit compat_trace(const SceneDev& S, V3 o, V3 d, uint2* stk, int stride, uint32_t& n_sh) {
    CompatHit h;
    h.surface = false;
    h.colour = V3{0.0f, 0.0f, 0.0f};
    uint32_t nv = 0, tv = 0;
    const Ray r{o.x, o.y, o.z, 0.01f, d.x, d.y, d.z, 1000.0f};
    const TraceResult res = traverse<false, false, STACK>(S.nodes, S.leaf, r, stk, stride, nv, tv);
    if (res.rec == 0xFFFFFFFFu) return h;
    const float4 r2 = S.leaf[(size_t)res.rec * kTriWords + 2];
    float hu, hv;
    hit_uv(S.leaf, res.rec, r, hu, hv);
    const Surf s = surface_at(S, d, __float_as_uint(r2.y), __float_as_uint(r2.z), hu, hv);
    if (s.emissive) {
        h.colour = s.le;
        return h;
    }
    h.surface = true;
    h.pos = s.p;
    h.nraw = s.nraw;
    const V3 v = d * -1.0f;
    for (uint32_t k = 0; k < S.n_alights; ++k) {
        const float4* L = S.alights + 4 * (size_t)k;
        const float4 w0 = L[0], w1 = L[1];
        if ((uint32_t)w0.w != 0u) continue;  // directional lights only
        const V3 l{w1.x, w1.y, w1.z};
        const V3 c = mul(brdf_eval(s.m, s.nraw, v, l), V3{w0.x, w0.y, w0.z}) * fmaxf(dot(s.nraw, l), 0.0f);
        if (!(c.x > 0.0f || c.y > 0.0f || c.z > 0.0f)) continue;
        const V3 so = s.p + s.nraw * 0.01f;
        const Ray sr{so.x, so.y, so.z, 0.001f, l.x, l.y, l.z, 10000.0f};
        ++n_sh;
        if (traverse<true, false, STACK>(S.nodes, S.leaf, sr, stk, stride, nv, tv).rec == 0xFFFFFFFFu)
            h.colour = h.colour + c;
    }
    return h;
}

template <int STACK>
__global__ __launch_bounds__(kCompatBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_compat_bounce(SceneDev S, GView G, float4* __restrict__ rad,
                                                                uint32_t n, uint64_t initstate,
                                                                unsigned long long* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kCompatBlock];
    __shared__ uint32_t red[2];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) red[0] = red[1] = 0u;
    __syncthreads();
    uint32_t n_ext = 0, n_sh = 0;
    if (i < n) {
        const GSurf g = load_g(G, i);
        if (g.valid && !(g.m.flags & 1u)) {
            const V3 nrm = g.ns, alb = g.m.albedo;
            const V3 ro = g.p + nrm * (0.01f + fmaxf(g.t * 0.001f, 0.01f));
            V3 nt, nb;
            ref_frame(nrm, nt, nb);
            Rng rng = rng_seed(initstate ^ kPhaseCompat, i);
            V3 sum{0.0f, 0.0f, 0.0f};
            for (int k = 0; k < 8; ++k) {
                const float r1 = rng.uniform(), r2 = rng.uniform();
                const V3 rd = ref_hemisphere(nrm, nt, nb, r1, r2);
                ++n_ext;
                const CompatHit h = compat_trace<STACK>(S, ro, rd, &stk[threadIdx.x], kCompatBlock, n_sh);
                sum = sum + mul(alb, h.colour) * (fmaxf(cosf(dot(rd, nrm)), 0.0f) / 8.0f);
                if (!h.surface) continue;
                V3 it, ib;
                ref_frame(h.nraw, it, ib);
                for (int b = 0; b < 2; ++b) {
                    const float q1 = rng.uniform(), q2 = rng.uniform();
                    const V3 rdi = ref_hemisphere(h.nraw, it, ib, q1, q2);
                    ++n_ext;
                    const CompatHit hi = compat_trace<STACK>(S, h.pos, rdi, &stk[threadIdx.x], kCompatBlock, n_sh);
                    sum = sum + mul(alb, hi.colour) * (0.25f * fmaxf(cosf(dot(rdi, h.nraw)), 0.0f) / 8.0f);
                }
            }
            sum = sum * (2.0f * kRefPi);
            const float4 o = rad[i];
            rad[i] = make_float4(o.x + sum.x, o.y + sum.y, o.z + sum.z, o.w);
        }
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
    if (threadIdx.x == 0 && stats && (red[0] | red[1])) {
        atomicAdd(&stats[1], (unsigned long long)red[0]);
        atomicAdd(&stats[2], (unsigned long long)red[1]);
    }
}


// ---- the reference's own ReSTIR DI (Q10) for SNKPT_MODE_REFERENCE_COMPAT ----
//
// DoInitialWRS, WriteInitialReservoirs, ComputeAndWriteSpatialReservoirReuse and
// CalculateRadianceFromReservoirs (RayGen.rgen:215-528) as the reference computes
// them, for the pass's point lights and emissive triangles:
//   candidates  32 point lights drawn uniformly (weight length(CalcPointlight) N_pl)
//               and 32 emissive triangles (instance uniform in the emissive list,
//               triangle uniform in it, folded uniform barycentrics; pdf
//               1 / (N_inst N_tri area)), weight length(ComputeTriangleRadiance)
//               / pdf: emission albedo CookTorranceBRDF / d^2, no emitter cosine;
//               W = visibility w_sum / M / p_hat (the point light's ray aimed at
//               pos - biased_pos / dist, as written at :345)
//   temporal    the history (normal test > 0.95, M capped at 20 M with w_sum
//               scaled) and the fresh reservoir are combined with weights
//               p_hat (visibility included) M W, no MIS; M = M_fresh + M_hist;
//               without a history the fresh reservoir is discarded (an empty one
//               is written)
//   spatial     own reservoir + k neighbours in a +-radius pixel box (redrawn until
//               inside the image), skipped when the previous frame's normal there
//               has dot < 0.95; every combine also counts 1 into M
//   shade       directional light (shadowed), the point light and the emissive
//               point times their W (shadowed)
// Each UpdateReservoir adds 1 to M, as in the reference.  Deviations: PCG32 in
// place of the wall-clock gold noise (numbers per pixel, stream = pixel; pass A
// state ^ kPhaseQA: per point-light candidate u_idx, u_wrs; per emissive
// candidate u_inst, u_tri, u_b1, u_b2, u_wrs; then per combine u_pl, u_em; pass B
// state ^ kPhaseQB: u_pl, u_em of the own combine, per neighbour (ux, uy) per try
// (at most 16 tries) and u_pl, u_em when it is combined); the emissive point is
// interpolated from the world-space triangle and its area is the world-space
// area (the reference's det(M) x object area, equal for rigid transforms); the
// history is found by reprojection into the previous camera (no motion vectors)
// and its normal is read at the nearest texel; surfaces of the previous frame
// without a surface never pass the normal tests.
constexpr uint64_t kPhaseQA = 0x51314141ull, kPhaseQB = 0x51314242ull;
constexpr int kQBlock = 128;
constexpr uint32_t kQInvalid = 0xFFFFFFFFu;

struct RefLights {
    const float4* pl;          // per point light, 3 float4: colour | position | (constant, linear, exp)
    uint32_t n_pl;
    const uint32_t* em_inst;   // emissive instance list
    uint32_t n_em;
    float4 dir_colour, dir_l;  // the block's directional light (its direction as stored)
};

struct QRes {
    float wsum_pl, w_pl;
    uint32_t pl, m_pl;
    uint32_t inst, tri;
    float b1, b2;
    float wsum_em, w_em;
    uint32_t m_em;
};
__device__ __forceinline__ QRes qres_empty() {
    return QRes{0.0f, 0.0f, kQInvalid, 0u, kQInvalid, kQInvalid, 0.0f, 0.0f, 0.0f, 0.0f, 0u};
}
__device__ __forceinline__ QRes qres_load(const float4* r, uint32_t i) {
    const float4 a = r[3 * (size_t)i], b = r[3 * (size_t)i + 1], c = r[3 * (size_t)i + 2];
    return QRes{a.x, a.y, __float_as_uint(a.z), __float_as_uint(a.w), __float_as_uint(b.x), __float_as_uint(b.y),
                b.z, b.w, c.x, c.y, __float_as_uint(c.z)};
}
__device__ __forceinline__ void qres_store(float4* r, uint32_t i, const QRes& x) {
    r[3 * (size_t)i] = make_float4(x.wsum_pl, x.w_pl, __uint_as_float(x.pl), __uint_as_float(x.m_pl));
    r[3 * (size_t)i + 1] = make_float4(__uint_as_float(x.inst), __uint_as_float(x.tri), x.b1, x.b2);
    r[3 * (size_t)i + 2] = make_float4(x.wsum_em, x.w_em, __uint_as_float(x.m_em), 0.0f);
}

__device__ __forceinline__ float vlen(V3 a) { return sqrtf(dot(a, a)); }

// CookTorranceBRDF (LightBuffers.glsl:103-124): (kd / PI + spec) n.l, kd without albedo
__device__ __forceinline__ V3 ct_ref(V3 lpos, V3 sp, V3 v, V3 f0, V3 n, float rough, float metal) {
    const V3 l = normalize(lpos - sp);
    const V3 hh = normalize(v + l);
    const float c1 = fminf(fmaxf(1.0f - fmaxf(dot(hh, v), 0.0f), 0.0f), 1.0f), c2 = c1 * c1;
    const float fw = (c2 * c2) * c1;
    const V3 F{f0.x + (1.0f - f0.x) * fw, f0.y + (1.0f - f0.y) * fw, f0.z + (1.0f - f0.z) * fw};
    const float a = rough * rough;
    const float D = ggx_d(fmaxf(dot(n, hh), 0.0f), a * a);
    const float r = rough + 1.0f, k = (r * r) / 8.0f;
    const float nv = fmaxf(dot(n, v), 0.0f), nl = fmaxf(dot(n, l), 0.0f);
    const float G = g1_schlick(nv, k) * g1_schlick(nl, k);
    const float den = 4.0f * nv * nl + 0.0001f;
    const float km = 1.0f - metal;
    const V3 spec = F * (D * G / den);
    const V3 kd{(1.0f - F.x) * km, (1.0f - F.y) * km, (1.0f - F.z) * km};
    return (kd * (1.0f / kRefPi) + spec) * nl;
}

// CalcPointlight (LightBuffers.glsl:145-157)
__device__ __forceinline__ V3 calc_point(const RefLights& L, uint32_t k, V3 v, V3 f0, V3 wp, V3 n, const Mat& m) {
    const float4 c = L.pl[3 * (size_t)k], p = L.pl[3 * (size_t)k + 1], at = L.pl[3 * (size_t)k + 2];
    const V3 lp{p.x, p.y, p.z};
    const float dist = vlen(lp - wp);
    const float atten = (at.x + at.y * dist) + at.z * (dist * dist);
    const V3 brdf = ct_ref(lp, wp, v, f0, n, m.roughness, m.metallic);
    return mul(mul(brdf, m.albedo), V3{c.x, c.y, c.z}) * (1.0f / atten);
}

// CalcDirectionalLight (LightBuffers.glsl:127-143)
__device__ __forceinline__ V3 calc_dir(const RefLights& L, V3 v, V3 f0, V3 n, const Mat& m) {
    const V3 l{L.dir_l.x, L.dir_l.y, L.dir_l.z};
    const V3 hh = normalize(v + l);
    const float c1 = fminf(fmaxf(1.0f - fmaxf(dot(hh, v), 0.0f), 0.0f), 1.0f), c2 = c1 * c1;
    const float fw = (c2 * c2) * c1;
    const V3 F{f0.x + (1.0f - f0.x) * fw, f0.y + (1.0f - f0.y) * fw, f0.z + (1.0f - f0.z) * fw};
    const float a = m.roughness * m.roughness;
    const float D = ggx_d(fmaxf(dot(n, hh), 0.0f), a * a);
    const float r = m.roughness + 1.0f, k = (r * r) / 8.0f;
    const float nv = fmaxf(dot(n, v), 0.0f), nl = fmaxf(dot(n, l), 0.0f);
    const float G = g1_schlick(nv, k) * g1_schlick(nl, k);
    const float den = 4.0f * nv * nl + 0.0001f;
    const float km = 1.0f - m.metallic;
    const V3 spec = F * (D * G / den);
    const V3 kd{(1.0f - F.x) * km * m.albedo.x, (1.0f - F.y) * km * m.albedo.y, (1.0f - F.z) * km * m.albedo.z};
    return mul((kd * (1.0f / kRefPi) + spec) * nl, V3{L.dir_colour.x, L.dir_colour.y, L.dir_colour.z});
}

// Point, normal (normalised, not turned), emission and world-space area of the
// emissive triangle `tri` of instance `inst` at barycentrics (b1, b2, 1 - b1 - b2)
// of its vertices 0, 1, 2 (TriangleShadowTest / ComputeTriangleRadiance, :241-282).
struct QTri {
    V3 p, n, le;
    float area;
};
__device__ __forceinline__ QTri tri_point(const SceneDev& S, uint32_t inst, uint32_t tri, float b1, float b2) {
    const InstD in = S.inst[inst];
    const float4* w = S.world + 3 * (size_t)(S.tri_off[inst] + tri);
    const float4 r0 = w[0], r1 = w[1], r2 = w[2];
    const V3 v0{r0.x, r0.y, r0.z}, v1{r0.w, r1.x, r1.y}, v2{r1.z, r1.w, r2.x};
    const float b3 = 1.0f - b1 - b2;
    QTri q;
    q.p = (v0 * b1 + v1 * b2) + v2 * b3;
    const uint32_t* ix = S.idx + in.index_offset + 3 * tri;
    const V3 n0 = ld3(S.nrm + 3 * (size_t)(ix[0] + in.vertex_offset));
    const V3 n1 = ld3(S.nrm + 3 * (size_t)(ix[1] + in.vertex_offset));
    const V3 n2 = ld3(S.nrm + 3 * (size_t)(ix[2] + in.vertex_offset));
    const V3 nl = (n0 * b1 + n1 * b2) + n2 * b3;
    const float4* x = S.xf + 6 * (size_t)inst;
    const float4 m0 = x[3], m1 = x[4], m2 = x[5];
    q.n = normalize(V3{(m0.x * nl.x + m0.y * nl.y) + m0.z * nl.z, (m1.x * nl.x + m1.y * nl.y) + m1.z * nl.z,
                       (m2.x * nl.x + m2.y * nl.y) + m2.z * nl.z});
    q.area = 0.5f * vlen(cross(v1 - v0, v1 - v2));
    const MatD& md = S.mats[in.material_idx];
    q.le = V3{md.albedo[0], md.albedo[1], md.albedo[2]};
    return q;
}

// ShadowTest (:101-121): 1 when nothing lies on (ro + t rd, t in [0.001, maxd])
template <int STACK>
__device__ __forceinline__ float qshadow(const SceneDev& S, V3 ro, V3 rd, float maxd, uint2* stk, uint32_t& n_sh) {
    uint32_t nv = 0, tv = 0;
    ++n_sh;
    const Ray r{ro.x, ro.y, ro.z, 0.001f, rd.x, rd.y, rd.z, maxd};
    return traverse<true, false, STACK>(S.nodes, S.leaf, r, stk, kQBlock, nv, tv).rec == 0xFFFFFFFFu ? 1.0f : 0.0f;
}

// ComputeTriangleRadiance (:258-282)
template <int STACK>
__device__ __forceinline__ V3 tri_radiance(const SceneDev& S, uint32_t inst, uint32_t tri, float b1, float b2, V3 v,
                                           V3 f0, V3 wp, V3 n, const Mat& m, V3 biased, bool shadow, uint2* stk,
                                           uint32_t& n_sh) {
    const QTri q = tri_point(S, inst, tri, b1, b2);
    const V3 brdf = ct_ref(q.p, wp, v, f0, n, m.roughness, m.metallic);
    const V3 d = wp - q.p;
    V3 rad = mul(mul(q.le, m.albedo), brdf) * (1.0f / fabsf(dot(d, d)));
    if (shadow) {
        const V3 tgt = q.p + q.n * 0.01f;
        const float pd = vlen(biased - tgt);
        rad = rad * qshadow<STACK>(S, biased, (tgt - biased) * (1.0f / pd), pd, stk, n_sh);
    }
    return rad;
}

__device__ __forceinline__ bool q_update(float& wsum, uint32_t& M, float weight, float u) {
    M += 1u;
    wsum += weight;
    return wsum > 0.0f && u < weight / wsum;
}

// Primary-vertex inputs of the reference's shading (RayGen.rgen:547-553)
struct QVert {
    V3 p, n, v, f0, biased;
    Mat m;
};
__device__ __forceinline__ QVert qvert(const GSurf& g) {
    QVert q;
    q.p = g.p;
    q.n = g.ns;
    q.v = g.wo;
    q.m = g.m;
    q.f0 = V3{0.04f + (g.m.albedo.x - 0.04f) * g.m.metallic, 0.04f + (g.m.albedo.y - 0.04f) * g.m.metallic,
              0.04f + (g.m.albedo.z - 0.04f) * g.m.metallic};
    q.biased = g.p + g.ns * (0.01f + fmaxf(g.t * 0.001f, 0.01f));
    return q;
}

// CombinePixelReservoir (:361-386)
template <int STACK>
__device__ __forceinline__ void q_combine(const SceneDev& S, const RefLights& L, const QVert& x, QRes& r,
                                          const QRes& nb, float& ph_pl, float& ph_em, Rng& rng, uint2* stk,
                                          uint32_t& n_sh) {
    const float u_pl = rng.uniform(), u_em = rng.uniform();
    if (nb.pl != kQInvalid) {
        const float4 pp = L.pl[3 * (size_t)nb.pl + 1];
        const V3 lp{pp.x, pp.y, pp.z};
        const float ld = vlen(lp - x.biased);
        const float ph = vlen(calc_point(L, nb.pl, x.v, x.f0, x.p, x.n, x.m)) *
                         qshadow<STACK>(S, x.biased, (lp - x.biased) * (1.0f / ld), ld - 0.001f, stk, n_sh);
        const float weight = ph * (float)nb.m_pl * nb.w_pl;
        if (q_update(r.wsum_pl, r.m_pl, weight, u_pl)) {
            r.pl = nb.pl;
            ph_pl = ph;
        }
    }
    if (nb.inst != kQInvalid) {
        const float ph = vlen(tri_radiance<STACK>(S, nb.inst, nb.tri, nb.b1, nb.b2, x.v, x.f0, x.p, x.n, x.m,
                                                  x.biased, true, stk, n_sh));
        const float weight = ph * nb.w_em * (float)nb.m_em;
        if (q_update(r.wsum_em, r.m_em, weight, u_em)) {
            r.inst = nb.inst;
            r.tri = nb.tri;
            r.b1 = nb.b1;
            r.b2 = nb.b2;
            ph_em = ph;
        }
    }
}

// Pass A: DoInitialWRS + WriteInitialReservoirs (:296-357, :411-468) -> tmp
template <int STACK>
__global__ __launch_bounds__(kQBlock) void k_q_initial(SceneDev S, RefLights L, GView G, GView Gp,
                                                       const float4* __restrict__ hist, float4* __restrict__ tmp,
                                                       uint32_t w, uint32_t h, PrevCam pc, int temporal,
                                                       uint64_t initstate, unsigned long long* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kQBlock];
    __shared__ uint32_t red;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) red = 0u;
    __syncthreads();
    uint32_t n_sh = 0;
    if (i < w * h) {
        const GSurf g = load_g(G, i);
        QRes out = qres_empty();
        if (g.valid) {
            const QVert x = qvert(g);
            Rng rng = rng_seed(initstate ^ kPhaseQA, i);
            uint2* sk = &stk[threadIdx.x];
            QRes r = qres_empty();
            float ph_pl = 0.0f, ph_em = 0.0f;
            for (uint32_t c = 0; c < 32u; ++c) {
                const float u = rng.uniform(), uw = rng.uniform();
                uint32_t k = 0;
                float ph = 0.0f, weight = 0.0f;
                if (L.n_pl > 0) {
                    k = min((uint32_t)(u * (float)L.n_pl), L.n_pl - 1u);
                    ph = vlen(calc_point(L, k, x.v, x.f0, x.p, x.n, x.m));
                    weight = ph / (1.0f / (float)L.n_pl);
                }
                if (q_update(r.wsum_pl, r.m_pl, weight, uw)) {
                    r.pl = k;
                    ph_pl = ph;
                }
            }
            for (uint32_t c = 0; c < 32u; ++c) {
                const float ui = rng.uniform(), ut = rng.uniform(), ub1 = rng.uniform(), ub2 = rng.uniform(),
                            uw = rng.uniform();
                if (L.n_em == 0) {
                    r.m_em += 1u;
                    continue;
                }
                const uint32_t inst = L.em_inst[min((uint32_t)(ui * (float)L.n_em), L.n_em - 1u)];
                const uint32_t ntri = S.inst[inst].num_indices / 3u;
                if (ntri == 0) {
                    r.m_em += 1u;
                    continue;
                }
                const uint32_t tri = min((uint32_t)(ut * (float)ntri), ntri - 1u);
                float b1 = ub1, b2 = ub2;
                if (b1 + b2 >= 1.0f) {
                    b1 = 1.0f - b1;
                    b2 = 1.0f - b2;
                }
                const QTri q = tri_point(S, inst, tri, b1, b2);
                const V3 rad = tri_radiance<STACK>(S, inst, tri, b1, b2, x.v, x.f0, x.p, x.n, x.m, x.biased, false,
                                                   sk, n_sh);
                const float ph = vlen(rad);
                const float pdf = 1.0f / (((float)L.n_em * (float)ntri) * q.area);
                const float weight = ph / pdf;
                if (q_update(r.wsum_em, r.m_em, weight, uw)) {
                    r.inst = inst;
                    r.tri = tri;
                    r.b1 = b1;
                    r.b2 = b2;
                    ph_em = ph;
                }
            }
            if (r.pl != kQInvalid) {
                const float4 pp = L.pl[3 * (size_t)r.pl + 1];
                const V3 lp{pp.x, pp.y, pp.z};
                const float d = vlen(lp - x.biased);
                r.w_pl = qshadow<STACK>(S, x.biased, lp - x.biased * (1.0f / d), d, sk, n_sh) *
                         ((r.wsum_pl / (float)r.m_pl) / ph_pl);
            }
            if (r.inst != kQInvalid) {
                const QTri q = tri_point(S, r.inst, r.tri, r.b1, r.b2);
                const V3 tgt = q.p + q.n * 0.01f;
                const float pd = vlen(x.biased - tgt);
                r.w_em = qshadow<STACK>(S, x.biased, (tgt - x.biased) * (1.0f / pd), pd, sk, n_sh) *
                         ((r.wsum_em / (float)r.m_em) / ph_em);
            }
            // temporal resample: history by reprojection into the previous camera
            int q = -1;
            if (temporal) {
                float cl[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    cl[k] = ((pc.pv[k] * g.p.x + pc.pv[4 + k] * g.p.y) + pc.pv[8 + k] * g.p.z) + pc.pv[12 + k];
                if (cl[3] > 1e-6f) {
                    const float fx = (cl[0] / cl[3] * 0.5f + 0.5f) * (float)w;
                    const float fy = (cl[1] / cl[3] * 0.5f + 0.5f) * (float)h;
                    if (fx > 0.0f && fy > 0.0f && fx < (float)w && fy < (float)h) q = (int)fy * (int)w + (int)fx;
                }
            }
            if (q >= 0) {
                const float4 pq = Gp.p[q], nq = Gp.n[q];
                if (!(pq.w > 0.0f) || !(dot(V3{nq.x, nq.y, nq.z}, x.n) > 0.95f)) q = -1;
            }
            if (q >= 0) {
                QRes hh = qres_load(hist, (uint32_t)q);
                if ((float)hh.m_pl > 20.0f * (float)r.m_pl) {
                    hh.wsum_pl *= 20.0f * (float)r.m_pl / (float)hh.m_pl;
                    hh.m_pl = 20u * r.m_pl;
                }
                if ((float)hh.m_em > 20.0f * (float)r.m_em) {
                    hh.wsum_em *= 20.0f * (float)r.m_em / (float)hh.m_em;
                    hh.m_em = 20u * r.m_em;
                }
                float tp_pl = 0.0f, tp_em = 0.0f;
                q_combine<STACK>(S, L, x, out, r, tp_pl, tp_em, rng, sk, n_sh);
                q_combine<STACK>(S, L, x, out, hh, tp_pl, tp_em, rng, sk, n_sh);
                out.m_em = hh.m_em + r.m_em;
                out.m_pl = hh.m_pl + r.m_pl;
                out.w_em = tp_em > 0.0f ? (out.wsum_em / (float)out.m_em) / tp_em : 0.0f;
                out.w_pl = tp_pl > 0.0f ? (out.wsum_pl / (float)out.m_pl) / tp_pl : 0.0f;
            }
        }
        qres_store(tmp, i, out);
    }
    for (int o = 32; o > 0; o >>= 1) n_sh += __shfl_xor(n_sh, o, 64);
    if ((threadIdx.x & 63) == 0 && n_sh) atomicAdd(&red, n_sh);
    __syncthreads();
    if (threadIdx.x == 0 && stats && red) atomicAdd(&stats[2], (unsigned long long)red);
}

// Pass B: ComputeAndWriteSpatialReservoirReuse + CalculateRadianceFromReservoirs
// + the directional light (:470-528, :560-565): tmp -> hist, radiance into rad
template <int STACK>
__global__ __launch_bounds__(kQBlock) void k_q_spatial(SceneDev S, RefLights L, GView G, GView Gp, int have_prev,
                                                       const float4* __restrict__ tmp, float4* __restrict__ hist,
                                                       float4* __restrict__ rad, uint32_t w, uint32_t h, uint32_t k,
                                                       float radius, uint64_t initstate,
                                                       unsigned long long* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kQBlock];
    __shared__ uint32_t red;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) red = 0u;
    __syncthreads();
    uint32_t n_sh = 0;
    if (i < w * h) {
        const GSurf g = load_g(G, i);
        QRes r = qres_empty();
        if (g.valid) {
            const QVert x = qvert(g);
            Rng rng = rng_seed(initstate ^ kPhaseQB, i);
            uint2* sk = &stk[threadIdx.x];
            float ph_pl = 0.0f, ph_em = 0.0f;
            const QRes own = qres_load(tmp, i);
            q_combine<STACK>(S, L, x, r, own, ph_pl, ph_em, rng, sk, n_sh);
            r.m_em += own.m_em;
            r.m_pl += own.m_pl;
            const float cx0 = (float)(i % w) / (float)w, cy0 = (float)(i / w) / (float)h;
            for (uint32_t s = 0; s < k; ++s) {
                float cx = -1.0f, cy = -1.0f;
                for (int t = 0; t < 16; ++t) {
                    const float ux = rng.uniform(), uy = rng.uniform();
                    const float ox = (ux - 0.5f) * 2.0f * radius / (float)w, oy = (uy - 0.5f) * 2.0f * radius / (float)h;
                    const float tx = cx0 + ox, ty = cy0 + oy;
                    if (tx >= 0.0f && ty >= 0.0f && tx <= 1.0f && ty <= 1.0f) {
                        cx = tx;
                        cy = ty;
                        break;
                    }
                }
                if (cx < 0.0f || !have_prev) continue;
                const uint32_t nx = min((uint32_t)((float)w * cx), w - 1u), ny = min((uint32_t)((float)h * cy), h - 1u);
                const uint32_t q = ny * w + nx;
                const float4 pq = Gp.p[q], nq = Gp.n[q];
                if (!(pq.w > 0.0f) || dot(V3{nq.x, nq.y, nq.z}, x.n) < 0.95f) continue;
                const QRes nb = qres_load(tmp, q);
                q_combine<STACK>(S, L, x, r, nb, ph_pl, ph_em, rng, sk, n_sh);
                r.m_em += nb.m_em;
                r.m_pl += nb.m_pl;
            }
            r.w_em = ph_em > 0.0f ? r.wsum_em / ph_em / (float)r.m_em : 0.0f;
            r.w_pl = ph_pl > 0.0f ? r.wsum_pl / ph_pl / (float)r.m_pl : 0.0f;
            // the pixel's light
            V3 light = calc_dir(L, x.v, x.f0, x.n, x.m);
            if (light.x > 0.0f || light.y > 0.0f || light.z > 0.0f)
                light = light * qshadow<STACK>(S, x.biased, V3{L.dir_l.x, L.dir_l.y, L.dir_l.z}, 1000.0f, sk, n_sh);
            if (L.n_pl > 0 && r.pl != kQInvalid) {
                const float4 pp = L.pl[3 * (size_t)r.pl + 1];
                const V3 lp{pp.x, pp.y, pp.z};
                const float ld = vlen(lp - x.biased);
                const V3 c = calc_point(L, r.pl, x.v, x.f0, x.p, x.n, x.m) *
                             qshadow<STACK>(S, x.biased, (lp - x.biased) * (1.0f / ld), ld, sk, n_sh);
                light = light + c * r.w_pl;
            }
            if (L.n_em > 0 && r.inst != kQInvalid) {
                const V3 c = tri_radiance<STACK>(S, r.inst, r.tri, r.b1, r.b2, x.v, x.f0, x.p, x.n, x.m, x.biased,
                                                 true, sk, n_sh);
                light = light + c * r.w_em;
            }
            const float4 o = rad[i];
            rad[i] = make_float4(o.x + light.x, o.y + light.y, o.z + light.z, o.w);
        }
        qres_store(hist, i, r);
    }
    for (int o = 32; o > 0; o >>= 1) n_sh += __shfl_xor(n_sh, o, 64);
    if ((threadIdx.x & 63) == 0 && n_sh) atomicAdd(&red, n_sh);
    __syncthreads();
    if (threadIdx.x == 0 && stats && red) atomicAdd(&stats[2], (unsigned long long)red);
}
}  // namespace

#define HIPR(expr)                                                \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr);    \
    } while (0)

// The camera and primary surfaces of this frame become the next frame's history.
static void remember_history(RestirState& X, RenderState& R, const snkpt_camera* cam) {
    std::memcpy(X.prev_pv, cam->proj_view, sizeof(X.prev_pv));
    X.prev_pos[0] = cam->cam_pos[0];
    X.prev_pos[1] = cam->cam_pos[1];
    X.prev_pos[2] = cam->cam_pos[2];
    X.cur ^= 1;
    X.have_prev = true;
    R.history_valid = true;
}

// SNKPT_MODE_REFERENCE_COMPAT after the primary vertices: the reference's own
// ReSTIR passes A and B (direct light) and BounceDiffuse (unless SNKPT_FLAG_NO_BOUNCE).
static snkpt_status compat_pass(snkpt_ctx* ctx, const snkpt_settings* s, const SceneDev& S, const GView& G,
                                const GView& Gp, bool temporal, uint64_t initstate, hipStream_t st) {
    RenderState& R = *ctx->render;
    RestirState& X = *R.restir;
    const uint32_t w = R.w, h = R.h, n = w * h;
    const Accel& A = ctx->accel;
    if (A.num_tris == 0) return SNKPT_OK;
    const uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
    if (need > 64) return fail(ctx, SNKPT_ERR_CAPACITY, "compat: BVH8 deeper than the 64-entry traversal stack");
    unsigned long long* stats = R.stats.as<unsigned long long>();
    const Emitters& E = ctx->emit;
    RefLights L;
    L.pl = E.ref_pl.as<float4>();
    L.n_pl = E.n_ref_pl;
    L.em_inst = E.em_list.as<uint32_t>();
    L.n_em = E.n_em_list;
    L.dir_colour = make_float4(E.ref_dir[0], E.ref_dir[1], E.ref_dir[2], 0.0f);
    L.dir_l = make_float4(E.ref_dir[4], E.ref_dir[5], E.ref_dir[6], 0.0f);
    PrevCam pc;
    std::memcpy(pc.pv, X.prev_pv, sizeof(pc.pv));
    pc.pos = V3{X.prev_pos[0], X.prev_pos[1], X.prev_pos[2]};
    const uint32_t k = s->num_spatial_resamples > 0 ? std::min((uint32_t)s->num_spatial_resamples, 64u) : 0u;
    const float radius = s->restir_radius > 0.0f ? s->restir_radius : 30.0f;
    const uint32_t qg = (n + kQBlock - 1) / kQBlock;
    float4* q_tmp = X.q_tmp.as<float4>();
    float4* q_hist = X.q_hist.as<float4>();
    const int have_prev = X.have_prev ? 1 : 0;
#define Q_LAUNCH(ST)                                                                                              \
    do {                                                                                                          \
        k_q_initial<ST><<<qg, kQBlock, 0, st>>>(S, L, G, Gp, q_hist, q_tmp, w, h, pc, temporal ? 1 : 0, initstate, \
                                                stats);                                                           \
        k_q_spatial<ST><<<qg, kQBlock, 0, st>>>(S, L, G, Gp, have_prev, q_tmp, q_hist, R.rad.as<float4>(), w, h,  \
                                                k, radius, initstate, stats);                                     \
    } while (0)
    if (need <= 16)
        Q_LAUNCH(16);
    else if (need <= 32)
        Q_LAUNCH(32);
    else
        Q_LAUNCH(64);
#undef Q_LAUNCH
    if (!(s->flags & SNKPT_FLAG_NO_BOUNCE)) {
        // BounceDiffuse over this frame's primary vertices
        const uint32_t grid = (n + kCompatBlock - 1) / kCompatBlock;
        if (need <= 16)
            k_compat_bounce<16><<<grid, kCompatBlock, 0, st>>>(S, G, R.rad.as<float4>(), n, initstate, stats);
        else if (need <= 32)
            k_compat_bounce<32><<<grid, kCompatBlock, 0, st>>>(S, G, R.rad.as<float4>(), n, initstate, stats);
        else
            k_compat_bounce<64><<<grid, kCompatBlock, 0, st>>>(S, G, R.rad.as<float4>(), n, initstate, stats);
    }
    HIPR(hipGetLastError());
    return SNKPT_OK;
}

snkpt_status restir_pass(snkpt_ctx* ctx, const snkpt_camera* cam, const CamDev& cd, const snkpt_settings* s,
                         uint64_t initstate, uint32_t pass, hipStream_t st) {
    RenderState& R = *ctx->render;
    if (!R.restir) R.restir = new RestirState();
    RestirState& X = *R.restir;
    const uint32_t w = R.w, h = R.h, n = w * h;
    if (n > X.cap) {
        for (int f = 0; f < 2; ++f)
            for (int k = 0; k < 5; ++k) HIPR(X.g[f][k].reserve((size_t)n * 16));
        HIPR(X.r_init.reserve((size_t)n * 16));
        HIPR(X.r_tmp.reserve((size_t)n * 16));
        HIPR(X.r_fin.reserve((size_t)n * 16));
        HIPR(X.m_init.reserve((size_t)n * 4));
        HIPR(X.m_tmp.reserve((size_t)n * 4));
        HIPR(X.m_fin.reserve((size_t)n * 4));
        HIPR(X.a_init.reserve((size_t)n * 16));
        HIPR(X.a_tmp.reserve((size_t)n * 16));
        HIPR(X.a_fin.reserve((size_t)n * 16));
        HIPR(X.q_tmp.reserve((size_t)n * 48));
        HIPR(X.q_hist.reserve((size_t)n * 48));
        X.cap = n;
        X.have_prev = false;
    }
    if (!R.history_valid) X.have_prev = false;
    SceneDev S = scene_dev(ctx);
    auto gview = [&](int f) {
        return GView{X.g[f][0].as<float4>(), X.g[f][1].as<float4>(), X.g[f][2].as<float4>(),
                     X.g[f][3].as<float4>(), X.g[f][4].as<float4>()};
    };
    GView G = gview(X.cur), Gp = gview(X.cur ^ 1);
    uint32_t* counts = R.counts.as<uint32_t>();
    unsigned long long* stats = R.stats.as<unsigned long long>();
    const bool compat = s->mode == SNKPT_MODE_REFERENCE_COMPAT;
    // the compatibility estimator replaces the path after the primary vertex
    const uint32_t max_depth = compat ? 1u : (s->max_depth ? s->max_depth : 1u);
    snkpt_status e = SNKPT_OK;
    // the reference-compatible mode runs its own candidates (pass A below)
    const uint32_t n_cand = compat ? 0u : (s->restir_candidates ? s->restir_candidates : 32u);
    if (R.use_gb) {
        // primary vertices from the caller's G-buffer
        k_counts3<<<1, 64, 0, st>>>(counts, n, 0u, 0u);
        k_primary<true><<<grid_for(n, kBlock), kBlock, 0, st>>>(
            S, G, queue_of(R, 0), hits_of(R), R.gb, cd, w, queue_of(R, 1), paths_of(R), R.rad.as<float4>(),
            X.r_init.as<float4>(), X.m_init.as<float>(), X.a_init.as<float4>(), counts, n, n_cand, max_depth,
            s->rr_start, initstate, stats);
    } else {
        // camera rays + primary hits
        generate_camera(ctx, cd, initstate, st);
        snkpt_hits hh{R.h_t.as<float>(), R.h_inst.as<uint32_t>(), R.h_prim.as<uint32_t>(), R.h_u.as<float>(),
                      R.h_v.as<float>()};
        e = launch_trace(ctx, R.q[0].org.as<float4>(), R.q[0].dir.as<float4>(), n, hh, false, nullptr, nullptr,
                         nullptr, st);
        if (e != SNKPT_OK) return e;
        k_counts3<<<1, 64, 0, st>>>(counts, n, 0u, 0u);
        k_primary<false><<<grid_for(n, kBlock), kBlock, 0, st>>>(
            S, G, queue_of(R, 0), hits_of(R), R.gb, cd, w, queue_of(R, 1), paths_of(R), R.rad.as<float4>(),
            X.r_init.as<float4>(), X.m_init.as<float>(), X.a_init.as<float4>(), counts, n, n_cand, max_depth,
            s->rr_start, initstate, stats);
    }
    const bool temporal = X.have_prev && s->num_temporal_resamples != 0 && !cam->prev_frame_invalid;
    if (compat) {
        if ((e = compat_pass(ctx, s, S, G, Gp, temporal, initstate, st)) != SNKPT_OK) return e;
        remember_history(X, R, cam);
        return SNKPT_OK;
    }
    const float4* r_sp_in = X.r_init.as<float4>();
    const float4* a_sp_in = X.a_init.as<float4>();
    const float* m_sp_in = X.m_init.as<float>();
    if (temporal) {
        PrevCam pc;
        std::memcpy(pc.pv, X.prev_pv, sizeof(pc.pv));
        pc.pos = V3{X.prev_pos[0], X.prev_pos[1], X.prev_pos[2]};
        float m_cap = (float)(s->restir_m_cap ? s->restir_m_cap : 20u);
        k_temporal<<<grid_for(n, kBlock), kBlock, 0, st>>>(S, G, Gp, X.r_init.as<float4>(), X.m_init.as<float>(),
                                                           X.r_fin.as<float4>(), X.m_fin.as<float>(),
                                                           X.r_tmp.as<float4>(), X.m_tmp.as<float>(), w, h, pc, m_cap,
                                                           initstate, R.use_gb ? R.gb.motion : nullptr, stats,
                                                           X.a_init.as<float4>(), X.a_fin.as<float4>(),
                                                           X.a_tmp.as<float4>());
        r_sp_in = X.r_tmp.as<float4>();
        a_sp_in = X.a_tmp.as<float4>();
        m_sp_in = X.m_tmp.as<float>();
    }
    const uint32_t k = s->num_spatial_resamples > 0 ? (uint32_t)s->num_spatial_resamples : 0u;
    // the register arrays hold one slot per input: the 5-neighbour variant (the
    // reference's configurations) keeps fewer live than the 8-neighbour one
    if (k <= 5)
        k_spatial<5><<<xcd_tile_blocks(w, h), kBlock, 0, st>>>(
            S, G, r_sp_in, m_sp_in, X.r_fin.as<float4>(), X.m_fin.as<float>(), w, h, k,
            s->restir_radius > 0.0f ? s->restir_radius : 30.0f, initstate);
    else
        k_spatial<kMaxSpatial><<<xcd_tile_blocks(w, h), kBlock, 0, st>>>(
            S, G, r_sp_in, m_sp_in, X.r_fin.as<float4>(), X.m_fin.as<float>(), w, h, k,
            s->restir_radius > 0.0f ? s->restir_radius : 30.0f, initstate);
    k_rshade<<<grid_for(n, kBlock), kBlock, 0, st>>>(S, G, X.r_fin.as<float4>(), X.m_fin.as<float>(), shadow_of(R),
                                                     counts, n, stats);
    if ((e = launch_shadow(ctx, R, R.rad.as<float4>(), st)) != SNKPT_OK) return e;
    k_restir_stats<<<1, 64, 0, st>>>(counts, stats);
    if (S.n_alights > 0) {
        // the shadow queue is reused for the analytic lights of the primary hits
        k_restir_stats2<<<1, 64, 0, st>>>(counts, nullptr);
        k_spatial_analytic<<<xcd_tile_blocks(w, h), kBlock, 0, st>>>(S, G, a_sp_in, X.a_fin.as<float4>(), w, h, k,
                                   s->restir_radius > 0.0f ? s->restir_radius : 30.0f, initstate);
        k_rshade_analytic<<<grid_for(n, kBlock), kBlock, 0, st>>>(S, G, X.a_fin.as<float4>(), shadow_of(R), counts,
                                                                  n);
        if ((e = launch_shadow(ctx, R, R.rad.as<float4>(), st)) != SNKPT_OK) return e;
        k_restir_stats2<<<1, 64, 0, st>>>(counts, stats);
    }
    remember_history(X, R, cam);
    (void)pass;
    // remaining bounces with the plain path tracer
    if (max_depth > 1) return pt_bounces(ctx, s, 1, 1, st);
    return SNKPT_OK;
}

}  // namespace snkpt
