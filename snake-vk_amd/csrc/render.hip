// The path-trace pass: RT::RecordRenderCmdBuf (Core/src/rendering/Raytracing.cpp:227-267)
// rebuilt as a wavefront path tracer.
//
// The reference runs two traceRaysKHR(W, H) dispatches whose raygen shader does
// primary shading from a G-buffer, ReSTIR and BounceDiffuse with recursive
// traceRayEXT calls (RayGen.rgen:530-570).  Here one frame is a sequence of
// kernels over structure-of-arrays queues in HBM:
//   generate  camera rays through pixel centres (Util.glsl:9-13), one path per pixel
//   extend    BVH8 closest hit of the ray queue (trace.hip)
//   shade     hit interpolation + material (RayClosestHit.rchit:88-123), emission with
//             MIS, next-event estimation to a power-sampled emissive triangle, GGX /
//             diffuse BRDF sampling, Russian roulette; survivors are appended to the
//             next ray queue and NEE samples to the shadow queue with one atomic per wave
//   shadow    any-hit traversal of the shadow queue (ShadowTest, RayGen.rgen:101-121),
//             unoccluded contributions added to the frame radiance
//   resolve   running mean over accumulated frames, RGBA32F or RGBA16F out
// Queue counts stay on the device, so a frame has no host round trip.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "render_common.h"
#include "pt_step.h"
#include "render_modes.h"

namespace snkpt {

namespace {

constexpr int kBlock = 256;
constexpr int kShadowBlock = 128;

__global__ void k_fill_f4(float4* p, uint32_t n, float4 v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_set_counts(uint32_t* c, uint32_t i0, uint32_t v0, uint32_t i1, uint32_t v1) {
    if (threadIdx.x == 0) {
        c[i0] = v0;
        c[i1] = v1;
    }
}

// Start of a bounce: count the extension rays queued for it (count_ext) and
// the shadow rays traced in the previous bounce (count_shadow), then clear the
// queues this bounce fills.  Counting from the queue sizes keeps per-wave
// statistics atomics out of the shade kernel.  The queue length of bounce
// b < kQueueHist goes to stats[kQueueHistAt + b] (the per-thread tail's choice).
constexpr uint32_t kQueueHist = 16, kQueueHistAt = 16;
__global__ void k_bounce_begin(uint32_t* c, uint32_t cur, uint32_t nxt, unsigned long long* stats,
                               uint32_t count_ext, uint32_t count_shadow, uint32_t bounce) {
    if (threadIdx.x == 0) {
        if (count_ext) stats[1] += c[cur];
        if (count_shadow) stats[2] += c[2];
        if (bounce < kQueueHist) stats[kQueueHistAt + bounce] = c[cur];
        c[nxt] = 0;
        c[2] = 0;
    }
}
__global__ void k_bounce_end(uint32_t* c, unsigned long long* stats) {
    if (threadIdx.x == 0) {
        stats[2] += c[2];
        c[2] = 0;
    }
}

// Pixel of queue slot i when a whole w x h frame is enqueued in 8 x 8 tiles:
// bands of 8 rows, each cut into tiles of 8 columns (the last band and the last
// tile of a band may be narrower), pixels row by row within a tile.  A
// bijection of [0, w h); 64 consecutive slots (one traversal chunk) are one tile.
__device__ __forceinline__ uint32_t tile_pixel(uint32_t i, uint32_t w, uint32_t h) {
    const uint32_t b = min(i / (8u * w), (h - 1u) / 8u);
    const uint32_t k = i - 8u * w * b;
    const uint32_t bh = min(8u, h - 8u * b);
    const uint32_t t = k / (8u * bh);
    const uint32_t j = k - 8u * bh * t;
    const uint32_t tw = min(8u, w - 8u * t);
    return (8u * b + j / tw) * w + 8u * t + j % tw;
}

// Camera rays of n paths.  Path p goes through pixel (start + p stride) mod
// (w h) and owns RNG stream p; with `tiled` (a whole frame: start 0, stride 1)
// queue slot i holds path tile_pixel(i), so the rays of one traversal chunk,
// and the queries and shadow rays their paths append, come from one 8 x 8
// tile.  Every path's work is the same in either order.
__global__ void k_generate(CamDev cam, uint32_t w, uint32_t h, uint32_t n, uint32_t start, uint32_t stride,
                           uint32_t tiled, uint64_t initstate, float4* __restrict__ q_org, float4* __restrict__ q_dir,
                           uint32_t* __restrict__ q_path, PathState ps, uint32_t* __restrict__ counts, uint32_t cur,
                           uint32_t* __restrict__ nrec) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) counts[cur] = n;
    if (i >= n) return;
    const uint32_t p = tiled ? tile_pixel(i, w, h) : i;
    uint32_t px = (uint32_t)(((uint64_t)start + (uint64_t)p * stride) % ((uint64_t)w * h));
    if (nrec) nrec[p] = 0;
    V3 d = camera_dir(cam, px % w, px / w, w, h);
    q_org[i] = make_float4(cam.pos.x, cam.pos.y, cam.pos.z, 0.0f);
    q_dir[i] = make_float4(d.x, d.y, d.z, 1e30f);
    q_path[i] = p;
    Rng r = rng_seed(initstate, p);
    ps.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    ps.rng[p] = make_uint2((uint32_t)r.state, (uint32_t)(r.state >> 32));
    ps.bounce[p] = 0;
}

// One path vertex of the unbiased path tracer over the queue of one scene.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_shade(SceneDev S, PtParams P, Queue in, HitArrays hit, Queue out,
                                                  ShadowQueue sq, PathState ps, float4* __restrict__ rad,
                                                  uint32_t* __restrict__ counts, uint32_t cur, uint32_t nxt,
                                                  unsigned long long* __restrict__ stats, ShadeHooks H) {
    __shared__ uint32_t lds_app[kBlock / 64 + 1];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const SceneOne ss{S};
    (void)stats;
    shade_step<KIND, kBlock>(ss, P, in, hit, out, sq, ps, rad, counts, nxt, H, i, i < counts[cur], lds_app);
}

// Bounce 0 from the G-buffer: the primary vertex of path i (pixel i) read from
// the rasterized G-buffer instead of a traced camera ray, then the same vertex
// logic and queue appends as the shade kernel.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_gshade(SceneDev S, GbufDev G, CamDev cam, uint32_t w, uint32_t h,
                                                   PtParams P, Queue out, ShadowQueue sq, PathState ps,
                                                   float4* __restrict__ rad, uint32_t* __restrict__ counts,
                                                   uint32_t nxt, uint64_t initstate, ShadeHooks H) {
    __shared__ uint32_t lds_app[kBlock / 64 + 1];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    VertexOut o;
    Rng rng{0, 0};
    if (i < w * h) {
        rng = rng_seed(initstate, i);
        if (KIND == 2) H.nrec[i] = 0;
        Surf s;
        V3 dir;
        float t;
        if (gbuffer_surface(G, cam, i, w, h, s, dir, t)) {
            const float4 rn = KIND == 2 ? rad[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            path_vertex_surf<KIND>(S, P, H, i, 0, dir, t, s, make_float4(1.0f, 1.0f, 1.0f, 0.0f), rng, rn, o);
            if (o.emit) {
                float4 r = rad[i];
                rad[i] = make_float4(r.x + o.c_emit.x, r.y + o.c_emit.y, r.z + o.c_emit.z, r.w);
            }
        }
    }
    vertex_appends<KIND, kBlock>(o, i, 0, rng, out, sq, ps, counts, nxt, H, lds_app);
}

// extension rays queued by bounce 0 and its shadow rays
__global__ void k_gbuf_stats(const uint32_t* c, unsigned long long* stats) {
    if (threadIdx.x == 0) {
        stats[1] += c[1];
        stats[2] += c[2];
    }
}

// Shadow rays: persistent any-hit traversal with ray refill; each wave owns a
// contiguous slice of the shadow queue.
template <int STACK>
__global__ __launch_bounds__(kShadowBlock) __attribute__((amdgpu_waves_per_eu(STACK <= 10 ? 8 : 1, 8))) void k_shadow(const uint4* __restrict__ nodes,
                                                         const float4* __restrict__ tris, ShadowQueue sq,
                                                         const uint32_t* __restrict__ counts,
                                                         float4* __restrict__ rad) {
    __shared__ uint2 stk[STACK * kShadowBlock];
    const uint32_t nn = counts[2];
    const uint32_t nwaves = gridDim.x * (kShadowBlock / 64);
    const uint32_t wave = blockIdx.x * (kShadowBlock / 64) + (threadIdx.x >> 6);
    auto load = [&](uint32_t i, Ray& ray, const uint4*& nd, const float4*& tr) {
        ray = shadow_ray(sq, i);
        nd = nodes;
        tr = tris;
    };
    auto done = [&](uint32_t i, const TravState& st, uint32_t, uint32_t) { shadow_done(sq, i, st.res, rad); };
    refill_trace<true, false, STACK, kRefillLanes>(nn, wave, nwaves, &stk[threadIdx.x], kShadowBlock, load, done);
}

// Whole paths, one per thread (KIND 0: radiance into rad[i]; KIND 2: training
// records as the wavefront's kind-2 shade).  Ray counts: one atomic per block.
constexpr int kMegaBlock = 128;
template <int KIND, int STACK>
__global__ __launch_bounds__(kMegaBlock) void k_pt_mega(SceneDev S, CamDev cam, uint32_t w, uint32_t h, uint32_t n,
                                                        uint32_t start, uint32_t stride, uint64_t initstate,
                                                        PtParams P, ShadeHooks H, float4* __restrict__ rad,
                                                        unsigned long long* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kMegaBlock];
    __shared__ uint32_t red[2];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (KIND == 2 && H.tseed) {  // a replayed frame's training paths: streams of this frame
        initstate = H.tseed->path_state;
        start = H.tseed->start;
    }
    if (threadIdx.x == 0) red[0] = red[1] = 0u;
    __syncthreads();
    uint32_t n_ext = 0, n_sh = 0;
    if (i < n) {
        rad[i] = whole_path<KIND, STACK>(S, cam, w, h, i, start, stride, initstate, P, H, &stk[threadIdx.x],
                                         kMegaBlock, rad[i], n_ext, n_sh);
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

// The late bounces of the plain path tracer, one thread per path (tail_paths).
template <int STACK>
__global__ __launch_bounds__(kMegaBlock) void k_pt_tail(SceneDev S, PtParams P, ShadeHooks H, Queue in,
                                                        const uint32_t* __restrict__ count, PathState ps,
                                                        float4* __restrict__ rad,
                                                        unsigned long long* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kMegaBlock];
    __shared__ uint32_t red[2];
    tail_paths<STACK, kMegaBlock>(SceneOne{S}, P, H, in, *count, ps, rad, &stk[threadIdx.x], stats, red);
}

// Output of the frame: its mean radiance, or with `accumulate` the weighted
// running mean (this frame's samples count `weight` each).
__global__ void k_resolve(const float4* __restrict__ rad, float4* __restrict__ accum, void* __restrict__ out,
                          uint32_t n, float spp, uint32_t accumulate, uint32_t fmt, float weight) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 r = rad[i];
    float4 o;
    if (accumulate) {
        float4 a = accum[i];
        a = make_float4(a.x + weight * r.x, a.y + weight * r.y, a.z + weight * r.z, a.w + weight * spp);
        accum[i] = a;
        float inv = a.w > 0.0f ? 1.0f / a.w : 0.0f;
        o = make_float4(a.x * inv, a.y * inv, a.z * inv, 1.0f);
    } else {
        float inv = 1.0f / spp;
        o = make_float4(r.x * inv, r.y * inv, r.z * inv, 1.0f);
    }
    if (fmt == SNKPT_FORMAT_RGBA16F) {
        __half* h = (__half*)out + 4 * (size_t)i;
        h[0] = __float2half(o.x);
        h[1] = __float2half(o.y);
        h[2] = __float2half(o.z);
        h[3] = __float2half(o.w);
    } else {
        ((float4*)out)[i] = o;
    }
}

__global__ void k_add_f64(const float4* __restrict__ rad, double* __restrict__ acc, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 r = rad[i];
    acc[3 * (size_t)i] += (double)r.x;
    acc[3 * (size_t)i + 1] += (double)r.y;
    acc[3 * (size_t)i + 2] += (double)r.z;
}

inline uint32_t nblk(uint32_t n, uint32_t b) { return n == 0 ? 1 : (n + b - 1) / b; }
// Resident blocks per CU of the per-thread tail (119 VGPRs: 4 waves per SIMD).
constexpr uint32_t kTailBlocksPerCu = 8;
constexpr uint32_t kTailNever = 0xFFFFFFFFu;

// inverse of a 4x4 (row-major) by cofactors, in double
bool inverse4(const double m[16], double inv[16]) {
    double a[16];
    a[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    a[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    a[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    a[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    a[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    a[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    a[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    a[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    a[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    a[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    a[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    a[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    a[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    a[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    a[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    a[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * a[0] + m[1] * a[4] + m[2] * a[8] + m[3] * a[12];
    if (det == 0.0) return false;
    for (int i = 0; i < 16; ++i) inv[i] = a[i] / det;
    return true;
}

}  // namespace

// WorldPosFromDepth (Util.glsl:1-7): world = inverse(view) * inverse(proj) * clip.
// The matrix is formed in double on the host and rounded to float once.
CamDev make_cam(const snkpt_camera* c) {
    double v[16], p[16], vi[16], pi[16];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            v[4 * r + k] = c->view[4 * k + r];  // column-major -> row-major
            p[4 * r + k] = c->proj[4 * k + r];
        }
    inverse4(v, vi);
    inverse4(p, pi);
    CamDev d;
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            double acc = 0;
            for (int j = 0; j < 4; ++j) acc += vi[4 * r + j] * pi[4 * j + k];
            d.m[4 * r + k] = (float)acc;
        }
    d.pos = V3{c->cam_pos[0], c->cam_pos[1], c->cam_pos[2]};
    return d;
}

SceneDev scene_dev(snkpt_ctx* ctx) {
    SceneDev S;
    S.inst = ctx->d_inst.as<InstD>();
    S.xf = ctx->d_inst_xf.as<float4>();
    S.tri_off = ctx->accel.inst_tri_offset.as<uint32_t>();
    S.mats = ctx->d_mats.as<MatD>();
    S.idx = ctx->d_idx.as<uint32_t>();
    S.nrm = ctx->d_nrm.as<float>();
    S.world = ctx->accel.world_tris.as<float4>();
    S.emit_rec = ctx->emit.rec.as<float4>();
    S.gid2e = ctx->emit.gid_to_emitter.as<uint32_t>();
    S.n_emit = ctx->emit.count;
    S.nodes = ctx->accel.nodes.as<uint4>();
    S.leaf = ctx->accel.leaf_tris.as<float4>();
    S.alights = ctx->emit.analytic.as<float4>();
    S.n_alights = ctx->emit.n_analytic;
    S.q_tri = ctx->emit.q_tri;
    return S;
}

#define HIPR(expr)                                                \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr);    \
    } while (0)

snkpt_status ensure_state(snkpt_ctx* ctx, uint32_t w, uint32_t h, hipStream_t st) {
    if (!ctx->render) ctx->render = new RenderState();
    RenderState& R = *ctx->render;
    const uint32_t n = w * h;
    if (n > R.cap) {
        for (int k = 0; k < 2; ++k) {
            HIPR(R.q[k].org.reserve((size_t)n * 16));
            HIPR(R.q[k].dir.reserve((size_t)n * 16));
            HIPR(R.q[k].path.reserve((size_t)n * 4));
        }
        HIPR(R.h_t.reserve((size_t)n * 4));
        HIPR(R.h_inst.reserve((size_t)n * 4));
        HIPR(R.h_prim.reserve((size_t)n * 4));
        HIPR(R.h_u.reserve((size_t)n * 4));
        HIPR(R.h_v.reserve((size_t)n * 4));
        HIPR(R.s_org.reserve((size_t)n * 16));
        HIPR(R.s_dir.reserve((size_t)n * 16));
        HIPR(R.s_contrib.reserve((size_t)n * 16));
        HIPR(R.p_thr.reserve((size_t)n * 16));
        HIPR(R.p_rng.reserve((size_t)n * 8));
        HIPR(R.p_bounce.reserve((size_t)n * 4));
        HIPR(R.rad.reserve((size_t)n * 16));
        HIPR(R.accum.reserve((size_t)n * 16));
        R.cap = n;
        ++ctx->generation;
    }
    HIPR(R.counts.reserve(64));
    HIPR(R.stats.reserve(8 * (kQueueHistAt + kQueueHist)));
    if (R.w != w || R.h != h) {
        R.w = w;
        R.h = h;
        R.tail_auto = 0;
        ++ctx->generation;
        k_fill_f4<<<nblk(n, kBlock), kBlock, 0, st>>>(R.accum.as<float4>(), n, make_float4(0, 0, 0, 0));
        R.accum_frames = 0;
        R.history_valid = false;
    }
    return SNKPT_OK;
}

Queue queue_of(RenderState& R, int k) {
    return Queue{R.q[k].org.as<float4>(), R.q[k].dir.as<float4>(), R.q[k].path.as<uint32_t>()};
}
HitArrays hits_of(RenderState& R) {
    return HitArrays{R.h_t.as<float>(), R.h_inst.as<uint32_t>(), R.h_prim.as<uint32_t>(), R.h_u.as<float>(),
                     R.h_v.as<float>()};
}
ShadowQueue shadow_of(RenderState& R) {
    return ShadowQueue{R.s_org.as<float4>(), R.s_dir.as<float4>(), R.s_contrib.as<float4>()};
}
PathState paths_of(RenderState& R) {
    return PathState{R.p_thr.as<float4>(), R.p_rng.as<uint2>(), R.p_bounce.as<uint32_t>()};
}

// The rest of every path queued in `cur` (up to n_max of them), one thread per
// path on a persistent grid.
snkpt_status launch_tail(snkpt_ctx* ctx, RenderState& R, const SceneDev& S, const PtParams& P, const ShadeHooks& H,
                         int cur, float4* rad, hipStream_t st, uint32_t n_max) {
    const Accel& A = ctx->accel;
    if (A.num_tris == 0) return SNKPT_OK;  // no hits, so nothing was queued
    const uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
    const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
    const uint32_t grid = std::min(nblk(n_max, kMegaBlock), cus * kTailBlocksPerCu);
    const uint32_t* count = R.counts.as<uint32_t>() + cur;
    unsigned long long* stats = R.stats.as<unsigned long long>();
#define SNKPT_TAIL(STK) \
    k_pt_tail<STK><<<grid, kMegaBlock, 0, st>>>(S, P, H, queue_of(R, cur), count, paths_of(R), rad, stats)
    if (need <= 10)
        SNKPT_TAIL(10);
    else if (need <= 16)
        SNKPT_TAIL(16);
    else if (need <= 32)
        SNKPT_TAIL(32);
    else if (need <= 64)
        SNKPT_TAIL(64);
    else
        return fail(ctx, SNKPT_ERR_CAPACITY, "path tail: BVH8 deeper than the 64-entry traversal stack");
#undef SNKPT_TAIL
    return SNKPT_OK;
}

snkpt_status launch_shadow(snkpt_ctx* ctx, RenderState& R, float4* rad, hipStream_t st, uint32_t n_max) {
    const Accel& A = ctx->accel;
    if (A.num_tris == 0) return SNKPT_OK;  // no hits, so no shadow rays were queued
    uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
    const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
    // persistent grid: as many blocks as the LDS of the stack size lets stay resident
    // (an even count: two waves per block spread evenly over the four SIMDs)
    const uint32_t resident = need <= 10 ? 16u : need <= 12 ? 12u : need <= 16 ? 10u : need <= 32 ? 4u : 2u;
    uint32_t grid = std::min(nblk(n_max ? n_max : R.w * R.h, kShadowBlock), cus * resident);
    if (need <= 10)
        k_shadow<10><<<grid, kShadowBlock, 0, st>>>(A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), shadow_of(R),
                                                     R.counts.as<uint32_t>(), rad);
    else if (need <= 12)
        k_shadow<12><<<grid, kShadowBlock, 0, st>>>(A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), shadow_of(R),
                                                     R.counts.as<uint32_t>(), rad);
    else if (need <= 16)
        k_shadow<16><<<grid, kShadowBlock, 0, st>>>(A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), shadow_of(R),
                                                     R.counts.as<uint32_t>(), rad);
    else if (need <= 32)
        k_shadow<32><<<grid, kShadowBlock, 0, st>>>(A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), shadow_of(R),
                                                     R.counts.as<uint32_t>(), rad);
    else if (need <= 64)
        k_shadow<64><<<grid, kShadowBlock, 0, st>>>(A.nodes.as<uint4>(), A.leaf_tris.as<float4>(), shadow_of(R),
                                                     R.counts.as<uint32_t>(), rad);
    else
        return fail(ctx, SNKPT_ERR_CAPACITY, "shadow: BVH8 deeper than the 64-entry traversal stack");
    return SNKPT_OK;
}

// Bounces first..max_depth-1 of the queued paths (queue `cur` already holds
// the rays of bounce `first`).
snkpt_status pt_bounces(snkpt_ctx* ctx, const snkpt_settings* s, int cur, uint32_t first, hipStream_t st,
                        const ShadeHooks* hooks, uint32_t max_depth_override, uint32_t n_paths) {
    RenderState& R = *ctx->render;
    const uint32_t n = n_paths ? std::min(n_paths, R.w * R.h) : R.w * R.h;
    SceneDev S = scene_dev(ctx);
    PtParams P{max_depth_override ? max_depth_override : (s->max_depth ? s->max_depth : 1u), s->rr_start};
    ShadeHooks H = hooks ? *hooks : ShadeHooks{};
    float4* rad = H.rad ? H.rad : R.rad.as<float4>();
    uint32_t* counts = R.counts.as<uint32_t>();
    unsigned long long* stats = R.stats.as<unsigned long long>();
    // bounce from which the surviving paths run one thread per path (0: never)
    uint32_t tail_from = 0;
    if (H.kind == 0) {
        const uint32_t t = s->tail_bounce ? s->tail_bounce
                           : R.tail_auto  ? R.tail_auto
                                          : std::max(1u, P.rr_start);
        tail_from = t < P.max_depth ? std::max(t, 1u) : 0u;
    }
    R.hist_first = first;
    R.hist_depth = P.max_depth;
    R.tail_used = 0;
    for (uint32_t b = first; b < P.max_depth; ++b) {
        int nxt = cur ^ 1;
        k_bounce_begin<<<1, 64, 0, st>>>(counts, (uint32_t)cur, (uint32_t)nxt, stats, b > first ? 1u : 0u,
                                         b > first ? 1u : 0u, b);
        if (tail_from && b >= tail_from) {
            snkpt_status e = launch_tail(ctx, R, S, P, H, cur, rad, st, n);
            if (e != SNKPT_OK) return e;
            R.tail_used = b;
            break;
        }
        snkpt_hits h{R.h_t.as<float>(), R.h_inst.as<uint32_t>(), R.h_prim.as<uint32_t>(), R.h_u.as<float>(),
                     R.h_v.as<float>()};
        snkpt_status e = launch_trace(ctx, R.q[cur].org.as<float4>(), R.q[cur].dir.as<float4>(), n, h, false,
                                      nullptr, nullptr, counts + cur, st);
        if (e != SNKPT_OK) return e;
        if (H.kind == 1)
            k_shade<1><<<nblk(n, kBlock), kBlock, 0, st>>>(S, P, queue_of(R, cur), hits_of(R), queue_of(R, nxt),
                                                           shadow_of(R), paths_of(R), rad, counts, (uint32_t)cur,
                                                           (uint32_t)nxt, stats, H);
        else if (H.kind == 2)
            k_shade<2><<<nblk(n, kBlock), kBlock, 0, st>>>(S, P, queue_of(R, cur), hits_of(R), queue_of(R, nxt),
                                                           shadow_of(R), paths_of(R), rad, counts, (uint32_t)cur,
                                                           (uint32_t)nxt, stats, H);
        else
            k_shade<0><<<nblk(n, kBlock), kBlock, 0, st>>>(S, P, queue_of(R, cur), hits_of(R), queue_of(R, nxt),
                                                           shadow_of(R), paths_of(R), rad, counts, (uint32_t)cur,
                                                           (uint32_t)nxt, stats, H);
        if ((e = launch_shadow(ctx, R, rad, st, n)) != SNKPT_OK) return e;
        cur = nxt;
    }
    if (first < P.max_depth) k_bounce_end<<<1, 64, 0, st>>>(counts, stats);
    return SNKPT_OK;
}

snkpt_status generate_camera(snkpt_ctx* ctx, const CamDev& cam, uint64_t initstate, hipStream_t st, uint32_t n_paths,
                             uint32_t start, uint32_t stride, uint32_t* nrec_clear, bool tiled) {
    RenderState& R = *ctx->render;
    const uint32_t n = n_paths ? n_paths : R.w * R.h;
    const uint32_t tl = (tiled && n == R.w * R.h && start == 0 && stride == 1) ? 1u : 0u;
    k_generate<<<nblk(n, kBlock), kBlock, 0, st>>>(cam, R.w, R.h, n, start, stride, tl, initstate,
                                                   R.q[0].org.as<float4>(), R.q[0].dir.as<float4>(),
                                                   R.q[0].path.as<uint32_t>(), paths_of(R), R.counts.as<uint32_t>(), 0,
                                                   nrec_clear);
    return SNKPT_OK;
}

snkpt_status primary_bounces(snkpt_ctx* ctx, const CamDev& cam, const snkpt_settings* s, uint64_t initstate,
                             hipStream_t st, const ShadeHooks* hooks, uint32_t max_depth_override) {
    RenderState& R = *ctx->render;
    if (!R.use_gb) {
        // the queue stages read every path's state by path index: tiles are free
        generate_camera(ctx, cam, initstate, st, 0, 0, 1, nullptr, true);
        return pt_bounces(ctx, s, 0, 0, st, hooks, max_depth_override);
    }
    const uint32_t n = R.w * R.h;
    SceneDev S = scene_dev(ctx);
    PtParams P{max_depth_override ? max_depth_override : (s->max_depth ? s->max_depth : 1u), s->rr_start};
    ShadeHooks H = hooks ? *hooks : ShadeHooks{};
    float4* rad = H.rad ? H.rad : R.rad.as<float4>();
    uint32_t* counts = R.counts.as<uint32_t>();
    unsigned long long* stats = R.stats.as<unsigned long long>();
    k_bounce_begin<<<1, 64, 0, st>>>(counts, 0u, 1u, stats, 0u, 0u, 0u);  // clears queue 1 and the shadow queue
    if (H.kind == 1)
        k_gshade<1><<<nblk(n, kBlock), kBlock, 0, st>>>(S, R.gb, cam, R.w, R.h, P, queue_of(R, 1), shadow_of(R),
                                                        paths_of(R), rad, counts, 1u, initstate, H);
    else if (H.kind == 2)
        k_gshade<2><<<nblk(n, kBlock), kBlock, 0, st>>>(S, R.gb, cam, R.w, R.h, P, queue_of(R, 1), shadow_of(R),
                                                        paths_of(R), rad, counts, 1u, initstate, H);
    else
        k_gshade<0><<<nblk(n, kBlock), kBlock, 0, st>>>(S, R.gb, cam, R.w, R.h, P, queue_of(R, 1), shadow_of(R),
                                                        paths_of(R), rad, counts, 1u, initstate, H);
    snkpt_status e = launch_shadow(ctx, R, rad, st, n);
    if (e != SNKPT_OK) return e;
    k_gbuf_stats<<<1, 64, 0, st>>>(counts, stats);
    if (P.max_depth > 1) return pt_bounces(ctx, s, 1, 1, st, hooks, max_depth_override, n);
    return SNKPT_OK;
}

// One pass of 1 sample per pixel of the plain path tracer, added into R.rad.
snkpt_status pt_pass(snkpt_ctx* ctx, const CamDev& cam, const snkpt_settings* s, uint64_t initstate,
                     hipStream_t st, uint32_t first_bounce_skip) {
    (void)first_bounce_skip;
    if ((s->flags & SNKPT_FLAG_WHOLE_PATHS) && !ctx->render->use_gb) {
        RenderState& R = *ctx->render;
        return whole_paths(ctx, cam, R.w, R.h, s, initstate, R.w * R.h, 0, 1, nullptr, R.rad.as<float4>(),
                           R.stats.as<unsigned long long>(), st);
    }
    return primary_bounces(ctx, cam, s, initstate, st);
}

snkpt_status whole_paths(snkpt_ctx* ctx, const CamDev& cam, uint32_t w, uint32_t h, const snkpt_settings* s,
                         uint64_t initstate, uint32_t n_paths, uint32_t start, uint32_t stride,
                         const ShadeHooks* hooks, float4* rad, unsigned long long* stats, hipStream_t st) {
    const Accel& A = ctx->accel;
    if (n_paths == 0) return SNKPT_OK;
    if (A.num_tris == 0) return SNKPT_OK;  // nothing to hit: radiance stays 0
    SceneDev S = scene_dev(ctx);
    PtParams P{s->max_depth ? s->max_depth : 1u, s->rr_start};
    const uint32_t need = A.depth > 0 ? A.depth - 1 : 0;
    const uint32_t grid = nblk(n_paths, kMegaBlock);
    const bool rec = hooks && hooks->kind == 2;
    ShadeHooks H = hooks ? *hooks : ShadeHooks{};
#define SNKPT_MEGA(STK)                                                                                        \
    do {                                                                                                       \
        if (rec)                                                                                               \
            k_pt_mega<2, STK><<<grid, kMegaBlock, 0, st>>>(S, cam, w, h, n_paths, start, stride, initstate,     \
                                                           P, H, rad, stats);                                  \
        else                                                                                                   \
            k_pt_mega<0, STK><<<grid, kMegaBlock, 0, st>>>(S, cam, w, h, n_paths, start, stride, initstate,     \
                                                           P, H, rad, stats);                                  \
    } while (0)
    if (need <= 16)
        SNKPT_MEGA(16);
    else if (need <= 32)
        SNKPT_MEGA(32);
    else if (need <= 64)
        SNKPT_MEGA(64);
    else
        return fail(ctx, SNKPT_ERR_CAPACITY, "whole paths: BVH8 deeper than the 64-entry traversal stack");
#undef SNKPT_MEGA
    return SNKPT_OK;
}

snkpt_status begin_frame(snkpt_ctx* ctx, uint32_t w, uint32_t h, hipStream_t st) {
    snkpt_status e = ensure_state(ctx, w, h, st);
    if (e != SNKPT_OK) return e;
    RenderState& R = *ctx->render;
    const uint32_t n = w * h;
    k_fill_f4<<<nblk(n, kBlock), kBlock, 0, st>>>(R.rad.as<float4>(), n, make_float4(0, 0, 0, 0));
    return SNKPT_OK;
}

// The per-thread tail's bounce for the next frame: the first bounce (>= 1) whose
// queue of paths fits twice the tail's resident lanes.  q: this frame's queue
// lengths, known for the bounces the wavefront began (up to the tail's start);
// when the tail started on a longer queue, the next frame tries one bounce later.
static void tail_feedback(snkpt_ctx* ctx, const unsigned long long* q) {
    RenderState& R = *ctx->render;
    if (R.hist_depth == 0) return;  // no path-traced bounces this frame
    const uint32_t cus = ctx->props.multiProcessorCount > 0 ? (uint32_t)ctx->props.multiProcessorCount : 256u;
    const unsigned long long fit = 2ull * cus * kTailBlocksPerCu * kMegaBlock;
    const uint32_t last = std::min(R.tail_used ? R.tail_used : R.hist_depth - 1, kQueueHist - 1);
    uint32_t pick = R.tail_used && R.tail_used + 1 < R.hist_depth ? R.tail_used + 1 : kTailNever;
    for (uint32_t b = std::max(R.hist_first, 1u); b <= last; ++b)
        if (q[b] <= fit) {
            pick = b;
            break;
        }
    R.tail_auto = pick;
    R.hist_depth = 0;
}

static snkpt_status read_stats(snkpt_ctx* ctx, hipStream_t st, uint64_t camera_rays) {
    RenderState& R = *ctx->render;
    unsigned long long h[kQueueHistAt + kQueueHist];
    HIPR(hipMemcpyAsync(h, R.stats.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    tail_feedback(ctx, h + kQueueHistAt);
    ctx->stats.camera_rays = camera_rays;
    ctx->stats.extension_rays = h[1];
    ctx->stats.shadow_rays = h[2];
    ctx->stats.cache_queries = h[3];
    ctx->stats.cache_train_samples = h[4];
    ctx->stats.restir_history_reused = h[5];
    return SNKPT_OK;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

snkpt_status render(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* s, void* out, uint32_t w,
                    uint32_t h, uint32_t fmt, hipStream_t st, const snkpt_gbuffer* gb) {
    if (!ctx->accel.built) return fail(ctx, SNKPT_ERR_NOT_BUILT, "snkpt_render: call snkpt_build_accel first");
    if ((uint64_t)w * h > (1ull << 26)) return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_render: image above 2^26 pixels");
    if (s->mode > SNKPT_MODE_REFERENCE_COMPAT) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render: unknown mode");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cap) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipStreamIsCapturing");
    if (cap != hipStreamCaptureStatusNone) {
        // a captured frame may launch kernels only: no allocation, no readback,
        // no host-side state that a replay would not repeat
        if ((s->mode != SNKPT_MODE_PT && s->mode != SNKPT_MODE_CACHE) || !(s->flags & SNKPT_FLAG_NO_STATS))
            return fail(ctx, SNKPT_ERR_INVALID_ARG,
                        "snkpt_render: under stream capture only SNKPT_MODE_PT or SNKPT_MODE_CACHE with "
                        "SNKPT_FLAG_NO_STATS");
        if (!ctx->render || ctx->render->w != w || ctx->render->h != h || ctx->render->counts.ptr == nullptr ||
            ctx->render->stats.ptr == nullptr)
            return fail(ctx, SNKPT_ERR_INVALID_ARG,
                        "snkpt_render: render one uncaptured frame of this size (the last one) before capturing");
        if (s->mode == SNKPT_MODE_CACHE && !cache_capture_ready(ctx, s, w, h))
            return fail(ctx, SNKPT_ERR_INVALID_ARG,
                        "snkpt_render: render one uncaptured cache-mode frame with these settings before capturing");
    }
    snkpt_status e = begin_frame(ctx, w, h, st);
    if (e != SNKPT_OK) return e;
    RenderState& R = *ctx->render;
    // the G-buffer is borrowed for this call only
    struct GbScope {
        RenderState& R;
        ~GbScope() {
            R.use_gb = false;
            R.gb = GbufDev{};
        }
    } gb_scope{R};
    if (gb) {
        R.use_gb = true;
        R.gb = GbufDev{gb->depth, gb->normal, gb->albedo, gb->rma, gb->mat_flags, gb->motion};
    }
    k_zero_u64<<<1, 64, 0, st>>>(R.stats.as<unsigned long long>(), 8);
    if (s->flags & SNKPT_FLAG_RESET) {
        k_fill_f4<<<nblk(w * h, kBlock), kBlock, 0, st>>>(R.accum.as<float4>(), w * h, make_float4(0, 0, 0, 0));
        R.accum_frames = 0;
    }
    CamDev cd = make_cam(cam);
    const uint32_t spp = s->spp ? s->spp : 1u;
    for (uint32_t k = 0; k < spp; ++k) {
        uint64_t is = pass_state(s->seed, s->frame, k);
        if (s->mode == SNKPT_MODE_PT)
            e = pt_pass(ctx, cd, s, is, st, 0);
        else if (s->mode == SNKPT_MODE_RESTIR || s->mode == SNKPT_MODE_REFERENCE_COMPAT)
            e = restir_pass(ctx, cam, cd, s, is, k, st);
        else
            e = cache_pass(ctx, cd, s, is, k, st);
        if (e != SNKPT_OK) return e;
    }
    k_resolve<<<nblk(w * h, kBlock), kBlock, 0, st>>>(R.rad.as<float4>(), R.accum.as<float4>(), out, w * h,
                                                      (float)spp, (s->flags & SNKPT_FLAG_ACCUMULATE) ? 1u : 0u, fmt,
                                                      s->accum_weight > 0.0f ? s->accum_weight : 1.0f);
    if (s->flags & SNKPT_FLAG_ACCUMULATE) R.accum_frames += spp;
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(ctx, le, "render launch");
    if (s->flags & SNKPT_FLAG_NO_STATS) return SNKPT_OK;
    return read_stats(ctx, st, gb ? 0ull : (uint64_t)w * h * spp);
}

snkpt_status oracle_accumulate(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* s, double* accum,
                               uint32_t w, uint32_t h, uint32_t spp, hipStream_t st) {
    if (!ctx->accel.built) return fail(ctx, SNKPT_ERR_NOT_BUILT, "oracle_accumulate: build the accel first");
    if ((uint64_t)w * h > (1ull << 26)) return fail(ctx, SNKPT_ERR_CAPACITY, "oracle_accumulate: image too large");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    CamDev cd = make_cam(cam);
    snkpt_status e = ensure_state(ctx, w, h, st);
    if (e != SNKPT_OK) return e;
    RenderState& R = *ctx->render;
    for (uint32_t k = 0; k < spp; ++k) {
        k_fill_f4<<<nblk(w * h, kBlock), kBlock, 0, st>>>(R.rad.as<float4>(), w * h, make_float4(0, 0, 0, 0));
        // oracle streams are disjoint from render streams: seed ^ "orac"
        if ((e = pt_pass(ctx, cd, s, pass_state(s->seed ^ 0x6F726163u, s->frame, k), st, 0)) != SNKPT_OK) return e;
        k_add_f64<<<nblk(w * h, kBlock), kBlock, 0, st>>>(R.rad.as<float4>(), accum, w * h);
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(ctx, le, "oracle_accumulate launch");
    return SNKPT_OK;
}

void render_state_free(RenderState* r) { delete r; }

}  // namespace snkpt
