// Wavefront state shared by the path-trace, ReSTIR and cache passes.
#pragma once

#include "cache_nn.h"
#include "render_common.h"

namespace snkpt {

struct CamDev {
    float m[16];  // inverse(view) * inverse(proj), row-major
    V3 pos;
};

// Camera ray through the centre of pixel (x, y), row 0 at the top (Util.glsl:9-13).
__device__ __forceinline__ V3 camera_dir(const CamDev& c, uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    float u = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
    float v = ((float)y + 0.5f) / (float)h * 2.0f - 1.0f;
    float wp[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) wp[r] = ((c.m[4 * r] * u + c.m[4 * r + 1] * v) + c.m[4 * r + 2]) + c.m[4 * r + 3];
    float iw = 1.0f / fmaxf(wp[3], 1e-6f);
    V3 p{wp[0] * iw, wp[1] * iw, wp[2] * iw};
    return normalize(p - c.pos);
}

// Rasterized G-buffer of the frame (device pointers, one texel per pixel):
// D32 depth, RGBA16F normal, RGBA8 SNORM albedo, RGBA16F roughness /
// metallic / ao, R16 material flags, optional RG16F motion.
struct GbufDev {
    const float* depth = nullptr;
    const uint16_t* normal = nullptr;
    const int8_t* albedo = nullptr;
    const uint16_t* rma = nullptr;
    const uint16_t* flags = nullptr;
    const uint16_t* motion = nullptr;
};

__device__ __forceinline__ float half_bits(uint16_t b) {
    _Float16 h;
    __builtin_memcpy(&h, &b, 2);
    return (float)h;
}
// R8_SNORM texel to float: max(c / 127, -1)
__device__ __forceinline__ float snorm8(int8_t c) { return fmaxf((float)c / 127.0f, -1.0f); }

// The primary vertex of pixel px from the G-buffer (RayGen.rgen:531-552):
// world position from depth through inverse(view) * inverse(proj)
// (WorldPosFromDepth, Util.glsl:1-7, the camera-ray unprojection with the
// texel's depth for 1), the texel's normal as shading and geometric normal
// (turned to face the camera), material from albedo / rma / flags, and the
// reference's ray bias 0.01 + max(0.001 |cam - p|, 0.01).  False where no
// surface was rasterized (depth >= 1) or the normal is zero.
__device__ __forceinline__ bool gbuffer_surface(const GbufDev& g, const CamDev& c, uint32_t px, uint32_t w,
                                                uint32_t h, Surf& s, V3& dir, float& t) {
    const float depth = g.depth[px];
    if (!(depth < 1.0f)) return false;
    const uint32_t x = px % w, y = px / w;
    const float u = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
    const float v = ((float)y + 0.5f) / (float)h * 2.0f - 1.0f;
    float wp[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) wp[r] = ((c.m[4 * r] * u + c.m[4 * r + 1] * v) + c.m[4 * r + 2] * depth) + c.m[4 * r + 3];
    const float iw = 1.0f / fmaxf(wp[3], 1e-6f);
    s.p = V3{wp[0] * iw, wp[1] * iw, wp[2] * iw};
    const V3 d = s.p - c.pos;
    const float l2 = dot(d, d);
    if (!(l2 > 0.0f)) return false;
    t = sqrtf(l2);
    dir = d * (1.0f / t);
    const uint2 nb = *(const uint2*)(g.normal + 4 * (size_t)px);
    V3 n{half_bits((uint16_t)(nb.x & 0xFFFFu)), half_bits((uint16_t)(nb.x >> 16)), half_bits((uint16_t)(nb.y & 0xFFFFu))};
    const float n2 = dot(n, n);
    if (!(n2 > 1e-20f)) return false;
    n = n * (1.0f / sqrtf(n2));
    if (dot(n, dir) > 0.0f) n = n * -1.0f;
    s.ng = n;
    s.ns = n;
    s.nraw = n;
    const char4 a = *(const char4*)(g.albedo + 4 * (size_t)px);
    s.m.albedo = V3{fminf(fmaxf(snorm8(a.x), 0.0f), 1.0f), fminf(fmaxf(snorm8(a.y), 0.0f), 1.0f),
                    fminf(fmaxf(snorm8(a.z), 0.0f), 1.0f)};
    const uint2 rb = *(const uint2*)(g.rma + 4 * (size_t)px);
    s.m.roughness = fminf(fmaxf(half_bits((uint16_t)(rb.x & 0xFFFFu)), 0.02f), 1.0f);
    s.m.metallic = fminf(fmaxf(half_bits((uint16_t)(rb.x >> 16)), 0.0f), 1.0f);
    s.m.flags = g.flags[px];
    s.emissive = (s.m.flags & 1u) != 0u;
    s.le = s.m.albedo;
    s.gid = 0xFFFFFFFFu;
    s.eps = 0.01f + fmaxf(t * 0.001f, 0.01f);
    return true;
}

struct PtParams {
    uint32_t max_depth, rr_start;
};
struct Queue {
    float4* org;
    float4* dir;
    uint32_t* path;
};
struct HitArrays {
    float* t;
    uint32_t* inst;
    uint32_t* prim;
    float* u;
    float* v;
};
struct ShadowQueue {
    float4* org;
    float4* dir;
    float4* contrib;  // rgb, pixel (bits)
};
struct PathState {
    float4* thr;       // throughput rgb, pdf of the last BRDF sample
    uint2* rng;        // PCG32 state (the stream is the pixel index)
    uint32_t* bounce;  // path vertex index of the next hit
};

// Optional behaviour of the shade kernel for the cache mode.
//   kind 0: plain path tracing
//   kind 1: at vertex `query_at`, a non-emissive hit writes a cache query
//           (position, direction, normal, roughness, albedo) and the path ends;
//           with `indirect` the vertex first takes its direct light by NEE
//           (weight 1) and the cache stands for the indirect part only
//   kind 2: every non-emissive vertex below `max_rec` is recorded for training
//           (sample fields, the radiance accumulated so far, the throughput)
struct ShadeHooks {
    uint32_t kind = 0;
    uint32_t query_at = 0xFFFFFFFFu;
    uint32_t indirect = 0;
    uint32_t mis = 0;            // with indirect: NEE and a traced BRDF probe under MIS at the query vertex
    float* q_samples = nullptr;  // 16 floats per query
    float4* q_meta = nullptr;    // throughput rgb, path (bits)
    uint32_t* q_count = nullptr;
    float* rec = nullptr;        // kRecFloats per (path, vertex)
    uint32_t* nrec = nullptr;    // recorded vertices per path
    uint32_t max_rec = 0;
    float4* rad = nullptr;       // radiance target (defaults to the frame radiance)
    const struct TrainSeedD* tseed = nullptr;  // training paths: RNG state and first pixel read on the device
};

// Random-stream state of a frame's training paths, computed on the device from
// the frame index when the frame is a replayed graph (batched cache form).
struct TrainSeedD {
    uint64_t path_state;  // RNG state of the paths (the frame's pass state ^ the training-path tag)
    uint64_t kseed;       // seed of the vertex selection (the frame's pass state ^ the target tag)
    uint32_t start;       // first pixel of the strided subset
    uint32_t pad[3];
};
constexpr int kRecFloats = 20;  // 16 sample floats (target = radiance so far) + throughput rgb + pad

struct QueueBufs {
    DeviceBuffer org, dir, path;
};

struct RestirState;
void restir_state_free(RestirState* r);

struct RenderState {
    uint32_t cap = 0, w = 0, h = 0;
    QueueBufs q[2];
    DeviceBuffer h_t, h_inst, h_prim, h_u, h_v;
    DeviceBuffer s_org, s_dir, s_contrib;
    DeviceBuffer p_thr, p_rng, p_bounce;
    DeviceBuffer rad, accum, counts, stats;
    uint32_t accum_frames = 0;
    // per-thread tail of the path tracer: the bounce chosen from the last frame
    // read back (0 = none yet, 0xFFFFFFFF = never), and what the last bounce
    // loop did (its first bounce, its depth (0 = none since the last readback),
    // the bounce its tail started at (0 = no tail))
    uint32_t tail_auto = 0, hist_first = 0, hist_depth = 0, tail_used = 0;
    bool history_valid = false;
    bool use_gb = false;  // this frame's primary vertices come from `gb`
    GbufDev gb;
    RestirState* restir = nullptr;
    ~RenderState() { restir_state_free(restir); }
};

CamDev make_cam(const snkpt_camera* c);
SceneDev scene_dev(snkpt_ctx* ctx);
snkpt_status ensure_state(snkpt_ctx* ctx, uint32_t w, uint32_t h, hipStream_t st);
Queue queue_of(RenderState& R, int k);
HitArrays hits_of(RenderState& R);
ShadowQueue shadow_of(RenderState& R);
PathState paths_of(RenderState& R);
// n_max: upper bound of the queued shadow rays (0 = one per pixel)
snkpt_status launch_shadow(snkpt_ctx* ctx, RenderState& R, float4* rad, hipStream_t st, uint32_t n_max = 0);
snkpt_status pt_pass(snkpt_ctx* ctx, const CamDev& cam, const snkpt_settings* s, uint64_t initstate,
                     hipStream_t st, uint32_t first_bounce_skip);
// n_paths: number of paths generated (0 = one per pixel); sizes the launches.
snkpt_status pt_bounces(snkpt_ctx* ctx, const snkpt_settings* s, int cur, uint32_t first, hipStream_t st,
                        const ShadeHooks* hooks = nullptr, uint32_t max_depth_override = 0, uint32_t n_paths = 0);
// Whole paths, one per thread: path j < n_paths through pixel (start + j * stride)
// mod (w h), radiance added into rad[j]; hooks of kind 2 record training vertices.
snkpt_status whole_paths(snkpt_ctx* ctx, const CamDev& cam, uint32_t w, uint32_t h, const snkpt_settings* s,
                         uint64_t initstate, uint32_t n_paths, uint32_t start, uint32_t stride,
                         const ShadeHooks* hooks, float4* rad, unsigned long long* stats, hipStream_t st);
// cache.hip: cache query and training for callers that trace the render paths themselves (the batch)
snkpt_status cache_ensure(snkpt_ctx* ctx, const snkpt_settings* s, hipStream_t st);
// the context's network for the multi-network queries (weight averages with `averaged`)
snkpt_status cache_ref(snkpt_ctx* ctx, bool averaged, CacheRefD* out);
// queries grouped by network: rows [off[k], off[k+1]) (64-row multiples, off[nseg]
// rows in all, at most rows_max) through network refs[k]; x0 holds rows_max x 64 halves
void cache_query_many(const CacheRefD* refs, const uint32_t* off, uint32_t nseg, const float* samples,
                      uint32_t rows_max, _Float16* x0, float4* y, hipStream_t st);
// every buffer of a frame's training half for n_paths paths through a w x h view
snkpt_status cache_prepare_train(snkpt_ctx* ctx, const snkpt_settings* s, uint32_t w, uint32_t h, uint32_t n_paths,
                                 hipStream_t st);
// tags of the training paths' random streams (TrainSeedD)
constexpr uint64_t kTrainPathsTag = 0x545241494eull;
constexpr uint64_t kTrainTargetTag = 0x7461726765ull;
// a cache-mode frame with these settings would allocate nothing (it may be captured)
bool cache_capture_ready(snkpt_ctx* ctx, const snkpt_settings* s, uint32_t w, uint32_t h);
snkpt_status cache_query(snkpt_ctx* ctx, const float* samples, uint32_t n, float4* y, bool averaged,
                         hipStream_t st);
// Training paths of a context enqueued on `side` (statistics kept per context),
// then, for several contexts at once, one readback of their sample counts and
// each one's optimiser step on `st` (their statistics added into `stats`).
snkpt_status cache_train_paths_forked(snkpt_ctx* ctx, const CamDev& cd, uint32_t w, uint32_t h,
                                      const snkpt_settings* s, uint64_t initstate, uint32_t n_paths,
                                      hipStream_t side, const TrainSeedD* tseed);
snkpt_status cache_train_steps(snkpt_ctx* const* ctxs, uint32_t n, const snkpt_settings* s,
                               unsigned long long* stats, hipStream_t st);
// Bounces 0.. of one path per pixel: camera rays through the wavefront, or,
// when the frame has a G-buffer, bounce 0 read from it and bounces 1.. traced.
snkpt_status primary_bounces(snkpt_ctx* ctx, const CamDev& cam, const snkpt_settings* s, uint64_t initstate,
                             hipStream_t st, const ShadeHooks* hooks = nullptr, uint32_t max_depth_override = 0);
// Camera rays for paths j < n_paths through pixels (start + j * stride) mod (w h).
// With `tiled` and a whole frame the queue holds the paths in 8 x 8 pixel tiles
// (queue slot != path); without it slot j holds path j, as per-pixel kernels
// that read the primary hits by pixel need.
snkpt_status generate_camera(snkpt_ctx* ctx, const CamDev& cam, uint64_t initstate, hipStream_t st,
                             uint32_t n_paths = 0, uint32_t start = 0, uint32_t stride = 1,
                             uint32_t* nrec_clear = nullptr, bool tiled = false);
inline uint32_t grid_for(uint32_t n, uint32_t b) { return n == 0 ? 1 : (n + b - 1) / b; }

// restir.hip / cache.hip
snkpt_status restir_pass(snkpt_ctx* ctx, const snkpt_camera* cam, const CamDev& cd, const snkpt_settings* s,
                         uint64_t initstate, uint32_t pass, hipStream_t st);
snkpt_status cache_pass(snkpt_ctx* ctx, const CamDev& cd, const snkpt_settings* s, uint64_t initstate,
                        uint32_t pass, hipStream_t st);

}  // namespace snkpt
