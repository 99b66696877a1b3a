// Batched path tracing: many views of many scenes of uneven sizes in one set
// of wavefront queues (the uneven batched workload; no reference counterpart:
// RT::RecordRenderCmdBuf, Raytracing.cpp:227-267, renders one view of one scene
// per call).
//
// All views of a group share the queues: path p of the group belongs to view
// J (path_off(J) <= p < path_off(J + 1)); its scene gives the BVH, triangles,
// materials and emitters the kernels use for that lane.  Each stage is ONE
// persistent launch per bounce over the whole group (grid = CUs x blocks per
// CU, grid-stride loops), so a 128^2 view costs no launches of its own and the
// big views balance over the whole chip.  A frame can be captured once as a
// hipGraph and replayed: per-frame values (frame index, accumulation flags)
// are read from device memory, never baked into the captured launches.
//
// Per-path results are bit-identical to snkpt_render of the same view alone:
// the same shade step runs with the same RNG stream (the pixel index within
// the view) and the same per-view seed.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pt_step.h"
#include "render_modes.h"
#include "snkpt_host.h"

namespace snkpt {
namespace {

constexpr int kBBlock = 256;        // generate / shade / resolve blocks
constexpr uint32_t kMaxViews = 65536;

struct JobD {
    CamDev cam;
    uint32_t scene, w, h;
    uint32_t path_off;   // first path of the view within its group
    uint32_t accum_off;  // first pixel of the view in the accumulation buffer
    uint32_t seed, fmt, pad;
    void* out;
};

struct FrameD {
    uint32_t frame;
    uint32_t accumulate;  // add this frame into the running mean and output the mean
    uint32_t reset;       // clear the running mean first
    uint32_t pad;
};

// Scene source of a batched launch: the view of a path selects its scene; the
// RNG stream of a path is its pixel index within the view.
struct SceneMany {
    const SceneDev* scenes;
    const JobD* jobs;
    const uint32_t* path_job;
    __device__ __forceinline__ const SceneDev& at(uint32_t path) const {
        return scenes[jobs[path_job[path]].scene];
    }
    __device__ __forceinline__ uint32_t stream(uint32_t path) const { return path - jobs[path_job[path]].path_off; }
};

__global__ void k_bzero(unsigned long long* stats, uint32_t n) {
    if (threadIdx.x < n) stats[threadIdx.x] = 0ull;
}

// Camera rays of views [j0, j1) (jobs sorted by path_off, group-relative).
__global__ __launch_bounds__(kBBlock) void k_bgenerate(const JobD* __restrict__ jobs, uint32_t j0, uint32_t j1,
                                                       uint32_t total, const FrameD* __restrict__ fr,
                                                       float4* __restrict__ q_org, float4* __restrict__ q_dir,
                                                       uint32_t* __restrict__ q_path, PathState ps,
                                                       uint32_t* __restrict__ path_job, float4* __restrict__ rad,
                                                       uint32_t* __restrict__ counts) {
    const uint32_t frame = fr->frame;
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[0] = total;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        // last view with path_off <= p
        uint32_t lo = j0, hi = j1;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].path_off <= p)
                lo = mid;
            else
                hi = mid;
        }
        const JobD& J = jobs[lo];
        const uint32_t px = p - J.path_off;
        V3 d = camera_dir(J.cam, px % J.w, px / J.w, J.w, J.h);
        q_org[p] = make_float4(J.cam.pos.x, J.cam.pos.y, J.cam.pos.z, 0.0f);
        q_dir[p] = make_float4(d.x, d.y, d.z, 1e30f);
        q_path[p] = p;
        Rng r = rng_seed(pass_state(J.seed, frame, 0), px);
        ps.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        ps.rng[p] = make_uint2((uint32_t)r.state, (uint32_t)(r.state >> 32));
        ps.bounce[p] = 0;
        path_job[p] = lo;
        rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_bbounce(uint32_t* c, uint32_t cur, uint32_t nxt, unsigned long long* stats, uint32_t count) {
    if (threadIdx.x == 0) {
        if (count) {
            stats[1] += c[cur];
            stats[2] += c[2];
        }
        c[nxt] = 0;
        c[2] = 0;
    }
}
__global__ void k_bbounce_end(uint32_t* c, unsigned long long* stats) {
    if (threadIdx.x == 0) {
        stats[2] += c[2];
        c[2] = 0;
    }
}

// Closest hit of the queued rays, each against its own scene's BVH8:
// persistent waves with ray refill over a contiguous slice each.
template <int STACK, int TB>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(STACK <= 10 ? 8 : 1, 8))) void k_btrace(SceneMany ss, const float4* __restrict__ org,
                                               const float4* __restrict__ dir, const uint32_t* __restrict__ q_path,
                                               const uint32_t* __restrict__ count, HitArrays hit) {
    __shared__ uint2 stk[STACK * TB];
    const uint32_t n = *count;
    const uint32_t nwaves = gridDim.x * (TB / 64);
    const uint32_t wave = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    auto load = [&](uint32_t i, Ray& ray, const uint4*& nd, const float4*& tr) {
        float4 a = org[i], d = dir[i];
        ray = Ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
        const SceneDev& S = ss.at(q_path[i]);
        nd = S.nodes;
        tr = S.leaf;
    };
    auto done = [&](uint32_t i, const TravState& st, uint32_t, uint32_t) {
        const TraceResult& res = st.res;
        if (res.rec != 0xFFFFFFFFu) {
            const float4* tris = ss.at(q_path[i]).leaf;
            float4 r2 = tris[(size_t)res.rec * kTriWords + 2];
            float4 a = org[i], d = dir[i];
            Ray ray{a.x, a.y, a.z, a.w, d.x, d.y, d.z, d.w};
            float u, v;
            hit_uv(tris, res.rec, ray, u, v);
            hit.t[i] = res.t;
            hit.inst[i] = __float_as_uint(r2.y);
            hit.prim[i] = __float_as_uint(r2.z);
            hit.u[i] = u;
            hit.v[i] = v;
        } else {
            hit.t[i] = -1.0f;
            hit.inst[i] = 0xFFFFFFFFu;
            hit.prim[i] = 0xFFFFFFFFu;
            hit.u[i] = 0.0f;
            hit.v[i] = 0.0f;
        }
    };
    refill_trace<false, false, STACK, kRefillLanes>(n, wave, nwaves, &stk[threadIdx.x], TB, load, done);
}

// The late bounces of the batch, one thread per path (tail_paths).
constexpr int kBTailBlock = 128;
constexpr uint32_t kBTailBlocksPerCu = 8;
template <int STACK>
__global__ __launch_bounds__(kBTailBlock) void k_btail(SceneMany ss, PtParams P, Queue in,
                                                       const uint32_t* __restrict__ count, PathState ps,
                                                       float4* __restrict__ rad,
                                                       unsigned long long* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kBTailBlock];
    __shared__ uint32_t red[2];
    tail_paths<STACK, kBTailBlock>(ss, P, ShadeHooks{}, in, *count, ps, rad, &stk[threadIdx.x], stats, red);
}

template <int KIND>
__global__ __launch_bounds__(kBBlock) void k_bshade(SceneMany ss, PtParams P, Queue in, HitArrays hit, Queue out,
                                                    ShadowQueue sq, PathState ps, float4* __restrict__ rad,
                                                    uint32_t* __restrict__ counts, uint32_t cur, uint32_t nxt,
                                                    ShadeHooks H) {
    __shared__ uint32_t lds_app[kBBlock / 64 + 1];
    const uint32_t n = counts[cur];
    // block-uniform loop: the appends inside the step are block-wide
    for (uint32_t base = blockIdx.x * kBBlock; base < n; base += gridDim.x * kBBlock) {
        const uint32_t i = base + threadIdx.x;
        shade_step<KIND, kBBlock>(ss, P, in, hit, out, sq, ps, rad, counts, nxt, H, i, i < n, lds_app);
    }
}

// Cache queries of the batch grouped by scene (each scene's network answers its
// own contiguous segment).  Per block: an LDS histogram over the scenes, one
// global atomic per (block, scene) to reserve room, ranks from LDS atomics.
constexpr uint32_t kMaxHistScenes = 1024;
__global__ __launch_bounds__(kBBlock) void k_bq_hist(const float4* __restrict__ meta, const uint32_t* __restrict__ nq,
                                                     const JobD* __restrict__ jobs,
                                                     const uint32_t* __restrict__ path_job, uint32_t nscenes,
                                                     uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kMaxHistScenes];
    for (uint32_t k = threadIdx.x; k < nscenes; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    const uint32_t n = *nq;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&h[jobs[path_job[__float_as_uint(meta[i].w)]].scene], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nscenes; k += blockDim.x)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}

__global__ __launch_bounds__(kBBlock) void k_bq_scatter(const float* __restrict__ samples,
                                                        const float4* __restrict__ meta,
                                                        const uint32_t* __restrict__ nq, const JobD* __restrict__ jobs,
                                                        const uint32_t* __restrict__ path_job, uint32_t nscenes,
                                                        uint32_t* __restrict__ cursor, float* __restrict__ s_out,
                                                        float4* __restrict__ m_out) {
    __shared__ uint32_t h[kMaxHistScenes];
    for (uint32_t k = threadIdx.x; k < nscenes; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    const uint32_t n = *nq;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t sc = 0, rank = 0;
    if (i < n) {
        sc = jobs[path_job[__float_as_uint(meta[i].w)]].scene;
        rank = atomicAdd(&h[sc], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nscenes; k += blockDim.x)
        if (h[k]) h[k] = atomicAdd(&cursor[k], h[k]);
    __syncthreads();
    if (i < n) {
        const uint32_t slot = h[sc] + rank;
        const float4* src = (const float4*)(samples + (size_t)i * 16);
        float4* dst = (float4*)(s_out + (size_t)slot * 16);
#pragma unroll
        for (int c = 0; c < 4; ++c) dst[c] = src[c];
        m_out[slot] = meta[i];
    }
}

__global__ void k_bq_clear(uint32_t* __restrict__ hist, uint32_t n) {
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) hist[k] = 0u;
}

// Segments of the sorted queries: scene k's hist[k] queries start at off[k]
// (each segment rounded up to 64 rows, so every wave of the query kernels
// serves one scene); off[nscenes] = all rows.  One block of kMaxHistScenes threads.
__global__ __launch_bounds__(kMaxHistScenes) void k_bq_layout(const uint32_t* __restrict__ hist, uint32_t nscenes,
                                                             uint32_t* __restrict__ off,
                                                             uint32_t* __restrict__ cursor) {
    __shared__ uint32_t v[kMaxHistScenes];
    const uint32_t k = threadIdx.x;
    const uint32_t mine = k < nscenes ? ((hist[k] + 63u) & ~63u) : 0u;
    v[k] = mine;
    __syncthreads();
    for (uint32_t d = 1; d < kMaxHistScenes; d <<= 1) {  // inclusive scan (Hillis-Steele)
        const uint32_t add = k >= d ? v[k - d] : 0u;
        __syncthreads();
        v[k] += add;
        __syncthreads();
    }
    if (k < nscenes) {
        off[k] = v[k] - mine;
        cursor[k] = v[k] - mine;
    }
    if (k == nscenes - 1) off[nscenes] = v[k];
}

// The padding rows of every segment: zero samples (finite inputs for the
// encoder) whose records name no path.
__global__ void k_bq_pad(const uint32_t* __restrict__ hist, const uint32_t* __restrict__ off,
                         float* __restrict__ s_out, float4* __restrict__ m_out) {
    const uint32_t k = blockIdx.x;
    const uint32_t row = off[k] + hist[k] + threadIdx.x;
    if (row >= off[k + 1]) return;
    float4* dst = (float4*)(s_out + (size_t)row * 16);
#pragma unroll
    for (int c = 0; c < 4; ++c) dst[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    m_out[row] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0xFFFFFFFFu));
}

// Training-path streams of each trained scene for this frame (its first view's
// seed, the frame index from device memory): what the eager path computes on the host.
__global__ void k_btrain_seeds(const uint2* __restrict__ cfg, uint32_t n, const FrameD* __restrict__ fr,
                               TrainSeedD* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t is = pass_state(cfg[k].x, fr->frame, 0);
    TrainSeedD t;
    t.path_state = is ^ kTrainPathsTag;
    t.kseed = is ^ kTrainTargetTag;
    t.start = (uint32_t)((is >> 17) % cfg[k].y);
    t.pad[0] = t.pad[1] = t.pad[2] = 0u;
    out[k] = t;
}

// rad[path] += throughput * max(cache output, 0) for the sorted queries (rows
// of the segments; padding rows name no path).
__global__ void k_bq_apply(const float4* __restrict__ meta, const float4* __restrict__ y, uint32_t n,
                           const uint32_t* __restrict__ n_dev, float4* __restrict__ rad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(n, *n_dev)) return;
    const float4 m = meta[i], v = y[i];
    const uint32_t path = __float_as_uint(m.w);
    if (path == 0xFFFFFFFFu) return;
    const float4 r = rad[path];
    rad[path] = make_float4(r.x + m.x * fmaxf(v.x, 0.0f), r.y + m.y * fmaxf(v.y, 0.0f), r.z + m.z * fmaxf(v.z, 0.0f),
                            r.w);
}

template <int STACK, int TB>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(STACK <= 10 ? 8 : 1, 8))) void k_bshadow(SceneMany ss, ShadowQueue sq, const uint32_t* __restrict__ counts,
                                                float4* __restrict__ rad) {
    __shared__ uint2 stk[STACK * TB];
    const uint32_t n = counts[2];
    const uint32_t nwaves = gridDim.x * (TB / 64);
    const uint32_t wave = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    auto load = [&](uint32_t i, Ray& ray, const uint4*& nd, const float4*& tr) {
        ray = shadow_ray(sq, i);
        const SceneDev& S = ss.at(__float_as_uint(sq.contrib[i].w));
        nd = S.nodes;
        tr = S.leaf;
    };
    auto done = [&](uint32_t i, const TravState& st, uint32_t, uint32_t) { shadow_done(sq, i, st.res, rad); };
    refill_trace<true, false, STACK, kRefillLanes>(n, wave, nwaves, &stk[threadIdx.x], TB, load, done);
}

// Radiance of the group's paths -> each view's output (and running mean).
__global__ __launch_bounds__(kBBlock) void k_bresolve(const JobD* __restrict__ jobs,
                                                      const uint32_t* __restrict__ path_job,
                                                      const float4* __restrict__ rad, float4* __restrict__ accum,
                                                      uint32_t total, const FrameD* __restrict__ fr) {
    const uint32_t accumulate = fr->accumulate, reset = fr->reset;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const JobD& J = jobs[path_job[p]];
        const uint32_t px = p - J.path_off;
        float4 r = rad[p];
        float4 o;
        if (accumulate) {
            float4 a = reset ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : accum[J.accum_off + px];
            a = make_float4(a.x + r.x, a.y + r.y, a.z + r.z, a.w + 1.0f);
            accum[J.accum_off + px] = a;
            float inv = 1.0f / a.w;
            o = make_float4(a.x * inv, a.y * inv, a.z * inv, 1.0f);
        } else {
            o = make_float4(r.x, r.y, r.z, 1.0f);
        }
        if (J.fmt == SNKPT_FORMAT_RGBA16F) {
            __half* h = (__half*)J.out + 4 * (size_t)px;
            h[0] = __float2half(o.x);
            h[1] = __float2half(o.y);
            h[2] = __float2half(o.z);
            h[3] = __float2half(o.w);
        } else {
            ((float4*)J.out)[px] = o;
        }
    }
}

}  // namespace
}  // namespace snkpt

using snkpt::DeviceBuffer;

struct snkpt_batch {
    int device = 0;
    int cus = 256;
    std::string last_error;
    struct View {
        uint32_t scene, w, h, seed, fmt;
        snkpt_camera cam;
        void* out;
    };
    struct Group {
        uint32_t j0, j1, total;
    };
    std::vector<snkpt_ctx*> scenes;
    std::vector<uint64_t> scene_gen;  // each scene context's generation when the tables were built
    std::vector<View> views;
    std::vector<Group> groups;
    snkpt_batch_tuning tuning{128, 16, 8, 0, 0.0f, 0, {0, 0}};
    bool tables_dirty = true;
    uint32_t stack = 16;  // traversal stack entries (deepest scene)
    uint32_t cap = 0;     // queue capacity allocated (paths)
    uint64_t accum_total = 0;
    uint64_t camera_rays = 0;
    DeviceBuffer d_scenes, d_jobs, d_frame;
    DeviceBuffer q_org[2], q_dir[2], q_path[2];
    DeviceBuffer h_t, h_inst, h_prim, h_u, h_v;
    DeviceBuffer s_org, s_dir, s_contrib;
    DeviceBuffer p_thr, p_rng, p_bounce, path_job, rad, accum, counts, stats;
    // cache form: queries, their per-scene sorting (64-row segments), encodings, outputs;
    // hist = scene counts | segment offsets | scatter cursors; the scenes' networks;
    // the trained scenes' seed and stride, and their training streams of the frame
    DeviceBuffer q_samples, q_meta, s_samples, s_meta, q_x0, q_y, hist, refs, train_cfg, train_seed;
    uint32_t qcap = 0;
    uint64_t cache_key = 0;  // settings and tables the cache form's buffers were prepared for
    std::vector<uint32_t> train_scenes, train_views, train_paths;
    snkpt::FrameD frame_host{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // cache form: the training paths (replayed on a side stream beside the main
    // graph) and the optimiser steps (after the join) as graphs of their own
    hipGraph_t graph_train = nullptr, graph_steps = nullptr;
    hipGraphExec_t exec_train = nullptr, exec_steps = nullptr;
    uint64_t graph_key = 0;
    hipStream_t graph_stream = nullptr;
    // cache form: side streams on which the scenes' training paths run during the render
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSide] = {};
    ~snkpt_batch() {
        drop_graph();
        for (int i = 0; i < kSide; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
    }
    void drop_graph() {
        for (hipGraphExec_t* e : {&exec, &exec_train, &exec_steps}) {
            if (*e) (void)hipGraphExecDestroy(*e);
            *e = nullptr;
        }
        for (hipGraph_t* g : {&graph, &graph_train, &graph_steps}) {
            if (*g) (void)hipGraphDestroy(*g);
            *g = nullptr;
        }
        graph_key = 0;
    }
};

namespace snkpt {
namespace {

snkpt_status bfail(snkpt_batch* B, snkpt_status code, const std::string& msg) {
    if (B) B->last_error = msg;
    return code;
}

#define BHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return bfail(B, SNKPT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Smallest traversal stack for the deepest scene: 10 and 12 entries let 16
// and 13 blocks of 128 lanes share a CU's 160 KB of LDS, against 10 for 16.
uint32_t stack_for(uint32_t depth) {
    uint32_t need = depth > 0 ? depth - 1 : 0;
    return need <= 10 ? 10u : need <= 12 ? 12u : need <= 16 ? 16u : need <= 32 ? 32u : 64u;
}

// Scene table, view table (group-relative path offsets), groups, buffers.
snkpt_status prepare(snkpt_batch* B, hipStream_t st) {
    if (!B->tables_dirty) return SNKPT_OK;
    if (B->views.empty()) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch: no views");
    std::vector<SceneDev> sd;
    uint32_t depth = 0;
    B->scene_gen.clear();
    for (snkpt_ctx* c : B->scenes) {
        B->scene_gen.push_back(c->generation);
        if (!c->accel.built) return bfail(B, SNKPT_ERR_NOT_BUILT, "batch: a scene's acceleration structure is not built");
        if (c->accel.num_tris == 0) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch: a scene has no triangles");
        sd.push_back(scene_dev(c));
        depth = std::max(depth, c->accel.depth);
    }
    B->stack = stack_for(depth);
    if (B->stack > 64) return bfail(B, SNKPT_ERR_CAPACITY, "batch: BVH8 deeper than the 64-entry traversal stack");
    // groups of consecutive views holding at most `queue_capacity` paths (a larger view forms its own group)
    const uint64_t capq = B->tuning.queue_capacity;
    B->groups.clear();
    std::vector<JobD> jobs(B->views.size());
    uint64_t accum = 0, cam_rays = 0;
    uint32_t j = 0;
    while (j < B->views.size()) {
        snkpt_batch::Group g{j, j, 0};
        uint64_t tot = 0;
        while (j < B->views.size()) {
            const auto& v = B->views[j];
            uint64_t n = (uint64_t)v.w * v.h;
            if (g.j1 > g.j0 && capq && tot + n > capq) break;
            JobD& J = jobs[j];
            J.cam = make_cam(&v.cam);
            J.scene = v.scene;
            J.w = v.w;
            J.h = v.h;
            J.path_off = (uint32_t)tot;
            J.accum_off = (uint32_t)accum;
            J.seed = v.seed;
            J.fmt = v.fmt;
            J.pad = 0;
            J.out = v.out;
            tot += n;
            accum += n;
            ++j;
            g.j1 = j;
        }
        if (tot > (1ull << 30)) return bfail(B, SNKPT_ERR_CAPACITY, "batch: a group above 2^30 paths");
        g.total = (uint32_t)tot;
        cam_rays += tot;
        B->groups.push_back(g);
    }
    if (accum > 0xFFFFFFFFull) return bfail(B, SNKPT_ERR_CAPACITY, "batch: more than 2^32 pixels");
    uint32_t need = 0;
    for (auto& g : B->groups) need = std::max(need, g.total);
    BHIP(B->d_scenes.reserve(sd.size() * sizeof(SceneDev)));
    BHIP(B->d_jobs.reserve(jobs.size() * sizeof(JobD)));
    BHIP(B->d_frame.reserve(sizeof(FrameD)));
    BHIP(hipMemcpyAsync(B->d_scenes.ptr, sd.data(), sd.size() * sizeof(SceneDev), hipMemcpyHostToDevice, st));
    BHIP(hipMemcpyAsync(B->d_jobs.ptr, jobs.data(), jobs.size() * sizeof(JobD), hipMemcpyHostToDevice, st));
    if (need > B->cap) {
        const size_t n = need;
        for (int k = 0; k < 2; ++k) {
            BHIP(B->q_org[k].reserve(n * 16));
            BHIP(B->q_dir[k].reserve(n * 16));
            BHIP(B->q_path[k].reserve(n * 4));
        }
        BHIP(B->h_t.reserve(n * 4));
        BHIP(B->h_inst.reserve(n * 4));
        BHIP(B->h_prim.reserve(n * 4));
        BHIP(B->h_u.reserve(n * 4));
        BHIP(B->h_v.reserve(n * 4));
        BHIP(B->s_org.reserve(n * 16));
        BHIP(B->s_dir.reserve(n * 16));
        BHIP(B->s_contrib.reserve(n * 16));
        BHIP(B->p_thr.reserve(n * 16));
        BHIP(B->p_rng.reserve(n * 8));
        BHIP(B->p_bounce.reserve(n * 4));
        BHIP(B->path_job.reserve(n * 4));
        BHIP(B->rad.reserve(n * 16));
        B->cap = need;
    }
    if (accum > B->accum_total) {
        BHIP(B->accum.reserve((size_t)accum * 16));
        B->accum_total = accum;
    }
    BHIP(B->counts.reserve(64));
    BHIP(B->stats.reserve(64));
    BHIP(hipStreamSynchronize(st));
    B->camera_rays = cam_rays;
    B->tables_dirty = false;
    return SNKPT_OK;
}

template <int STACK, int TB>
void launch_traces(snkpt_batch* B, const SceneMany& ss, int cur, uint32_t grid, hipStream_t st, bool shadow) {
    if (!shadow) {
        HitArrays hit{B->h_t.as<float>(), B->h_inst.as<uint32_t>(), B->h_prim.as<uint32_t>(), B->h_u.as<float>(),
                      B->h_v.as<float>()};
        k_btrace<STACK, TB><<<grid, TB, 0, st>>>(ss, B->q_org[cur].as<float4>(), B->q_dir[cur].as<float4>(),
                                                 B->q_path[cur].as<uint32_t>(), B->counts.as<uint32_t>() + cur, hit);
    } else {
        ShadowQueue sq{B->s_org.as<float4>(), B->s_dir.as<float4>(), B->s_contrib.as<float4>()};
        k_bshadow<STACK, TB><<<grid, TB, 0, st>>>(ss, sq, B->counts.as<uint32_t>(), B->rad.as<float4>());
    }
}

template <int TB>
void launch_traces_tb(snkpt_batch* B, const SceneMany& ss, int cur, uint32_t grid, hipStream_t st, bool shadow) {
    if (B->stack <= 10)
        launch_traces<10, TB>(B, ss, cur, grid, st, shadow);
    else if (B->stack <= 12)
        launch_traces<12, TB>(B, ss, cur, grid, st, shadow);
    else if (B->stack <= 16)
        launch_traces<16, TB>(B, ss, cur, grid, st, shadow);
    else if (B->stack <= 32)
        launch_traces<32, TB>(B, ss, cur, grid, st, shadow);
    else
        launch_traces<64, TB>(B, ss, cur, grid, st, shadow);
}

void launch_btail(snkpt_batch* B, const SceneMany& ss, const PtParams& P, int cur, hipStream_t st) {
    Queue qin{B->q_org[cur].as<float4>(), B->q_dir[cur].as<float4>(), B->q_path[cur].as<uint32_t>()};
    PathState ps{B->p_thr.as<float4>(), B->p_rng.as<uint2>(), B->p_bounce.as<uint32_t>()};
    const uint32_t grid = (uint32_t)B->cus * kBTailBlocksPerCu;
    const uint32_t* count = B->counts.as<uint32_t>() + cur;
    float4* rad = B->rad.as<float4>();
    unsigned long long* stats = B->stats.as<unsigned long long>();
    if (B->stack <= 10)
        k_btail<10><<<grid, kBTailBlock, 0, st>>>(ss, P, qin, count, ps, rad, stats);
    else if (B->stack <= 16)
        k_btail<16><<<grid, kBTailBlock, 0, st>>>(ss, P, qin, count, ps, rad, stats);
    else if (B->stack <= 32)
        k_btail<32><<<grid, kBTailBlock, 0, st>>>(ss, P, qin, count, ps, rad, stats);
    else
        k_btail<64><<<grid, kBTailBlock, 0, st>>>(ss, P, qin, count, ps, rad, stats);
}

// Bounce of the plain batch frame from which the tail runs (0: never).
uint32_t batch_tail_from(const snkpt_batch_tuning& t, const PtParams& P) {
    const uint32_t b = t.tail_bounce ? t.tail_bounce : P.rr_start;
    return b >= 1 && b < P.max_depth ? b : 0u;
}

// Enqueues one frame of every view (no host synchronisation: capturable).
snkpt_status record_frame(snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    const uint32_t depth = s->max_depth ? s->max_depth : 1u;
    PtParams P{depth, s->rr_start};
    SceneMany ss{B->d_scenes.as<SceneDev>(), B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>()};
    PathState ps{B->p_thr.as<float4>(), B->p_rng.as<uint2>(), B->p_bounce.as<uint32_t>()};
    ShadowQueue sq{B->s_org.as<float4>(), B->s_dir.as<float4>(), B->s_contrib.as<float4>()};
    HitArrays hit{B->h_t.as<float>(), B->h_inst.as<uint32_t>(), B->h_prim.as<uint32_t>(), B->h_u.as<float>(),
                  B->h_v.as<float>()};
    uint32_t* counts = B->counts.as<uint32_t>();
    unsigned long long* stats = B->stats.as<unsigned long long>();
    const uint32_t tb = B->tuning.trace_block == 64 ? 64u : 128u;
    const uint32_t tgrid = (uint32_t)B->cus * std::max(1u, B->tuning.trace_blocks_per_cu);
    const uint32_t sgrid = (uint32_t)B->cus * std::max(1u, B->tuning.shade_blocks_per_cu);
    const uint32_t ggrid = (uint32_t)B->cus * 8u;
    const uint32_t tail = batch_tail_from(B->tuning, P);
    k_bzero<<<1, 64, 0, st>>>(stats, 8);
    for (const auto& g : B->groups) {
        k_bgenerate<<<ggrid, kBBlock, 0, st>>>(B->d_jobs.as<JobD>(), g.j0, g.j1, g.total, B->d_frame.as<FrameD>(),
                                               B->q_org[0].as<float4>(), B->q_dir[0].as<float4>(),
                                               B->q_path[0].as<uint32_t>(), ps, B->path_job.as<uint32_t>(),
                                               B->rad.as<float4>(), counts);
        int cur = 0;
        for (uint32_t b = 0; b < depth; ++b) {
            const int nxt = cur ^ 1;
            k_bbounce<<<1, 64, 0, st>>>(counts, (uint32_t)cur, (uint32_t)nxt, stats, b > 0 ? 1u : 0u);
            if (tail && b >= tail) {
                launch_btail(B, ss, P, cur, st);
                break;
            }
            if (tb == 64)
                launch_traces_tb<64>(B, ss, cur, tgrid, st, false);
            else
                launch_traces_tb<128>(B, ss, cur, tgrid, st, false);
            Queue qin{B->q_org[cur].as<float4>(), B->q_dir[cur].as<float4>(), B->q_path[cur].as<uint32_t>()};
            Queue qout{B->q_org[nxt].as<float4>(), B->q_dir[nxt].as<float4>(), B->q_path[nxt].as<uint32_t>()};
            k_bshade<0><<<sgrid, kBBlock, 0, st>>>(ss, P, qin, hit, qout, sq, ps, B->rad.as<float4>(), counts,
                                                   (uint32_t)cur, (uint32_t)nxt, ShadeHooks{});
            if (tb == 64)
                launch_traces_tb<64>(B, ss, cur, tgrid, st, true);
            else
                launch_traces_tb<128>(B, ss, cur, tgrid, st, true);
            cur = nxt;
        }
        k_bbounce_end<<<1, 64, 0, st>>>(counts, stats);
        k_bresolve<<<ggrid, kBBlock, 0, st>>>(B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>(),
                                              B->rad.as<float4>(), B->accum.as<float4>(), g.total,
                                              B->d_frame.as<FrameD>());
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return bfail(B, SNKPT_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(le));
    return SNKPT_OK;
}

__global__ void k_bset(uint32_t* c, uint32_t i, uint32_t v) {
    if (threadIdx.x == 0) c[i] = v;
}
__global__ void k_bstat_add(const uint32_t* c, uint32_t i, unsigned long long* stat) {
    if (threadIdx.x == 0) *stat += c[i];
}

// Everything a cache-form frame with these settings needs before it is
// recorded (eagerly, outside any capture): each scene's cache, the query and
// segment buffers, the scenes' networks as a device table, each trained
// scene's training buffers and seed/stride record, the side streams.  Redone
// only when the settings or the tables change.
uint64_t cache_key(const snkpt_batch* B, const snkpt_settings* s) {
    uint64_t k = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { k = (k ^ v) * 1099511628211ull; };
    mix(s->max_depth);
    mix(s->cache_train_paths);
    mix(s->flags & (SNKPT_FLAG_NO_TRAIN | SNKPT_FLAG_CACHE_INDIRECT | SNKPT_FLAG_CACHE_MIS));
    mix(s->cache_ema > 0.0f ? 1u : 0u);
    mix(B->cap);
    for (snkpt_ctx* c : B->scenes) mix(c->generation);
    return k | 1ull;
}

snkpt_status prepare_cache_frame(snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    const uint32_t nsc = (uint32_t)B->scenes.size();
    if (nsc > kMaxHistScenes) return bfail(B, SNKPT_ERR_CAPACITY, "batch cache form: more than 1024 scenes");
    const uint64_t key = cache_key(B, s);
    if (key == B->cache_key) return SNKPT_OK;
    B->drop_graph();
    for (snkpt_ctx* c : B->scenes) {
        snkpt_status e = cache_ensure(c, s, st);
        if (e != SNKPT_OK) return bfail(B, e, std::string("batch cache form: ") + c->last_error);
    }
    const uint32_t cap = B->cap;
    const size_t rows = (size_t)cap + 64u * nsc;  // sorted queries with each scene's segment padded to 64 rows
    if (cap > B->qcap) {
        BHIP(B->q_samples.reserve((size_t)cap * 64));
        BHIP(B->q_meta.reserve((size_t)cap * 16));
        B->qcap = cap;
    }
    BHIP(B->s_samples.reserve(rows * 64));
    BHIP(B->s_meta.reserve(rows * 16));
    BHIP(B->q_x0.reserve(rows * 128));
    BHIP(B->q_y.reserve(rows * 16));
    BHIP(B->hist.reserve(((size_t)3 * nsc + 1) * 4));
    std::vector<CacheRefD> refs(nsc);
    const bool avg = s->cache_ema > 0.0f;
    for (uint32_t k = 0; k < nsc; ++k) {
        snkpt_status e = cache_ref(B->scenes[k], avg, &refs[k]);
        if (e != SNKPT_OK) return bfail(B, e, "batch cache form: " + B->scenes[k]->last_error);
    }
    BHIP(B->refs.reserve((size_t)nsc * sizeof(CacheRefD)));
    BHIP(hipMemcpyAsync(B->refs.ptr, refs.data(), (size_t)nsc * sizeof(CacheRefD), hipMemcpyHostToDevice, st));
    // training: full paths through each scene's first view, budget split by the scenes' pixels
    B->train_scenes.clear();
    B->train_views.clear();
    B->train_paths.clear();
    std::vector<uint2> cfg;
    if (!(s->flags & SNKPT_FLAG_NO_TRAIN)) {
        if (!B->fork_ev) {
            for (int i = 0; i < snkpt_batch::kSide; ++i) {
                BHIP(hipStreamCreateWithFlags(&B->side[i], hipStreamNonBlocking));
                BHIP(hipEventCreateWithFlags(&B->join_ev[i], hipEventDisableTiming));
            }
            BHIP(hipEventCreateWithFlags(&B->fork_ev, hipEventDisableTiming));
        }
        std::vector<uint64_t> px(nsc, 0);
        std::vector<int64_t> first(nsc, -1);
        uint64_t total = 0;
        for (size_t v = 0; v < B->views.size(); ++v) {
            const auto& V = B->views[v];
            px[V.scene] += (uint64_t)V.w * V.h;
            total += (uint64_t)V.w * V.h;
            if (first[V.scene] < 0) first[V.scene] = (int64_t)v;
        }
        const uint64_t budget = s->cache_train_paths ? s->cache_train_paths : 65536u;
        for (uint32_t k = 0; k < nsc; ++k) {
            if (first[k] < 0) continue;
            const auto& V = B->views[(size_t)first[k]];
            const uint64_t npx = (uint64_t)V.w * V.h;
            uint64_t nt = std::max<uint64_t>(1024u, budget * px[k] / std::max<uint64_t>(total, 1));
            nt = std::min<uint64_t>(nt, npx);
            snkpt_status e = cache_prepare_train(B->scenes[k], s, V.w, V.h, (uint32_t)nt, st);
            if (e != SNKPT_OK) return bfail(B, e, "batch cache training: " + B->scenes[k]->last_error);
            B->train_scenes.push_back(k);
            B->train_views.push_back((uint32_t)first[k]);
            B->train_paths.push_back((uint32_t)nt);
            cfg.push_back(make_uint2(V.seed, std::max<uint32_t>((uint32_t)(npx / nt), 1u)));
        }
    }
    if (!cfg.empty()) {
        BHIP(B->train_cfg.reserve(cfg.size() * sizeof(uint2)));
        BHIP(B->train_seed.reserve(cfg.size() * sizeof(TrainSeedD)));
        BHIP(hipMemcpyAsync(B->train_cfg.ptr, cfg.data(), cfg.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    }
    BHIP(hipStreamSynchronize(st));
    B->cache_key = key;
    return SNKPT_OK;
}

bool cache_trains(const snkpt_batch* B, const snkpt_settings* s) {
    return !(s->flags & SNKPT_FLAG_NO_TRAIN) && !B->train_scenes.empty();
}

// Cache form, training half: each trained scene's training-path streams for
// this frame, from the frame index in device memory ...
void record_train_seeds(snkpt_batch* B, hipStream_t st) {
    const uint32_t nt = (uint32_t)B->train_scenes.size();
    k_btrain_seeds<<<(nt + 63) / 64, 64, 0, st>>>(B->train_cfg.as<uint2>(), nt, B->d_frame.as<FrameD>(),
                                                 B->train_seed.as<TrainSeedD>());
}

// ... then full paths through each trained scene's first view (they need no
// network); scene j's go to streams[j % n], which must follow the seeds.
snkpt_status record_train_paths(snkpt_batch* B, const snkpt_settings* s, hipStream_t* streams, int n) {
    const uint32_t nt = (uint32_t)B->train_scenes.size();
    TrainSeedD* tseed = B->train_seed.as<TrainSeedD>();
    for (uint32_t j = 0; j < nt; ++j) {
        snkpt_ctx* c = B->scenes[B->train_scenes[j]];
        const auto& V = B->views[B->train_views[j]];
        const CamDev cd = make_cam(&V.cam);
        snkpt_status e = cache_train_paths_forked(c, cd, V.w, V.h, s, pass_state(V.seed, s->frame, 0),
                                                  B->train_paths[j], streams[j % (uint32_t)n], tseed + j);
        if (e != SNKPT_OK) return bfail(B, e, std::string("batch cache training: ") + c->last_error);
    }
    return SNKPT_OK;
}

// Cache form, optimiser steps of every trained scene (after its training paths).
snkpt_status record_cache_steps(snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    std::vector<snkpt_ctx*> trained;
    for (uint32_t k : B->train_scenes) trained.push_back(B->scenes[k]);
    snkpt_status e = cache_train_steps(trained.data(), (uint32_t)trained.size(), s,
                                       B->stats.as<unsigned long long>(), st);
    if (e != SNKPT_OK) return bfail(B, e, "batch cache training step failed");
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return bfail(B, SNKPT_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(le));
    return SNKPT_OK;
}

// Cache form, render half: the batched paths end in a cache query at vertex Q
// (or take NEE there and query the indirect light); the queries are grouped by
// scene on the device (64-row segments) and each scene's own network answers
// its segment.
snkpt_status record_cache_render(snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    const uint32_t nsc = (uint32_t)B->scenes.size();
    const uint32_t depth = s->max_depth ? s->max_depth : 1u;
    const uint32_t Q = s->cache_query_depth ? s->cache_query_depth : 2u;
    const uint32_t rows = B->cap + 64u * nsc;
    // with depth >= Q the vertex before the query keeps its continuation; with depth < Q this is plain PT.
    // The MIS probe of an indirect query vertex is one more bounce.
    const bool mis = (s->flags & SNKPT_FLAG_CACHE_INDIRECT) && (s->flags & SNKPT_FLAG_CACHE_MIS);
    PtParams P{(mis && depth >= Q) ? Q + 1 : std::min(depth, Q), s->rr_start};
    SceneMany ss{B->d_scenes.as<SceneDev>(), B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>()};
    PathState ps{B->p_thr.as<float4>(), B->p_rng.as<uint2>(), B->p_bounce.as<uint32_t>()};
    ShadowQueue sq{B->s_org.as<float4>(), B->s_dir.as<float4>(), B->s_contrib.as<float4>()};
    HitArrays hit{B->h_t.as<float>(), B->h_inst.as<uint32_t>(), B->h_prim.as<uint32_t>(), B->h_u.as<float>(),
                  B->h_v.as<float>()};
    uint32_t* counts = B->counts.as<uint32_t>();
    unsigned long long* stats = B->stats.as<unsigned long long>();
    ShadeHooks H;
    H.kind = 1;
    H.query_at = Q - 1;
    H.indirect = (s->flags & SNKPT_FLAG_CACHE_INDIRECT) ? 1u : 0u;
    H.mis = mis ? 1u : 0u;
    H.q_samples = B->q_samples.as<float>();
    H.q_meta = B->q_meta.as<float4>();
    H.q_count = counts + 3;
    const uint32_t tb = B->tuning.trace_block == 64 ? 64u : 128u;
    const uint32_t tgrid = (uint32_t)B->cus * std::max(1u, B->tuning.trace_blocks_per_cu);
    const uint32_t sgrid = (uint32_t)B->cus * std::max(1u, B->tuning.shade_blocks_per_cu);
    const uint32_t ggrid = (uint32_t)B->cus * 8u;
    uint32_t* d_hist = B->hist.as<uint32_t>();
    uint32_t* d_off = d_hist + nsc;  // nsc + 1 entries
    uint32_t* d_cursor = d_off + nsc + 1;
    const CacheRefD* refs = B->refs.as<CacheRefD>();
    k_bzero<<<1, 64, 0, st>>>(stats, 8);
    for (const auto& g : B->groups) {
        k_bgenerate<<<ggrid, kBBlock, 0, st>>>(B->d_jobs.as<JobD>(), g.j0, g.j1, g.total, B->d_frame.as<FrameD>(),
                                               B->q_org[0].as<float4>(), B->q_dir[0].as<float4>(),
                                               B->q_path[0].as<uint32_t>(), ps, B->path_job.as<uint32_t>(),
                                               B->rad.as<float4>(), counts);
        k_bset<<<1, 64, 0, st>>>(counts, 3, 0u);
        int cur = 0;
        for (uint32_t b = 0; b < P.max_depth; ++b) {
            const int nxt = cur ^ 1;
            k_bbounce<<<1, 64, 0, st>>>(counts, (uint32_t)cur, (uint32_t)nxt, stats, b > 0 ? 1u : 0u);
            if (tb == 64)
                launch_traces_tb<64>(B, ss, cur, tgrid, st, false);
            else
                launch_traces_tb<128>(B, ss, cur, tgrid, st, false);
            Queue qin{B->q_org[cur].as<float4>(), B->q_dir[cur].as<float4>(), B->q_path[cur].as<uint32_t>()};
            Queue qout{B->q_org[nxt].as<float4>(), B->q_dir[nxt].as<float4>(), B->q_path[nxt].as<uint32_t>()};
            k_bshade<1><<<sgrid, kBBlock, 0, st>>>(ss, P, qin, hit, qout, sq, ps, B->rad.as<float4>(), counts,
                                                   (uint32_t)cur, (uint32_t)nxt, H);
            if (tb == 64)
                launch_traces_tb<64>(B, ss, cur, tgrid, st, true);
            else
                launch_traces_tb<128>(B, ss, cur, tgrid, st, true);
            cur = nxt;
        }
        k_bbounce_end<<<1, 64, 0, st>>>(counts, stats);
        k_bstat_add<<<1, 64, 0, st>>>(counts, 3, stats + 3);
        // group the queries by scene and query each scene's network, all on the device
        k_bq_clear<<<1, 256, 0, st>>>(d_hist, nsc);
        k_bq_hist<<<grid_for(g.total, kBBlock), kBBlock, 0, st>>>(B->q_meta.as<float4>(), counts + 3,
                                                                  B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>(),
                                                                  nsc, d_hist);
        k_bq_layout<<<1, kMaxHistScenes, 0, st>>>(d_hist, nsc, d_off, d_cursor);
        k_bq_pad<<<nsc, 64, 0, st>>>(d_hist, d_off, B->s_samples.as<float>(), B->s_meta.as<float4>());
        k_bq_scatter<<<grid_for(g.total, kBBlock), kBBlock, 0, st>>>(B->q_samples.as<float>(), B->q_meta.as<float4>(),
                                                                     counts + 3, B->d_jobs.as<JobD>(),
                                                                     B->path_job.as<uint32_t>(), nsc, d_cursor,
                                                                     B->s_samples.as<float>(), B->s_meta.as<float4>());
        cache_query_many(refs, d_off, nsc, B->s_samples.as<float>(), rows, B->q_x0.as<_Float16>(),
                         B->q_y.as<float4>(), st);
        k_bq_apply<<<grid_for(rows, 256), 256, 0, st>>>(B->s_meta.as<float4>(), B->q_y.as<float4>(), rows,
                                                        d_off + nsc, B->rad.as<float4>());
        k_bresolve<<<ggrid, kBBlock, 0, st>>>(B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>(),
                                              B->rad.as<float4>(), B->accum.as<float4>(), g.total,
                                              B->d_frame.as<FrameD>());
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return bfail(B, SNKPT_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(le));
    return SNKPT_OK;
}

// One frame of every view in the cache form, eagerly: the training paths go
// round robin onto the side streams while the render half runs on `st`; the
// optimiser steps follow the join.  Launches only (prepare_cache_frame
// allocated everything; every count stays on the device).
snkpt_status record_cache_frame(snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    const bool train = cache_trains(B, s);
    snkpt_status e;
    if (train) {
        // the seeds go on `st` before the fork, so every side stream sees them
        record_train_seeds(B, st);
        BHIP(hipEventRecord(B->fork_ev, st));
        for (int i = 0; i < snkpt_batch::kSide; ++i) BHIP(hipStreamWaitEvent(B->side[i], B->fork_ev, 0));
        if ((e = record_train_paths(B, s, B->side, snkpt_batch::kSide)) != SNKPT_OK) return e;
        for (int i = 0; i < snkpt_batch::kSide; ++i) BHIP(hipEventRecord(B->join_ev[i], B->side[i]));
    }
    if ((e = record_cache_render(B, s, st)) != SNKPT_OK) return e;
    if (train) {
        for (int i = 0; i < snkpt_batch::kSide; ++i) BHIP(hipStreamWaitEvent(st, B->join_ev[i], 0));
        return record_cache_steps(B, s, st);
    }
    return SNKPT_OK;
}

// Records what `rec` enqueues on `st` into a graph and instantiates it.
template <class Rec>
snkpt_status capture(snkpt_batch* B, hipStream_t st, hipGraph_t* graph, hipGraphExec_t* exec, Rec rec) {
    BHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    snkpt_status re = rec();
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(st, &g);
    if (re != SNKPT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return re;
    }
    if (ce != hipSuccess) return bfail(B, SNKPT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    *graph = g;
    BHIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return SNKPT_OK;
}

uint64_t graph_key(const snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    uint64_t k = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { k = (k ^ v) * 1099511628211ull; };
    mix(s->max_depth);
    mix(s->rr_start);
    mix(B->tuning.trace_block);
    mix(B->tuning.trace_blocks_per_cu);
    mix(B->tuning.shade_blocks_per_cu);
    mix(B->tuning.queue_capacity);
    mix(B->tuning.tail_bounce);
    mix((uint64_t)(uintptr_t)st);
    mix(B->views.size());
    mix(B->scenes.size());
    mix(s->mode);
    if (s->mode == SNKPT_MODE_CACHE) {  // values the cache form's launches take as arguments
        mix(s->flags & (SNKPT_FLAG_NO_TRAIN | SNKPT_FLAG_CACHE_INDIRECT | SNKPT_FLAG_CACHE_MIS));
        mix(s->cache_query_depth);
        mix(s->cache_train_paths);
        mix(s->cache_train_steps);
        uint32_t bits;
        std::memcpy(&bits, &s->cache_lr, 4);
        mix(bits);
        std::memcpy(&bits, &s->cache_ema, 4);
        mix(bits);
        mix(B->cache_key);
    }
    return k | 1ull;
}

snkpt_status batch_render(snkpt_batch* B, const snkpt_settings* s, int use_graph, hipStream_t st) {
    if (!s) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: settings is NULL");
    if (s->mode != SNKPT_MODE_PT && s->mode != SNKPT_MODE_CACHE)
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: SNKPT_MODE_PT or SNKPT_MODE_CACHE");
    if (s->spp > 1) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: one sample per pixel per frame");
    if (use_graph && st == nullptr)
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: graph capture needs a created (non-null) stream");
    if (hipSetDevice(B->device) != hipSuccess) return bfail(B, SNKPT_ERR_HIP, "hipSetDevice");
    // a scene rebuilt, re-lit or re-uploaded since the tables were made: its device
    // pointers, emitter counts or BVH depth may have changed
    for (size_t k = 0; k < B->scenes.size(); ++k)
        if (k >= B->scene_gen.size() || B->scenes[k]->generation != B->scene_gen[k]) B->tables_dirty = true;
    const bool rebuilt = B->tables_dirty;
    snkpt_status e = prepare(B, st);
    if (e != SNKPT_OK) return e;
    if (rebuilt) B->drop_graph();
    B->frame_host.frame = s->frame;
    B->frame_host.accumulate = (s->flags & SNKPT_FLAG_ACCUMULATE) ? 1u : 0u;
    B->frame_host.reset = (s->flags & SNKPT_FLAG_RESET) ? 1u : 0u;
    BHIP(hipMemcpyAsync(B->d_frame.ptr, &B->frame_host, sizeof(FrameD), hipMemcpyHostToDevice, st));
    const bool cache = s->mode == SNKPT_MODE_CACHE;
    if (cache && (e = prepare_cache_frame(B, s, st)) != SNKPT_OK) return e;
    if (!use_graph) return cache ? record_cache_frame(B, s, st) : record_frame(B, s, st);
    const bool train = cache && cache_trains(B, s);
    const uint64_t key = graph_key(B, s, st);
    if (!B->exec || key != B->graph_key) {
        B->drop_graph();
        // HIP replays a graph's branches one after another, so the training paths,
        // which overlap the render when run eagerly, get a graph of their own that
        // is replayed on a side stream beside the main one
        if (train) {
            if ((e = capture(B, B->side[0], &B->graph_train, &B->exec_train,
                             [&] {
                                 record_train_seeds(B, B->side[0]);
                                 return record_train_paths(B, s, B->side, 1);
                             })) != SNKPT_OK)
                return e;
            if ((e = capture(B, st, &B->graph_steps, &B->exec_steps,
                             [&] { return record_cache_steps(B, s, st); })) != SNKPT_OK)
                return e;
        }
        if ((e = capture(B, st, &B->graph, &B->exec, [&] {
                 return cache ? record_cache_render(B, s, st) : record_frame(B, s, st);
             })) != SNKPT_OK)
            return e;
        B->graph_key = key;
    }
    if (train) {
        BHIP(hipEventRecord(B->fork_ev, st));
        BHIP(hipStreamWaitEvent(B->side[0], B->fork_ev, 0));
        BHIP(hipGraphLaunch(B->exec_train, B->side[0]));
        BHIP(hipEventRecord(B->join_ev[0], B->side[0]));
    }
    BHIP(hipGraphLaunch(B->exec, st));
    if (train) {
        BHIP(hipStreamWaitEvent(st, B->join_ev[0], 0));
        BHIP(hipGraphLaunch(B->exec_steps, st));
    }
    return SNKPT_OK;
}

double time_frames(snkpt_batch* B, const snkpt_settings* s, hipStream_t st, int frames) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.0;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.0;
    }
    snkpt_settings t = *s;
    t.flags &= ~(uint32_t)(SNKPT_FLAG_ACCUMULATE | SNKPT_FLAG_RESET);
    double ms = -1.0;
    if (batch_render(B, &t, 0, st) == SNKPT_OK && hipEventRecord(a, st) == hipSuccess) {
        bool ok = true;
        for (int f = 0; f < frames && ok; ++f) {
            t.frame = s->frame + 1 + (uint32_t)f;
            ok = batch_render(B, &t, 0, st) == SNKPT_OK;
        }
        float el = 0.0f;
        if (ok && hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
            hipEventElapsedTime(&el, a, b) == hipSuccess)
            ms = el / frames;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

// Coordinate descent over the launch shapes, the queue capacity and the tail bounce, timing
// whole frames of this batch (results do not depend on the choice).
snkpt_status batch_autotune(snkpt_batch* B, const snkpt_settings* s, hipStream_t st, snkpt_batch_tuning* out) {
    if (hipSetDevice(B->device) != hipSuccess) return bfail(B, SNKPT_ERR_HIP, "hipSetDevice");
    snkpt_batch_tuning best = B->tuning;
    auto trial = [&](const snkpt_batch_tuning& t, double& ms) -> snkpt_status {
        B->tuning = t;
        B->tables_dirty = true;  // the capacity may regroup the views
        B->drop_graph();
        ms = time_frames(B, s, st, 3);
        if (ms < 0.0) return bfail(B, SNKPT_ERR_HIP, "batch_autotune: a trial frame failed: " + B->last_error);
        return SNKPT_OK;
    };
    double best_ms = 0.0;
    snkpt_status e = trial(best, best_ms);
    if (e != SNKPT_OK) return e;
    uint64_t total = 0;
    for (const auto& v : B->views) total += (uint64_t)v.w * v.h;
    const uint32_t tbs[] = {64, 128};
    const uint32_t tpc[] = {4, 8, 16, 32};
    const uint32_t spc[] = {2, 4, 8, 16};
    const uint64_t caps[] = {0, total / 2, total / 4};
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t a : tbs)
            for (uint32_t c : tpc) {
                snkpt_batch_tuning t = best;
                t.trace_block = a;
                t.trace_blocks_per_cu = c;
                double ms;
                if ((e = trial(t, ms)) != SNKPT_OK) return e;
                if (ms < best_ms) best_ms = ms, best = t;
            }
        for (uint32_t c : spc) {
            snkpt_batch_tuning t = best;
            t.shade_blocks_per_cu = c;
            double ms;
            if ((e = trial(t, ms)) != SNKPT_OK) return e;
            if (ms < best_ms) best_ms = ms, best = t;
        }
        for (uint64_t c : caps) {
            if (c > 0xFFFFFFFFull) continue;
            snkpt_batch_tuning t = best;
            t.queue_capacity = (uint32_t)c;
            double ms;
            if ((e = trial(t, ms)) != SNKPT_OK) return e;
            if (ms < best_ms) best_ms = ms, best = t;
        }
        // tail bounce: never, or from each bounce after the camera rays' on
        for (uint32_t b = 0; b < std::max(s->max_depth, 1u); ++b) {
            snkpt_batch_tuning t = best;
            t.tail_bounce = b == 0 ? 0xFFFFFFFFu : b;
            double ms;
            if ((e = trial(t, ms)) != SNKPT_OK) return e;
            if (ms < best_ms) best_ms = ms, best = t;
        }
    }
    best.ms_per_frame = (float)best_ms;
    B->tuning = best;
    B->tables_dirty = true;
    B->drop_graph();
    if (out) *out = best;
    return SNKPT_OK;
}

}  // namespace
}  // namespace snkpt

// ------------------------------------------------------------------ C ABI ----
using namespace snkpt;

extern "C" {

snkpt_status snkpt_batch_create(int hip_device, snkpt_batch** out) {
    if (!out) return SNKPT_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || hip_device < 0 || hip_device >= n) return SNKPT_ERR_INVALID_ARG;
    snkpt_batch* B = new (std::nothrow) snkpt_batch();
    if (!B) return SNKPT_ERR_OOM;
    B->device = hip_device;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, hip_device) == hipSuccess && p.multiProcessorCount > 0) B->cus = p.multiProcessorCount;
    *out = B;
    return SNKPT_OK;
}

void snkpt_batch_destroy(snkpt_batch* B) {
    if (!B) return;
    (void)hipSetDevice(B->device);
    delete B;
}

const char* snkpt_batch_last_error(const snkpt_batch* B) { return B ? B->last_error.c_str() : "null batch"; }

snkpt_status snkpt_batch_add_scene(snkpt_batch* B, snkpt_ctx* scene, uint32_t* out_scene) {
    if (!B || !scene) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_add_scene: NULL argument");
    if (scene->device != B->device) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_add_scene: context on another device");
    if (B->scenes.size() >= 65536) return bfail(B, SNKPT_ERR_CAPACITY, "batch_add_scene: too many scenes");
    B->scenes.push_back(scene);
    B->tables_dirty = true;
    if (out_scene) *out_scene = (uint32_t)(B->scenes.size() - 1);
    return SNKPT_OK;
}

snkpt_status snkpt_batch_add_view(snkpt_batch* B, uint32_t scene, const snkpt_camera* cam, uint32_t w, uint32_t h,
                                  uint32_t seed, void* out_rgba_device, uint32_t out_format, uint32_t* out_view) {
    if (!B || !cam || !out_rgba_device) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_add_view: NULL argument");
    if (scene >= B->scenes.size()) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_add_view: unknown scene");
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 26))
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_add_view: size must be 1..2^26 pixels");
    if (out_format > SNKPT_FORMAT_RGBA16F) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_add_view: unknown format");
    if (B->views.size() >= kMaxViews) return bfail(B, SNKPT_ERR_CAPACITY, "batch_add_view: too many views");
    snkpt_batch::View v{scene, w, h, seed, out_format, *cam, out_rgba_device};
    B->views.push_back(v);
    B->tables_dirty = true;
    if (out_view) *out_view = (uint32_t)(B->views.size() - 1);
    return SNKPT_OK;
}

snkpt_status snkpt_batch_render(snkpt_batch* B, const snkpt_settings* s, int use_graph, snkpt_stream stream) {
    if (!B) return SNKPT_ERR_INVALID_ARG;
    return batch_render(B, s, use_graph, (hipStream_t)stream);
}

snkpt_status snkpt_batch_autotune(snkpt_batch* B, const snkpt_settings* s, snkpt_stream stream,
                                  snkpt_batch_tuning* out) {
    if (!B || !s) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_autotune: NULL argument");
    if (s->mode != SNKPT_MODE_PT) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_autotune: only SNKPT_MODE_PT is batched");
    return batch_autotune(B, s, (hipStream_t)stream, out);
}

snkpt_status snkpt_batch_set_tuning(snkpt_batch* B, const snkpt_batch_tuning* t) {
    if (!B || !t) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_set_tuning: NULL argument");
    if (t->trace_block != 64 && t->trace_block != 128)
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_set_tuning: trace_block must be 64 or 128");
    if (t->trace_blocks_per_cu == 0 || t->trace_blocks_per_cu > 64 || t->shade_blocks_per_cu == 0 ||
        t->shade_blocks_per_cu > 64)
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_set_tuning: blocks per CU must be 1..64");
    B->tuning = *t;
    B->tables_dirty = true;
    return SNKPT_OK;
}

snkpt_status snkpt_batch_get_tuning(const snkpt_batch* B, snkpt_batch_tuning* out) {
    if (!B || !out) return SNKPT_ERR_INVALID_ARG;
    *out = B->tuning;
    return SNKPT_OK;
}

snkpt_status snkpt_batch_frame_stats(snkpt_batch* B, snkpt_frame_stats* out, snkpt_stream stream) {
    if (!B || !out) return SNKPT_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    if (!B->stats.ptr) return SNKPT_OK;
    unsigned long long h[8];
    BHIP(hipMemcpyAsync(h, B->stats.ptr, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
    BHIP(hipStreamSynchronize((hipStream_t)stream));
    out->camera_rays = B->camera_rays;
    out->extension_rays = h[1];
    out->shadow_rays = h[2];
    out->cache_queries = h[3];
    out->cache_train_samples = h[4];
    return SNKPT_OK;
}

}  // extern "C"
