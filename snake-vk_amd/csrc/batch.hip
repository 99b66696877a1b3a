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

// rad[path] += throughput * max(cache output, 0) for the n sorted queries.
__global__ void k_bq_apply(const float4* __restrict__ meta, const float4* __restrict__ y, uint32_t n,
                           float4* __restrict__ rad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 m = meta[i], v = y[i];
    const uint32_t path = __float_as_uint(m.w);
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
    snkpt_batch_tuning tuning{128, 16, 8, 0, 0.0f, {0, 0, 0}};
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
    // cache form: queries, their per-scene sorting, outputs
    DeviceBuffer q_samples, q_meta, s_samples, s_meta, q_y, hist;
    uint32_t qcap = 0;
    snkpt::FrameD frame_host{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
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
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
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

// One frame of every view in the cache form: the batched paths end in a cache
// query at vertex Q (or take NEE there and query the indirect light); the
// queries are grouped by scene and each scene's own network answers them; then
// every scene takes one training step from full paths through its first view
// (cache_train_paths split over the scenes in proportion to their pixels).
// Eager only: the grouping and the training read counts back to the host.
snkpt_status record_cache_frame(snkpt_batch* B, const snkpt_settings* s, hipStream_t st) {
    const uint32_t nsc = (uint32_t)B->scenes.size();
    if (nsc > kMaxHistScenes) return bfail(B, SNKPT_ERR_CAPACITY, "batch cache form: more than 1024 scenes");
    for (snkpt_ctx* c : B->scenes) {
        snkpt_status e = cache_ensure(c, s, st);
        if (e != SNKPT_OK) return bfail(B, e, std::string("batch cache form: ") + c->last_error);
    }
    const uint32_t depth = s->max_depth ? s->max_depth : 1u;
    const uint32_t Q = s->cache_query_depth ? s->cache_query_depth : 2u;
    const uint32_t cap = B->cap;
    if (cap > B->qcap) {
        BHIP(B->q_samples.reserve((size_t)cap * 64));
        BHIP(B->q_meta.reserve((size_t)cap * 16));
        BHIP(B->s_samples.reserve((size_t)cap * 64));
        BHIP(B->s_meta.reserve((size_t)cap * 16));
        BHIP(B->q_y.reserve((size_t)cap * 16));
        B->qcap = cap;
    }
    BHIP(B->hist.reserve((size_t)nsc * 8));
    const bool avg = s->cache_ema > 0.0f;
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
    k_bzero<<<1, 64, 0, st>>>(stats, 8);
    // training: full paths through each scene's first view, traced on side
    // streams (round robin) while the render below runs; they need no network
    const bool train = !(s->flags & SNKPT_FLAG_NO_TRAIN);
    std::vector<snkpt_ctx*> trained;
    if (train) {
        if (!B->fork_ev) {
            for (int i = 0; i < snkpt_batch::kSide; ++i) {
                BHIP(hipStreamCreateWithFlags(&B->side[i], hipStreamNonBlocking));
                BHIP(hipEventCreateWithFlags(&B->join_ev[i], hipEventDisableTiming));
            }
            BHIP(hipEventCreateWithFlags(&B->fork_ev, hipEventDisableTiming));
        }
        BHIP(hipEventRecord(B->fork_ev, st));
        for (int i = 0; i < snkpt_batch::kSide; ++i) BHIP(hipStreamWaitEvent(B->side[i], B->fork_ev, 0));
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
            uint64_t nt = std::max<uint64_t>(1024u, budget * px[k] / std::max<uint64_t>(total, 1));
            nt = std::min<uint64_t>(nt, (uint64_t)V.w * V.h);
            const CamDev cd = make_cam(&V.cam);
            snkpt_status e = cache_train_paths_forked(B->scenes[k], cd, V.w, V.h, s, pass_state(V.seed, s->frame, 0),
                                                      (uint32_t)nt, B->side[trained.size() % snkpt_batch::kSide]);
            if (e != SNKPT_OK) return bfail(B, e, std::string("batch cache training: ") + B->scenes[k]->last_error);
            trained.push_back(B->scenes[k]);
        }
        for (int i = 0; i < snkpt_batch::kSide; ++i) BHIP(hipEventRecord(B->join_ev[i], B->side[i]));
    }
    std::vector<uint32_t> hist(nsc), off(nsc);
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
        // group the queries by scene and query each scene's network
        uint32_t nq = 0;
        BHIP(hipMemcpyAsync(&nq, counts + 3, 4, hipMemcpyDeviceToHost, st));
        BHIP(hipStreamSynchronize(st));
        if (nq > 0) {
            uint32_t* d_hist = B->hist.as<uint32_t>();
            uint32_t* d_cursor = d_hist + nsc;
            std::fill(hist.begin(), hist.end(), 0u);
            BHIP(hipMemcpyAsync(d_hist, hist.data(), (size_t)nsc * 4, hipMemcpyHostToDevice, st));
            k_bq_hist<<<grid_for(nq, kBBlock), kBBlock, 0, st>>>(B->q_meta.as<float4>(), counts + 3,
                                                                 B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>(),
                                                                 nsc, d_hist);
            BHIP(hipMemcpyAsync(hist.data(), d_hist, (size_t)nsc * 4, hipMemcpyDeviceToHost, st));
            BHIP(hipStreamSynchronize(st));
            uint32_t run = 0;
            for (uint32_t k = 0; k < nsc; ++k) {
                off[k] = run;
                run += hist[k];
            }
            BHIP(hipMemcpyAsync(d_cursor, off.data(), (size_t)nsc * 4, hipMemcpyHostToDevice, st));
            k_bq_scatter<<<grid_for(nq, kBBlock), kBBlock, 0, st>>>(B->q_samples.as<float>(), B->q_meta.as<float4>(),
                                                                    counts + 3, B->d_jobs.as<JobD>(),
                                                                    B->path_job.as<uint32_t>(), nsc, d_cursor,
                                                                    B->s_samples.as<float>(), B->s_meta.as<float4>());
            for (uint32_t k = 0; k < nsc; ++k) {
                if (!hist[k]) continue;
                snkpt_status e = cache_query(B->scenes[k], B->s_samples.as<float>() + (size_t)off[k] * 16, hist[k],
                                             B->q_y.as<float4>() + off[k], avg, st);
                if (e != SNKPT_OK) return bfail(B, e, std::string("batch cache query: ") + B->scenes[k]->last_error);
            }
            k_bq_apply<<<grid_for(nq, 256), 256, 0, st>>>(B->s_meta.as<float4>(), B->q_y.as<float4>(), nq,
                                                          B->rad.as<float4>());
        }
        k_bresolve<<<ggrid, kBBlock, 0, st>>>(B->d_jobs.as<JobD>(), B->path_job.as<uint32_t>(),
                                              B->rad.as<float4>(), B->accum.as<float4>(), g.total,
                                              B->d_frame.as<FrameD>());
    }
    if (train) {
        // join the training paths, then one step per scene in scene order
        for (int i = 0; i < snkpt_batch::kSide; ++i) BHIP(hipStreamWaitEvent(st, B->join_ev[i], 0));
        snkpt_status e = cache_train_steps(trained.data(), (uint32_t)trained.size(), s, stats, st);
        if (e != SNKPT_OK) return bfail(B, e, "batch cache training step failed");
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return bfail(B, SNKPT_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(le));
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
    mix((uint64_t)(uintptr_t)st);
    mix(B->views.size());
    mix(B->scenes.size());
    return k | 1ull;
}

snkpt_status batch_render(snkpt_batch* B, const snkpt_settings* s, int use_graph, hipStream_t st) {
    if (!s) return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: settings is NULL");
    if (s->mode != SNKPT_MODE_PT && s->mode != SNKPT_MODE_CACHE)
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: SNKPT_MODE_PT or SNKPT_MODE_CACHE");
    if (s->mode == SNKPT_MODE_CACHE && use_graph)
        return bfail(B, SNKPT_ERR_INVALID_ARG, "batch_render: the cache form runs eagerly (use_graph = 0)");
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
    if (s->mode == SNKPT_MODE_CACHE) return record_cache_frame(B, s, st);
    if (!use_graph) return record_frame(B, s, st);
    const uint64_t key = graph_key(B, s, st);
    if (!B->exec || key != B->graph_key) {
        B->drop_graph();
        BHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        snkpt_status re = record_frame(B, s, st);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(st, &g);
        if (re != SNKPT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return re;
        }
        if (ce != hipSuccess) return bfail(B, SNKPT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        B->graph = g;
        BHIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->graph_key = key;
    }
    BHIP(hipGraphLaunch(B->exec, st));
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

// Coordinate descent over the launch shapes and the queue capacity, timing
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
