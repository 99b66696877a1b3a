// Internal host-side declarations of the snkpt library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "snkpt.h"

namespace snkpt {

// Owning device allocation; grows only.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    // Ensures at least `n` bytes; contents are not preserved.
    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        release();
        size_t alloc = n < 256 ? 256 : n;
        hipError_t e = hipMalloc(&ptr, alloc);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = alloc;
        return hipSuccess;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Acceleration structure built by build_accel().
// Launch shape of the persistent traversal kernel (refill_trace): refill once
// `refill` lanes of a wave are idle, `steps` traversal steps between two
// refill checks (3 or 4), `blocks_per_cu` resident 128-lane blocks per CU.
// refill == 0: not set (the defaults apply).
struct TraceShape {
    uint32_t refill = 0, steps = 4, blocks_per_cu = 16;
    float ms = 0.0f;  // the autotuner's measurement
};

struct Accel {
    bool built = false;
    TraceShape shape[2];      // [closest hit, any hit]
    bool shape_fixed = false; // set by the caller: the autotuner leaves it alone
    uint32_t num_tris = 0;
    uint32_t num_nodes = 0;
    uint32_t depth = 0;
    uint32_t binary_depth = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double build_ms = 0;
    DeviceBuffer world_tris;  // 48 B records in input (global triangle id) order
    DeviceBuffer leaf_tris;   // 48 B records in BVH8 leaf order
    DeviceBuffer nodes;       // 80 B BVH8 nodes, BFS order, root = 0
    DeviceBuffer keys, ids;   // sorted Morton keys / triangle ids
    DeviceBuffer bin_child;   // internal binary nodes: left, right (tagged)
    DeviceBuffer bin_aabb;    // internal then leaf AABBs, 6 floats each
    DeviceBuffer bin_leafify; // u32 per internal node
    DeviceBuffer bin_dpc, bin_dpk;  // collapse costs (8 floats) and split choices (8 x 3 bits) per internal node
    DeviceBuffer inst_tri_offset;  // per instance: first global triangle id (n_inst + 1 entries)
    // build scratch, kept between builds so a per-frame rebuild allocates nothing
    DeviceBuffer keys2, ids2, hist, hist_scan, sums, bounds, first, count, pint, pleaf, flags, up_live, up_cnt, cost,
        height, slot_child, nint, ntri, sint, stri, work, totals;
};

// Emitter table: one entry per emissive triangle, sampled proportionally to power.
struct Emitters {
    uint32_t count = 0;
    double total_power = 0;
    // one 64 B record per emitter (4 float4, world space): v0 | alias probability
    // bits, v1 | alias index bits, v2 | selection probability, emission rgb | area.
    // A ReSTIR candidate or NEE sample reads the record its uniform draw lands on
    // and, when the alias redirects, the alias's record: one or two 64 B reads.
    DeviceBuffer rec;
    DeviceBuffer gid_to_emitter;  // per global triangle: emitter index or ~0u
    DeviceBuffer analytic;        // directional / point / spot lights, 4 float4 each
    uint32_t n_analytic = 0;
    float q_tri = 1.0f;           // NEE: probability of sampling the triangles rather than an analytic light
    // the light block as the reference's shaders read it (SNKPT_MODE_REFERENCE_COMPAT)
    DeviceBuffer ref_pl;          // every point light of the block, 3 float4: colour | position | (c, l, e)
    uint32_t n_ref_pl = 0;
    DeviceBuffer em_list;         // the emissive instance list
    uint32_t n_em_list = 0;
    float ref_dir[8] = {0};       // directional light colour, direction (as stored)
};

struct RenderState;  // defined in render.hip
struct CacheState;   // defined in cache.hip
struct TaaState;     // defined in taa.hip

}  // namespace snkpt

struct snkpt_ctx {
    int device = 0;
    std::string last_error;
    hipDeviceProp_t props{};

    // host copies of the scene input
    std::vector<float> pos, nrm, uv, tan;
    std::vector<uint32_t> idx;
    std::vector<snkpt_instance> inst;
    std::vector<float> mat4s;
    std::vector<snkpt_material> mats;
    std::vector<uint8_t> light_block;
    std::vector<uint32_t> emissive_inst;
    bool emissive_auto = true;

    // device copies
    snkpt::DeviceBuffer d_pos, d_nrm, d_uv, d_idx, d_inst, d_mat4, d_mats;
    snkpt::DeviceBuffer d_inst_xf;  // per instance: 3x4 object->world + 3x3 normal matrix (float x 24)
    std::vector<float> prev_mat4s;  // previous frame's transforms (contact velocities); empty = current
    snkpt::DeviceBuffer d_prev_xf;  // per instance: previous 3x4 object->world (float x 12)
    bool dirty = true;
    // host arrays changed since their last upload (build_accel uploads only these)
    bool mesh_dirty = true, inst_dirty = true, mats_dirty = true;
    // largest index per (index offset, count) range, for validating instances; cleared with the meshes
    std::unordered_map<uint64_t, uint32_t> range_max;

    snkpt::Accel accel;
    snkpt::Emitters emit;
    snkpt::RenderState* render = nullptr;
    snkpt::CacheState* cache = nullptr;
    snkpt::TaaState* taa = nullptr;
    snkpt_frame_stats stats{};
    // State generation: advanced by every call that changes scene data or may move
    // a device buffer that launched work reads (scene input, build, render-buffer
    // growth, cache reset).  Batches compare it to rebuild their tables; callers
    // that capture frames re-capture when it changes.
    uint64_t generation = 1;
};

namespace snkpt {

snkpt_status fail(snkpt_ctx* ctx, snkpt_status code, const std::string& msg);
snkpt_status hip_fail(snkpt_ctx* ctx, hipError_t e, const char* what);

snkpt_status build_accel(snkpt_ctx* ctx, hipStream_t stream);
// Exclusive scan of n u32 (n <= 2^24); `sums` needs ceil(n / 4096) entries; the
// grand total goes to *total (device memory) when total != nullptr.
void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, uint32_t* total,
                    hipStream_t st);
snkpt_status accel_export(snkpt_ctx* ctx, uint32_t which, void* dst, size_t bytes, size_t* out_bytes);
snkpt_status build_emitters(snkpt_ctx* ctx, hipStream_t stream);

// trace.hip
snkpt_status launch_trace(snkpt_ctx* ctx, const float4* org_tmin, const float4* dir_tmax, uint32_t n,
                          const snkpt_hits& hits, bool any_hit, uint32_t* node_visits, uint32_t* tri_visits,
                          const uint32_t* count_dev, hipStream_t stream);

// render.hip
// gb: primary vertex from a rasterized G-buffer instead of camera rays (NULL: camera rays)
snkpt_status render(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* s, void* out, uint32_t w,
                    uint32_t h, uint32_t fmt, hipStream_t stream, const snkpt_gbuffer* gb = nullptr);
snkpt_status oracle_accumulate(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* s, double* accum,
                               uint32_t w, uint32_t h, uint32_t spp, hipStream_t stream);
void render_state_free(RenderState* r);
void cache_state_free(CacheState* c);
void taa_state_free(TaaState* t);

// contacts.hip
snkpt_status trace_contacts(snkpt_ctx* ctx, const float4* org_tmin, const float4* dir_tmax, uint32_t n,
                            const snkpt_contacts& out, float timestep, hipStream_t st);

// taa.hip
snkpt_status taa_resolve(snkpt_ctx* ctx, const void* current, void* out, const uint16_t* velocity, uint32_t w,
                         uint32_t h, uint32_t fmt, uint32_t reset, hipStream_t st);

// trace.hip: launch shape of the traversal kernel for the built structure
snkpt_status autotune_trace(snkpt_ctx* ctx, hipStream_t st);

// cache.hip
snkpt_status cache_reset(snkpt_ctx* ctx, uint32_t seed, const float* lo, const float* hi, hipStream_t st);
snkpt_status cache_infer(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out, hipStream_t st);
// queries whose workspace is a static device array (no allocation); at most kNoAllocQueries
constexpr uint32_t kNoAllocQueries = 65536;
snkpt_status cache_scratch_init(int device);
snkpt_status cache_infer_noalloc(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out, hipStream_t st);
snkpt_status cache_train(snkpt_ctx* ctx, const float* samples, uint32_t n, float lr, int apply, float* loss_out,
                         hipStream_t st);
// one optimiser step on min(n, *n_dev) samples (n alone when n_dev is null)
snkpt_status cache_train_run(snkpt_ctx* ctx, const float* samples, uint32_t n, const uint32_t* n_dev, float lr,
                             int apply, float ema, float* loss_out, hipStream_t st);
snkpt_status cache_params(snkpt_ctx* ctx, uint32_t which, void* dst, size_t bytes, size_t* out_bytes, hipStream_t st);
snkpt_status cache_set_params(snkpt_ctx* ctx, uint32_t which, const void* src, size_t bytes, hipStream_t st);

// scenegen.cpp
snkpt_status scene_generate(uint32_t config, uint32_t variant, snkpt_scene** out);
void scene_free(snkpt_scene* s);
snkpt_status scene_camera(uint32_t config, uint32_t variant, uint32_t frame, uint32_t w, uint32_t h,
                          snkpt_camera* out);
uint32_t c5_view_side(uint32_t view);

}  // namespace snkpt
