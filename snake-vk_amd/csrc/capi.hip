// C ABI entry points (snkpt.h): context lifetime, scene input, dispatch.
//
// Mirrors the buffer-upload side of the reference pass: MeshBufferManager
// (MeshBufferManager.cpp:22-88), RaytracingInstanceBufferSystem
// (RaytracingBufferSystem.cpp:69-130), GlobalMaterialBufferManager
// (ResourceBufferManagers.cpp:64-101) and LightBufferSystem
// (LightBufferSystem.cpp:65-108).  Where the reference asserts and breaks into
// the debugger (util.h:18-30) this library returns a status and keeps the
// message for snkpt_last_error.
#include <cmath>
#include <cstring>
#include <new>

#include "snkpt_host.h"

namespace snkpt {

snkpt_status fail(snkpt_ctx* ctx, snkpt_status code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

snkpt_status hip_fail(snkpt_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(ctx, SNKPT_ERR_OOM, std::string(what) + ": " + hipGetErrorString(e));
    return fail(ctx, SNKPT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static snkpt_status upload(snkpt_ctx* ctx, DeviceBuffer& buf, const void* src, size_t bytes) {
    hipError_t e = buf.reserve(bytes ? bytes : 4);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    if (bytes) {
        e = hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy");
    }
    return SNKPT_OK;
}

// Per instance: object->world rows (3 x float4) and the normal matrix
// transpose(inverse(mat3(M))) rows (3 x float4), as RayClosestHit.rchit:110-111.
static void instance_xf(const float* m, float out[24]) {
    // m column-major: element (row r, col c) = m[4c + r]
    for (int r = 0; r < 3; ++r) {
        out[4 * r + 0] = m[r];
        out[4 * r + 1] = m[4 + r];
        out[4 * r + 2] = m[8 + r];
        out[4 * r + 3] = m[12 + r];
    }
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = m[4 * c + r];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv_det = det != 0.0 ? 1.0 / det : 0.0;
    // cofactor matrix / det = transpose(inverse(a))
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
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out[12 + 4 * r + c] = (float)(cof[r][c] * inv_det);
        out[12 + 4 * r + 3] = 0.0f;
    }
}

}  // namespace snkpt

using namespace snkpt;

extern "C" {

snkpt_status snkpt_create(int hip_device, snkpt_ctx** out) {
    if (!out) return SNKPT_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || hip_device < 0 || hip_device >= n) return SNKPT_ERR_INVALID_ARG;
    snkpt_ctx* ctx = new (std::nothrow) snkpt_ctx();
    if (!ctx) return SNKPT_ERR_OOM;
    ctx->device = hip_device;
    if (hipSetDevice(hip_device) != hipSuccess || hipGetDeviceProperties(&ctx->props, hip_device) != hipSuccess ||
        cache_scratch_init(hip_device) != SNKPT_OK) {
        delete ctx;
        return SNKPT_ERR_HIP;
    }
    *out = ctx;
    return SNKPT_OK;
}

void snkpt_destroy(snkpt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    render_state_free(ctx->render);
    cache_state_free(ctx->cache);
    taa_state_free(ctx->taa);
    delete ctx;
}

const char* snkpt_last_error(const snkpt_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

snkpt_status snkpt_upload_mesh(snkpt_ctx* ctx, const float* pos, const float* nrm, const float* uv, const float* tan,
                               uint32_t n_verts, const uint32_t* idx, uint32_t n_idx, uint32_t* out_vertex_offset,
                               uint32_t* out_index_offset) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if ((n_verts && !pos) || (n_idx && !idx) || n_idx % 3 != 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_upload_mesh: null arrays or index count not a multiple of 3");
    for (uint32_t i = 0; i < n_idx; ++i)
        if (idx[i] >= n_verts) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_upload_mesh: index out of range");
    uint64_t vo = ctx->pos.size() / 3, io = ctx->idx.size();
    if (vo + n_verts >= (1ull << 32) || io + n_idx >= (1ull << 32))
        return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_upload_mesh: global buffers exceed 2^32 entries");
    ctx->pos.insert(ctx->pos.end(), pos, pos + 3 * (size_t)n_verts);
    if (nrm)
        ctx->nrm.insert(ctx->nrm.end(), nrm, nrm + 3 * (size_t)n_verts);
    else
        ctx->nrm.resize(ctx->nrm.size() + 3 * (size_t)n_verts, 0.0f);
    if (uv)
        ctx->uv.insert(ctx->uv.end(), uv, uv + 2 * (size_t)n_verts);
    else
        ctx->uv.resize(ctx->uv.size() + 2 * (size_t)n_verts, 0.0f);
    if (tan)
        ctx->tan.insert(ctx->tan.end(), tan, tan + 3 * (size_t)n_verts);
    else
        ctx->tan.resize(ctx->tan.size() + 3 * (size_t)n_verts, 0.0f);
    ctx->idx.insert(ctx->idx.end(), idx, idx + n_idx);
    if (out_vertex_offset) *out_vertex_offset = (uint32_t)vo;
    if (out_index_offset) *out_index_offset = (uint32_t)io;
    ctx->dirty = true;
    ctx->mesh_dirty = true;
    ctx->range_max.clear();
    ctx->accel.built = false;
    return SNKPT_OK;
}

snkpt_status snkpt_set_instances(snkpt_ctx* ctx, const snkpt_instance* inst, uint32_t n, const float* mat4s,
                                 uint32_t n_mat4) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if ((n && !inst) || (n_mat4 && !mat4s)) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_instances: null arrays");
    if (n > SNKPT_MAX_INSTANCES) return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_set_instances: more than 4096 instances");
    if (n_mat4 > SNKPT_MAX_TRANSFORMS)
        return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_set_instances: more than 4096 transforms");
    ctx->inst.assign(inst, inst + n);
    ctx->mat4s.assign(mat4s, mat4s + 16 * (size_t)n_mat4);
    ctx->dirty = true;
    ctx->inst_dirty = true;
    ctx->accel.built = false;
    return SNKPT_OK;
}

snkpt_status snkpt_set_materials(snkpt_ctx* ctx, const snkpt_material* m, uint32_t n) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if (n && !m) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_materials: null array");
    if (n > SNKPT_MAX_MATERIALS) return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_set_materials: more than 4096 materials");
    for (uint32_t i = 0; i < n; ++i)
        if (m[i].albedo_tex_idx != SNKPT_INVALID_INDEX)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_materials: textures are not supported on this path");
    ctx->mats.assign(m, m + n);
    ctx->dirty = true;
    ctx->mats_dirty = true;
    ctx->accel.built = false;
    return SNKPT_OK;
}

snkpt_status snkpt_set_lights(snkpt_ctx* ctx, const void* light_block, size_t bytes,
                              const uint32_t* emissive_instances, uint32_t n_emissive) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if (light_block) {
        if (bytes < 112) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_lights: light block shorter than its header");
        uint32_t np, ns;
        std::memcpy(&np, (const uint8_t*)light_block + 96, 4);
        std::memcpy(&ns, (const uint8_t*)light_block + 100, 4);
        if (np > SNKPT_MAX_POINTLIGHTS || ns > SNKPT_MAX_SPOTLIGHTS)
            return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_set_lights: light counts above 1024");
        if (bytes < 112 + (size_t)np * 48)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_lights: block shorter than its point lights");
        if (ns && bytes < 112 + 48 * (size_t)SNKPT_MAX_POINTLIGHTS + (size_t)ns * 64)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_lights: block shorter than its spot lights");
        ctx->light_block.assign((const uint8_t*)light_block, (const uint8_t*)light_block + bytes);
    } else {
        ctx->light_block.clear();
    }
    if (emissive_instances) {
        ctx->emissive_inst.assign(emissive_instances, emissive_instances + n_emissive);
        ctx->emissive_auto = false;
    } else {
        ctx->emissive_inst.clear();
        ctx->emissive_auto = true;
    }
    ctx->dirty = true;
    ctx->accel.built = false;
    return SNKPT_OK;
}

// Per instance: the previous frame's object->world rows (3 x float4) for
// contact velocities (transforms_prev_frame, ParticleUpdate.rchit:56-62); the
// current transforms until snkpt_set_prev_transforms supplies others.
static snkpt_status upload_prev_xf(snkpt_ctx* ctx) {
    const std::vector<float>& m = ctx->prev_mat4s.size() == ctx->mat4s.size() ? ctx->prev_mat4s : ctx->mat4s;
    std::vector<float> xf(12 * std::max<size_t>(1, ctx->inst.size()), 0.0f);
    const size_t n_mat4 = m.size() / 16;
    for (size_t i = 0; i < ctx->inst.size(); ++i) {
        const uint32_t ti = ctx->inst[i].transform_idx;
        if (ti >= n_mat4) return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": transform_idx out of range");
        const float* a = &m[16 * (size_t)ti];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) xf[12 * i + 4 * r + c] = a[4 * c + r];
    }
    return upload(ctx, ctx->d_prev_xf, xf.data(), xf.size() * 4);
}

snkpt_status snkpt_set_prev_transforms(snkpt_ctx* ctx, const float* mat4s, uint32_t n_mat4) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if (n_mat4 && !mat4s) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_prev_transforms: null array");
    if ((size_t)n_mat4 * 16 != ctx->mat4s.size())
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_set_prev_transforms: count differs from the current transforms");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    ctx->prev_mat4s.assign(mat4s, mat4s + 16 * (size_t)n_mat4);
    return upload_prev_xf(ctx);
}

snkpt_status snkpt_trace_contacts(snkpt_ctx* ctx, const snkpt_rays* rays, uint32_t n, const snkpt_contacts* out,
                                  float timestep, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!rays || !out || (n && (!rays->org_tmin || !rays->dir_tmax || !out->t || !out->normal || !out->position ||
                                !out->velocity)))
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_trace_contacts: null rays or contact arrays");
    if (!(timestep > 0.0f)) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_trace_contacts: timestep must be > 0");
    if ((((uintptr_t)rays->org_tmin | (uintptr_t)rays->dir_tmax | (uintptr_t)out->normal | (uintptr_t)out->position |
          (uintptr_t)out->velocity) & 15u) != 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_trace_contacts: float4 arrays must be 16-byte aligned");
    return trace_contacts(ctx, (const float4*)rays->org_tmin, (const float4*)rays->dir_tmax, n, *out, timestep,
                          (hipStream_t)stream);
}

snkpt_status snkpt_build_accel(snkpt_ctx* ctx, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    hipStream_t st = (hipStream_t)stream;
    snkpt_status s;
    // only what changed since the last build is uploaded again (a per-frame
    // rebuild after new instance transforms moves 64 B per transform, not the meshes)
    if (ctx->mesh_dirty) {
        if ((s = upload(ctx, ctx->d_pos, ctx->pos.data(), ctx->pos.size() * 4)) != SNKPT_OK) return s;
        if ((s = upload(ctx, ctx->d_nrm, ctx->nrm.data(), ctx->nrm.size() * 4)) != SNKPT_OK) return s;
        if ((s = upload(ctx, ctx->d_uv, ctx->uv.data(), ctx->uv.size() * 4)) != SNKPT_OK) return s;
        if ((s = upload(ctx, ctx->d_idx, ctx->idx.data(), ctx->idx.size() * 4)) != SNKPT_OK) return s;
        ctx->mesh_dirty = false;
    }
    if (ctx->inst_dirty) {
        if ((s = upload(ctx, ctx->d_inst, ctx->inst.data(), ctx->inst.size() * sizeof(snkpt_instance))) != SNKPT_OK)
            return s;
        if ((s = upload(ctx, ctx->d_mat4, ctx->mat4s.data(), ctx->mat4s.size() * 4)) != SNKPT_OK) return s;
    }
    if (ctx->mats_dirty) {
        if ((s = upload(ctx, ctx->d_mats, ctx->mats.data(), ctx->mats.size() * sizeof(snkpt_material))) != SNKPT_OK)
            return s;
        ctx->mats_dirty = false;
    }
    std::vector<float> xf(24 * ctx->inst.size());
    const uint32_t n_mat4 = (uint32_t)(ctx->mat4s.size() / 16);
    for (size_t i = 0; i < ctx->inst.size(); ++i) {
        if (ctx->inst[i].transform_idx >= n_mat4)
            return fail(ctx, SNKPT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": transform_idx out of range");
        instance_xf(&ctx->mat4s[16 * (size_t)ctx->inst[i].transform_idx], &xf[24 * i]);
    }
    if ((s = upload(ctx, ctx->d_inst_xf, xf.data(), xf.size() * 4)) != SNKPT_OK) return s;
    if (ctx->prev_mat4s.size() != ctx->mat4s.size()) ctx->prev_mat4s.clear();  // stale after a new transform set
    if ((s = upload_prev_xf(ctx)) != SNKPT_OK) return s;
    ctx->inst_dirty = false;
    ctx->dirty = false;
    if ((s = build_accel(ctx, st)) != SNKPT_OK) return s;
    return autotune_trace(ctx, st);
}

snkpt_status snkpt_get_trace_tuning(const snkpt_ctx* ctx, snkpt_trace_tuning* out) {
    if (!ctx || !out) return SNKPT_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    const Accel& A = ctx->accel;
    for (int k = 0; k < 2; ++k) {
        out->refill_lanes[k] = A.shape[k].refill;
        out->steps_per_check[k] = A.shape[k].steps;
        out->blocks_per_cu[k] = A.shape[k].blocks_per_cu;
        out->ms[k] = A.shape[k].ms;
    }
    out->fixed = A.shape_fixed ? 1u : 0u;
    return SNKPT_OK;
}

snkpt_status snkpt_set_trace_tuning(snkpt_ctx* ctx, const snkpt_trace_tuning* t) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    Accel& A = ctx->accel;
    if (!t) {
        A.shape_fixed = false;  // back to the autotuner (at the next build)
        return SNKPT_OK;
    }
    for (int k = 0; k < 2; ++k)
        if (t->refill_lanes[k] < 1 || t->refill_lanes[k] > 64 || (t->steps_per_check[k] != 3 &&
            t->steps_per_check[k] != 4) || t->blocks_per_cu[k] < 1 || t->blocks_per_cu[k] > 64)
            return fail(ctx, SNKPT_ERR_INVALID_ARG,
                        "snkpt_set_trace_tuning: refill_lanes 1..64, steps_per_check 3 or 4, blocks_per_cu 1..64");
    ++ctx->generation;
    for (int k = 0; k < 2; ++k) A.shape[k] = TraceShape{t->refill_lanes[k], t->steps_per_check[k], t->blocks_per_cu[k], 0.0f};
    A.shape_fixed = true;
    return SNKPT_OK;
}

snkpt_status snkpt_get_accel_info(const snkpt_ctx* ctx, snkpt_accel_info* out) {
    if (!ctx || !out) return SNKPT_ERR_INVALID_ARG;
    const Accel& A = ctx->accel;
    if (!A.built) return SNKPT_ERR_NOT_BUILT;
    std::memset(out, 0, sizeof(*out));
    out->num_triangles = A.num_tris;
    out->num_nodes = A.num_nodes;
    out->num_leaf_tris = A.num_tris;
    out->depth = A.depth;
    out->binary_depth = A.binary_depth;
    out->num_emitters = ctx->emit.count;
    for (int a = 0; a < 3; ++a) {
        out->scene_lo[a] = A.lo[a];
        out->scene_hi[a] = A.hi[a];
    }
    out->build_ms = A.build_ms;
    return SNKPT_OK;
}

snkpt_status snkpt_accel_export(snkpt_ctx* ctx, uint32_t which, void* host_dst, size_t bytes, size_t* out_bytes) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    return accel_export(ctx, which, host_dst, bytes, out_bytes);
}

snkpt_status snkpt_trace(snkpt_ctx* ctx, const snkpt_rays* rays, uint32_t n, const snkpt_hits* hits, int any_hit,
                         snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!rays || !hits) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_trace: null rays or hits");
    return launch_trace(ctx, (const float4*)rays->org_tmin, (const float4*)rays->dir_tmax, n, *hits, any_hit != 0,
                        nullptr, nullptr, nullptr, (hipStream_t)stream);
}

snkpt_status snkpt_trace_counted(snkpt_ctx* ctx, const snkpt_rays* rays, uint32_t n, const snkpt_hits* hits,
                                 int any_hit, uint32_t* node_visits, uint32_t* tri_visits, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!rays || !hits) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_trace_counted: null rays or hits");
    if (!node_visits && !tri_visits)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_trace_counted: no counter arrays given");
    return launch_trace(ctx, (const float4*)rays->org_tmin, (const float4*)rays->dir_tmax, n, *hits, any_hit != 0,
                        node_visits, tri_visits, nullptr, (hipStream_t)stream);
}

snkpt_status snkpt_render(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* settings, void* out_rgba,
                          uint32_t w, uint32_t h, uint32_t out_format, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!cam || !settings || !out_rgba || w == 0 || h == 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render: null camera/settings/output or empty image");
    if (out_format != SNKPT_FORMAT_RGBA32F && out_format != SNKPT_FORMAT_RGBA16F)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render: unknown output format");
    return render(ctx, cam, settings, out_rgba, w, h, out_format, (hipStream_t)stream);
}

snkpt_status snkpt_render_gbuffer(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* settings,
                                  const snkpt_gbuffer* gbuffer, void* out_rgba, uint32_t w, uint32_t h,
                                  uint32_t out_format, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!cam || !settings || !gbuffer || !out_rgba || w == 0 || h == 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render_gbuffer: null camera/settings/G-buffer/output or "
                                                "empty image");
    if (!gbuffer->depth || !gbuffer->normal || !gbuffer->albedo || !gbuffer->rma || !gbuffer->mat_flags)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render_gbuffer: depth, normal, albedo, rma and mat_flags "
                                                "images are required (motion is optional)");
    if ((((uintptr_t)gbuffer->depth | (uintptr_t)gbuffer->albedo | (uintptr_t)gbuffer->motion) & 3u) != 0 ||
        (((uintptr_t)gbuffer->normal | (uintptr_t)gbuffer->rma) & 7u) != 0 || ((uintptr_t)gbuffer->mat_flags & 1u) != 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render_gbuffer: images must be aligned to their texel size");
    if (out_format != SNKPT_FORMAT_RGBA32F && out_format != SNKPT_FORMAT_RGBA16F)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_render_gbuffer: unknown output format");
    return render(ctx, cam, settings, out_rgba, w, h, out_format, (hipStream_t)stream, gbuffer);
}

snkpt_status snkpt_taa_resolve(snkpt_ctx* ctx, const void* current_rgba, void* out_rgba, const uint16_t* velocity,
                               uint32_t w, uint32_t h, uint32_t format, uint32_t reset, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!current_rgba || !out_rgba || !velocity || w == 0 || h == 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_taa_resolve: null image or empty size");
    if (format != SNKPT_FORMAT_RGBA32F && format != SNKPT_FORMAT_RGBA16F)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_taa_resolve: unknown format");
    if ((uint64_t)w * h > (1ull << 26) || (h + 15u) / 16u > 65535u || (w + 15u) / 16u > 65535u)
        return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_taa_resolve: image above 2^26 pixels or 2^20 rows / columns");
    const uintptr_t px = format == SNKPT_FORMAT_RGBA32F ? 15u : 7u;
    if ((((uintptr_t)current_rgba | (uintptr_t)out_rgba) & px) != 0 || ((uintptr_t)velocity & 3u) != 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_taa_resolve: images must be aligned to their texel size");
    return taa_resolve(ctx, current_rgba, out_rgba, velocity, w, h, format, reset, (hipStream_t)stream);
}

snkpt_status snkpt_get_generation(const snkpt_ctx* ctx, uint64_t* out) {
    if (!ctx || !out) return SNKPT_ERR_INVALID_ARG;
    *out = ctx->generation;
    return SNKPT_OK;
}

snkpt_status snkpt_get_frame_stats(snkpt_ctx* ctx, snkpt_frame_stats* out) {
    if (!ctx || !out) return SNKPT_ERR_INVALID_ARG;
    *out = ctx->stats;
    return SNKPT_OK;
}

snkpt_status snkpt_oracle_accumulate(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* settings,
                                     double* accum_rgb, uint32_t w, uint32_t h, uint32_t spp, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (!cam || !settings || !accum_rgb || w == 0 || h == 0)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_oracle_accumulate: null argument or empty image");
    return oracle_accumulate(ctx, cam, settings, accum_rgb, w, h, spp, (hipStream_t)stream);
}

snkpt_status snkpt_cache_reset(snkpt_ctx* ctx, uint32_t seed, const float* lo3, const float* hi3,
                               snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    ++ctx->generation;
    if ((lo3 == nullptr) != (hi3 == nullptr)) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_cache_reset: lo3/hi3");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    return cache_reset(ctx, seed, lo3, hi3, (hipStream_t)stream);
}

snkpt_status snkpt_cache_infer(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out_rgba,
                               snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (n && (!samples || !out_rgba)) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_cache_infer: null pointer");
    if (n > (1u << 26)) return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_cache_infer: more than 2^26 samples");
    return cache_infer(ctx, samples, n, out_rgba, (hipStream_t)stream);
}

snkpt_status snkpt_cache_infer_noalloc(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out_rgba,
                                       snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (n && (!samples || !out_rgba))
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_cache_infer_noalloc: null pointer");
    return cache_infer_noalloc(ctx, samples, n, out_rgba, (hipStream_t)stream);
}

snkpt_status snkpt_cache_train(snkpt_ctx* ctx, const float* samples, uint32_t n, float lr, int apply_adam,
                               float* out_loss, snkpt_stream stream) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (n && !samples) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_cache_train: null samples");
    if (n > (1u << 22)) return fail(ctx, SNKPT_ERR_CAPACITY, "snkpt_cache_train: more than 2^22 samples per step");
    return cache_train(ctx, samples, n, lr, apply_adam, out_loss, (hipStream_t)stream);
}

snkpt_status snkpt_cache_params(snkpt_ctx* ctx, uint32_t which, void* host_dst, size_t bytes, size_t* out_bytes) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    return cache_params(ctx, which, host_dst, bytes, out_bytes, nullptr);
}

snkpt_status snkpt_cache_set_params(snkpt_ctx* ctx, uint32_t which, const void* host_src, size_t bytes) {
    if (!ctx || !host_src) return SNKPT_ERR_INVALID_ARG;
    return cache_set_params(ctx, which, host_src, bytes, nullptr);
}

snkpt_status snkpt_scene_generate(uint32_t config, uint32_t variant, snkpt_scene** out) {
    return scene_generate(config, variant, out);
}

void snkpt_scene_free(snkpt_scene* scene) { scene_free(scene); }

snkpt_status snkpt_scene_camera(uint32_t config, uint32_t variant, uint32_t frame, uint32_t w, uint32_t h,
                                snkpt_camera* out) {
    return scene_camera(config, variant, frame, w, h, out);
}

snkpt_status snkpt_scene_view(uint32_t config, uint32_t view, uint32_t* out_variant, uint32_t* out_w,
                              uint32_t* out_h) {
    if (config != 5 || view >= 64 || !out_variant || !out_w || !out_h) return SNKPT_ERR_INVALID_ARG;
    *out_variant = view / 4;
    *out_w = *out_h = c5_view_side(view);
    return SNKPT_OK;
}

snkpt_status snkpt_scene_load(snkpt_ctx* ctx, const snkpt_scene* s) {
    if (!ctx || !s) return SNKPT_ERR_INVALID_ARG;
    ctx->range_max.clear();
    ctx->pos.clear();
    ctx->nrm.clear();
    ctx->uv.clear();
    ctx->tan.clear();
    ctx->idx.clear();
    snkpt_status st = snkpt_upload_mesh(ctx, s->pos, s->nrm, s->uv, s->tan, s->n_verts, s->idx, s->n_idx, nullptr,
                                        nullptr);
    if (st != SNKPT_OK) return st;
    if ((st = snkpt_set_instances(ctx, s->inst, s->n_inst, s->mat4s, s->n_mat4)) != SNKPT_OK) return st;
    if ((st = snkpt_set_materials(ctx, s->mats, s->n_mats)) != SNKPT_OK) return st;
    return snkpt_set_lights(ctx, nullptr, 0, s->emissive_instances, s->n_emissive);
}

}  // extern "C"
