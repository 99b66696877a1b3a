/*
 * snkpt.h — C ABI of the MI355X-native path-trace pass.
 *
 * This library replaces the ray-tracing pass of orangefrog45/Snake-VK:
 *   RT::Init / RT::RecordRenderCmdBuf      Core/headers/rendering/Raytracing.h:11-31,
 *                                          Core/src/rendering/Raytracing.cpp:227-278
 * and the driver acceleration structures it traces against
 *   BLAS::GenerateFromMeshData, TLAS::BuildFromInstances  Core/src/resources/TLAS.cpp:10-200.
 *
 * Scene mesh / material / emitter buffers and a camera go in; an HDR radiance
 * image comes out.  Everything is plain C: pointers, sizes and status codes.
 * Host arrays are copied before a call returns.  Device pointers (rays, hits,
 * output images) belong to the caller and must be HIP device memory on the
 * context's device.  Work is enqueued on the caller's stream (a hipStream_t
 * passed as `snkpt_stream`, NULL = the null stream).
 *
 * Threading: a context is externally synchronised (one host thread at a time).
 * Several contexts may share a device.
 */
#ifndef SNKPT_H
#define SNKPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct snkpt_ctx snkpt_ctx;
typedef void* snkpt_stream; /* hipStream_t */
typedef int32_t snkpt_status;

#define SNKPT_OK 0
#define SNKPT_ERR_INVALID_ARG (-1)
#define SNKPT_ERR_OOM (-2)
#define SNKPT_ERR_HIP (-3)
#define SNKPT_ERR_CAPACITY (-4) /* a queue / table / stack limit would be exceeded: nothing is truncated */
#define SNKPT_ERR_NOT_BUILT (-5) /* trace/render before snkpt_build_accel */

/* Capacities of the reference's buffers (RaytracingBufferSystem.cpp:60-63,
 * TransformBufferSystem.h:70-76, ResourceBufferManagers.cpp:10,
 * LightBuffers.glsl:41-49).  The same limits are enforced here. */
#define SNKPT_MAX_INSTANCES 4096u
#define SNKPT_MAX_TRANSFORMS 4096u
#define SNKPT_MAX_MATERIALS 4096u
#define SNKPT_MAX_POINTLIGHTS 1024u
#define SNKPT_MAX_SPOTLIGHTS 1024u

#define SNKPT_INVALID_INDEX 0xFFFFFFFFu
#define SNKPT_MAT_FLAG_EMISSIVE 1u /* TexMatBuffers.glsl:3; emission = albedo.xyz */
#define SNKPT_MAT_FLAG_LAMBERT 2u  /* extension: pure albedo/pi BRDF (no reference counterpart) */

/* InstanceData — RaytracingInstances.glsl:7-13, RaytracingBufferSystem.h:14-28.
 * One record per (entity, submesh).  The hit's instance id is the record index
 * (gl_InstanceCustomIndexEXT, TLAS.cpp:193) and its primitive id is the
 * submesh-local triangle index (gl_PrimitiveID).  Vertex i of triangle p is
 *   idx[mesh_buffer_index_offset + 3p + i] + mesh_buffer_vertex_offset. */
typedef struct snkpt_instance {
    uint32_t transform_idx;
    uint32_t mesh_buffer_index_offset; /* in indices, a multiple of 3 */
    uint32_t mesh_buffer_vertex_offset;
    uint32_t material_idx;
    uint32_t num_mesh_indices;
} snkpt_instance; /* 20 B */

/* Material — TexMatBuffers.glsl:5-20, writer ResourceBufferManagers.cpp:64-101.
 * Texture indices must be SNKPT_INVALID_INDEX (textures are not part of this path). */
typedef struct snkpt_material {
    uint32_t albedo_tex_idx, normal_tex_idx, roughness_tex_idx, metallic_tex_idx, ao_tex_idx;
    float roughness, metallic, ao;
    float albedo[4];
    uint32_t flags, pad0, pad1, pad2;
} snkpt_material; /* 64 B */

/* CommonUBO — CommonUBO.glsl:1-13, SceneInfoBufferSystem.cpp:13-23,66-91.
 * view = lookAt, proj = perspective(radians(fov/2), aspect, near, far) with the
 * Vulkan Y flip (CameraComponent.h:15-19).  Column-major like glm.  Camera rays
 * are cam_pos -> unproject(pixel centre, depth 1) (Util.glsl:9-13). */
typedef struct snkpt_camera {
    float view[16];
    float proj[16];
    float proj_view[16];
    float cam_pos[4];
    float cam_forward[4];
    uint32_t frame_idx;
    float time_elapsed; /* unused: randomness comes from (seed, frame) only */
    float delta_time;   /* unused */
    uint32_t prev_frame_invalid; /* disables temporal reuse (RayGen.rgen:422) */
} snkpt_camera; /* 240 B */

/* Light SSBO header — LightBuffers.glsl:1-63, LightBufferSystem.cpp:65-108.
 * snkpt_set_lights accepts the block as laid out there: a 96 B directional
 * light, {num_pointlights, num_spotlights, pad, pad}, then 1024 x 48 B point
 * lights and 1024 x 64 B spot lights (the tail may be omitted). */
typedef struct snkpt_pointlight {
    float colour[4];
    float pos[4];
    float atten_exp, atten_linear, atten_constant, padding;
} snkpt_pointlight; /* 48 B */

#define SNKPT_MODE_PT 0u     /* unbiased path tracing, NEE + MIS */
#define SNKPT_MODE_RESTIR 1u /* ReSTIR DI for direct light at the primary hit + PT for the rest */
#define SNKPT_MODE_CACHE 2u  /* PT terminated into the MLP radiance cache, cache trained online */
/* The reference's own (biased) estimator, for regression against its look: its ReSTIR DI at the
 * primary hit (RayGen.rgen:215-528: 32 point-light candidates drawn uniformly and 32 emissive
 * triangles with pdf 1 / (N_inst N_tri area), target = length of the unshadowed contribution
 * without the emitter cosine, temporal and spatial combines weighted p_hat M W with visibility and
 * no MIS, a fresh reservoir without history discarded, each combine counting 1 into M; the
 * directional light shadowed) plus BounceDiffuse (RayGen.rgen:139-213): 8 uniform-hemisphere rays
 * from the biased primary position, each adding albedo * payload * max(cos(dot(rd, n)), 0) / 8, and
 * 2 more from every non-emissive hit at a quarter weight, all times 2 PI; the payload of a hit is
 * its emission, or the directional light through its BRDF and a shadow ray
 * (RayClosestHit.rchit:87-142).  Random numbers are PCG32 instead of the wall-clock gold noise;
 * max_depth and rr_start are ignored. */
#define SNKPT_MODE_REFERENCE_COMPAT 3u

#define SNKPT_FLAG_ACCUMULATE 1u /* add this frame to the running mean and output the mean */
#define SNKPT_FLAG_RESET 2u      /* clear the running mean first */
#define SNKPT_FLAG_NO_TRAIN 4u   /* cache mode: query only, skip the training step */
#define SNKPT_FLAG_WHOLE_PATHS 16u   /* plain PT: one thread runs each whole path (no queues); same image */
#define SNKPT_FLAG_CACHE_INDIRECT 8u /* cache mode: NEE at the query vertex, the cache learns and answers
                                        only the indirect radiance there */
#define SNKPT_FLAG_CACHE_MIS 32u     /* with SNKPT_FLAG_CACHE_INDIRECT: the query vertex's direct light is NEE
                                        and one traced BRDF sample under the power heuristic (one more
                                        closest-hit ray per query) instead of NEE alone, which bounds the
                                        1/r^2 outliers of shading points next to an emitter (snkpt_render,
                                        snkpt_batch_render) */
#define SNKPT_FLAG_NO_BOUNCE 128u   /* SNKPT_MODE_REFERENCE_COMPAT: direct light only (no BounceDiffuse term) */
#define SNKPT_FLAG_CACHE_TRAIN_AT_QUERY 256u /* cache mode: the training samples are the training paths' vertices at
                                                the query depth only (where the cache is asked), not every vertex */
#define SNKPT_FLAG_NO_STATS 64u      /* snkpt_render returns without reading the frame's ray counts back
                                        (snkpt_get_frame_stats keeps the previous values) and without
                                        waiting for the GPU.  With it a SNKPT_MODE_PT frame can be recorded
                                        into a HIP graph by stream capture, once an uncaptured frame of the
                                        same size has sized the context's buffers (no counterpart in the
                                        reference, whose command buffers are re-recorded per frame) */

#define SNKPT_FORMAT_RGBA32F 0u
#define SNKPT_FORMAT_RGBA16F 1u

/* RtSettings (Raytracing.h:13-16) plus what the unbiased pass needs. */
typedef struct snkpt_settings {
    int32_t num_temporal_resamples; /* 0 disables temporal reuse */
    int32_t num_spatial_resamples;  /* spatial neighbours (reference default 3, config 5) */
    uint32_t mode;                  /* SNKPT_MODE_* */
    uint32_t spp;                   /* samples per pixel per call (>= 1) */
    uint32_t max_depth;             /* path vertices after the camera (1 = direct light only) */
    uint32_t rr_start;              /* first bounce with Russian roulette; 0 = off */
    uint32_t seed;                  /* RNG seed; streams are (seed, frame, pixel, sample) */
    uint32_t frame;                 /* frame index for RNG streams */
    uint32_t restir_candidates;     /* RIS candidates per pixel (LIGHT_RIS_ITERS, RayGen.rgen:95) */
    uint32_t restir_m_cap;          /* temporal M cap (M_CAP, RayGen.rgen:97) */
    float restir_radius;            /* spatial radius in pixels (RayGen.rgen:488) */
    uint32_t cache_query_depth;     /* cache mode: path vertex answered by the cache (1 = primary hit; 0 = 2) */
    uint32_t cache_train_paths;     /* cache mode: training paths per frame */
    float cache_lr;                 /* cache mode: Adam learning rate */
    uint32_t flags;                 /* SNKPT_FLAG_* */
    float cache_ema;                /* cache mode: decay of the weight average used for queries (0 = raw weights) */
    uint32_t cache_train_steps;     /* cache mode: optimiser steps per frame on its training samples (0 = 1, <= 16) */
    float accum_weight;             /* SNKPT_FLAG_ACCUMULATE (snkpt_render): weight of this frame's samples in the
                                       running mean (0 = 1); rising weights let a still-learning cache's early frames
                                       fade out of the mean */
    uint32_t tail_bounce;           /* SNKPT_MODE_PT (and the path-traced bounces of ReSTIR and G-buffer frames):
                                       bounce from which the paths still alive run one thread per path instead of one
                                       launch per bounce; same radiance bit for bit.  0 = chosen from the context's
                                       previous frame (the first bounce whose queue of paths fits twice the resident
                                       lanes); >= max_depth = never */
    uint32_t reserved[1];
} snkpt_settings;

/* Ray and hit streams for snkpt_trace: structure-of-arrays, device pointers.
 * org_tmin[i] = (ox, oy, oz, tmin), dir_tmax[i] = (dx, dy, dz, tmax), 16 B aligned. */
typedef struct snkpt_rays {
    const float* org_tmin;
    const float* dir_tmax;
} snkpt_rays;

/* Closest hit: t, instance id, primitive id, barycentrics (u, v) with
 * bary = (1-u-v, u, v) as in RayClosestHit.rchit:88.  Miss: t = -1 and
 * inst = prim = SNKPT_INVALID_INDEX (RayMiss.rmiss:22).  Any-hit: t and ids of
 * the first occluder found (any one), or the miss values.  u / v may be NULL. */
typedef struct snkpt_hits {
    float* t;
    uint32_t* inst;
    uint32_t* prim;
    float* u;
    float* v;
} snkpt_hits;

typedef struct snkpt_accel_info {
    uint32_t num_triangles;   /* world-space triangles (all instances) */
    uint32_t num_nodes;       /* BVH8 nodes, 80 B each */
    uint32_t num_leaf_tris;   /* leaf triangle records, 48 B each (== num_triangles) */
    uint32_t depth;           /* BVH8 depth (root = 1) */
    uint32_t binary_depth;    /* LBVH depth before the collapse */
    uint32_t num_emitters;    /* emissive triangles */
    float scene_lo[3];
    float scene_hi[3];
    double build_ms;          /* device time of the last build */
} snkpt_accel_info;

/* Launch shape of the context's persistent traversal kernel (snkpt_trace,
 * snkpt_trace_counted and the extension rays of snkpt_render), [0] closest hit,
 * [1] any hit: a wave hands new rays to its idle lanes once refill_lanes of them
 * are idle, taking steps_per_check traversal steps between two such checks, with
 * blocks_per_cu resident 128-lane blocks per CU.  snkpt_build_accel picks it with
 * an autotuner (timed on 2^20 bounce-like rays leaving the scene's triangles; kept per device,
 * stack size and scene size, so rebuilds of similar scenes reuse it) unless the
 * caller fixed it with snkpt_set_trace_tuning (NULL hands it back to the
 * autotuner).  The shape never changes a hit.  No reference counterpart (the
 * driver schedules traceRaysKHR). */
typedef struct snkpt_trace_tuning {
    uint32_t refill_lanes[2];    /* 1..64 */
    uint32_t steps_per_check[2]; /* 3 or 4 */
    uint32_t blocks_per_cu[2];   /* 1..64 (capped by the LDS the stack needs) */
    float ms[2];                 /* the autotuner's time for its 2^20-ray sample (0: set by hand) */
    uint32_t fixed;              /* 1: set by snkpt_set_trace_tuning */
    uint32_t reserved[3];
} snkpt_trace_tuning;

/* ---- context ------------------------------------------------------------- */
snkpt_status snkpt_create(int hip_device, snkpt_ctx** out);
void snkpt_destroy(snkpt_ctx* ctx);
const char* snkpt_last_error(const snkpt_ctx* ctx);

/* ---- scene input ----------------------------------------------------------
 * MeshBufferManager::LoadMeshFromData (MeshBufferManager.cpp:22-88): appends a
 * mesh to the global SoA buffers.  pos/nrm/tan are vec3 (12 B stride), uv is
 * vec2; nrm/uv/tan may be NULL (zeros).  idx are submesh-local u32.  Returns
 * the vertex and index offsets of the mesh in the global buffers. */
snkpt_status snkpt_upload_mesh(snkpt_ctx* ctx, const float* pos, const float* nrm, const float* uv,
                               const float* tan, uint32_t n_verts, const uint32_t* idx, uint32_t n_idx,
                               uint32_t* out_vertex_offset, uint32_t* out_index_offset);
/* RaytracingInstanceBufferSystem::UpdateInstanceBuffer (RaytracingBufferSystem.cpp:69-103)
 * + TransformBufferSystem::UpdateTransformBuffer (TransformBufferSystem.h:92-122):
 * instance records and column-major mat4 transforms.  Replaces the previous set. */
snkpt_status snkpt_set_instances(snkpt_ctx* ctx, const snkpt_instance* inst, uint32_t n,
                                 const float* mat4s, uint32_t n_mat4);
/* GlobalMaterialBufferManager::UpdateMaterialUBO (ResourceBufferManagers.cpp:64-101). */
snkpt_status snkpt_set_materials(snkpt_ctx* ctx, const snkpt_material* m, uint32_t n);
/* LightBufferSystem::UpdateLightSSBO (LightBufferSystem.cpp:65-108) and
 * UpdateEmissiveIdxBuffer (RaytracingBufferSystem.cpp:105-130).  light_block may
 * be NULL (no analytic lights).  emissive_instances lists the emissive instance
 * records; NULL means "every instance whose material has the emissive flag". */
snkpt_status snkpt_set_lights(snkpt_ctx* ctx, const void* light_block, size_t bytes,
                              const uint32_t* emissive_instances, uint32_t n_emissive);

/* BLAS::GenerateFromMeshData + TlasSystem::BuildTlas (TLAS.cpp:10-200,
 * TlasSystem.cpp:27-45): flattens every instance to world space and builds one
 * BVH8 on the device (Morton30 -> radix sort -> Karras -> SAH collapse ->
 * 8-bit quantised children), plus the emitter table.  Synchronises `stream`. */
snkpt_status snkpt_build_accel(snkpt_ctx* ctx, snkpt_stream stream);
snkpt_status snkpt_get_accel_info(const snkpt_ctx* ctx, snkpt_accel_info* out);
snkpt_status snkpt_get_trace_tuning(const snkpt_ctx* ctx, snkpt_trace_tuning* out);
snkpt_status snkpt_set_trace_tuning(snkpt_ctx* ctx, const snkpt_trace_tuning* tuning);

/* Copies one build array to host memory for bit-exact checks.  `which`:
 *   0 sorted Morton keys (u32 / triangle)     1 sorted triangle ids (u32 / triangle)
 *   2 binary nodes: left, right (u32 x 2 / internal node; leaves are tagged 0x80000000|leaf)
 *   3 binary AABBs (float x 6 / node, internal nodes first then leaves)
 *   4 BVH8 nodes (80 B / node)                5 leaf triangles (48 B / record)
 *   6 world triangles in input order (float x 9 / triangle)
 *   7 binary-node collapse flags (u32 / internal node: bit0 left-is-leaf-cluster, bit1 right)
 * `bytes` must be at least the array's size; *out_bytes receives it. */
snkpt_status snkpt_accel_export(snkpt_ctx* ctx, uint32_t which, void* host_dst, size_t bytes,
                                size_t* out_bytes);

/* ---- tracing --------------------------------------------------------------
 * traceRayEXT with gl_RayFlagsOpaqueEXT (closest hit, RayGen.rgen:164-176) or
 * TerminateOnFirstHit|Opaque|SkipClosestHit (any hit, RayGen.rgen:101-121).
 * Both faces are hit (Q5: no cull flags).  Ray i's result is at hits[i]. */
snkpt_status snkpt_trace(snkpt_ctx* ctx, const snkpt_rays* rays, uint32_t n, const snkpt_hits* hits,
                         int any_hit, snkpt_stream stream);
/* Same traversal, also counting BVH8 nodes and triangles visited per ray into
 * node_visits / tri_visits (device u32 arrays, may be NULL). */
snkpt_status snkpt_trace_counted(snkpt_ctx* ctx, const snkpt_rays* rays, uint32_t n,
                                 const snkpt_hits* hits, int any_hit, uint32_t* node_visits,
                                 uint32_t* tri_visits, snkpt_stream stream);

/* Contact record of a short collision ray (ParticleUpdate.rchit:38-62), one per
 * ray, device arrays.  Miss: t = -1, zero vectors, ids SNKPT_INVALID_INDEX. */
typedef struct snkpt_contacts {
    float* t;          /* gl_RayTmaxEXT at the first hit found */
    float* normal;     /* float4 (xyz, 0): normalize(transpose(inverse(mat3(M))) * interpolated normal) */
    float* position;   /* float4: origin + direction * t */
    float* velocity;   /* float4: (position - M_prev * interpolated object-space point) / timestep */
    uint32_t* inst;    /* may be NULL */
    uint32_t* prim;    /* may be NULL */
} snkpt_contacts;

/* ParticleUpdate.rgen:92-167 (TraceCollisionRay, ConstrainOffsetWithScene,
 * CalculateGravity): rays traced with gl_RayFlagsTerminateOnFirstHitEXT (the
 * any-hit traversal of snkpt_trace, same first hit) and the contact record
 * of ParticleUpdate.rchit.  Rays as for snkpt_trace (org_tmin, dir_tmax). */
snkpt_status snkpt_trace_contacts(snkpt_ctx* ctx, const snkpt_rays* rays, uint32_t n, const snkpt_contacts* out,
                                  float timestep, snkpt_stream stream);
/* transforms_prev_frame (TransformBufferSystem.h:92-122): the previous frame's
 * column-major mat4s, same count as snkpt_set_instances' (host array, copied).
 * Until set, and after a new transform set, the current transforms are used
 * (zero velocity for static geometry). */
snkpt_status snkpt_set_prev_transforms(snkpt_ctx* ctx, const float* mat4s, uint32_t n_mat4);

/* ---- Vulkan interop ----------------------------------------------------------
 * The reference's output image and scene buffers are Vulkan memory
 * (Raytracing.cpp:96-147, S_VkBuffer.cpp:5-22).  Exported with
 * VK_KHR_external_memory_fd (vkGetMemoryFdKHR, OPAQUE_FD; the allocation made
 * with VkExportMemoryAllocateInfo), a range of it is imported as a device
 * pointer for every entry point above and below.  On success the library owns
 * the fd (do not close it); release the mapping before freeing the Vulkan
 * memory.  A linear-tiled image's rows are tightly packed only when its row
 * pitch is w * texel size. */
typedef struct snkpt_external snkpt_external;
snkpt_status snkpt_import_memory_fd(snkpt_ctx* ctx, int fd, uint64_t allocation_size, uint64_t offset,
                                    uint64_t size, void** out_dev_ptr, snkpt_external** out);
snkpt_status snkpt_release_memory(snkpt_ctx* ctx, snkpt_external* mem);
/* VK_KHR_external_semaphore_fd (OPAQUE_FD; timeline != 0 for a timeline
 * semaphore): the frame fence between the Vulkan queue and the HIP stream.
 * wait / signal are enqueued on the stream (value: timeline semaphores only). */
typedef struct snkpt_semaphore snkpt_semaphore;
snkpt_status snkpt_import_semaphore_fd(snkpt_ctx* ctx, int fd, int timeline, snkpt_semaphore** out);
snkpt_status snkpt_semaphore_wait(snkpt_ctx* ctx, snkpt_semaphore* sem, uint64_t value, snkpt_stream stream);
snkpt_status snkpt_semaphore_signal(snkpt_ctx* ctx, snkpt_semaphore* sem, uint64_t value, snkpt_stream stream);
snkpt_status snkpt_release_semaphore(snkpt_ctx* ctx, snkpt_semaphore* sem);

/* ---- the pass -------------------------------------------------------------
 * RT::RecordRenderCmdBuf (Raytracing.cpp:227-267): renders one frame of
 * settings->spp samples per pixel into out_rgba (device, w*h RGBA, row-major,
 * row 0 at the top).  With SNKPT_FLAG_ACCUMULATE the output is the running mean
 * over all accumulated frames.  On a stream under capture only SNKPT_MODE_PT and
 * SNKPT_MODE_CACHE frames with SNKPT_FLAG_NO_STATS are accepted, at the width
 * and height of the last uncaptured frame (a cache-mode frame also needs an
 * uncaptured cache-mode frame with the same depth and training paths first;
 * SNKPT_ERR_INVALID_ARG otherwise); a replay renders the captured camera, seed
 * and frame index again (a cache-mode replay also takes one more optimiser
 * step, as an uncaptured frame does), and is valid until the context's
 * generation (snkpt_get_generation) changes: a larger or differently shaped
 * frame, any scene-input call, snkpt_build_accel, snkpt_cache_reset or
 * snkpt_destroy invalidates it. */
snkpt_status snkpt_render(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* settings,
                          void* out_rgba, uint32_t w, uint32_t h, uint32_t out_format,
                          snkpt_stream stream);

/* Rasterized G-buffer of the primary visibility, in the reference's image
 * formats (EditorLayer.cpp:227-302, written by GBuffer.frag:46-72).  Device
 * pointers, w*h texels each, row-major, row 0 at the top, tightly packed,
 * each aligned to its texel size. */
typedef struct snkpt_gbuffer {
    const float* depth;        /* D32_SFLOAT: clip-space depth in [0, 1]; >= 1 (cleared) = nothing rasterized */
    const uint16_t* normal;    /* R16G16B16A16_SFLOAT: world-space shading normal in rgb */
    const int8_t* albedo;      /* R8G8B8A8_SNORM: albedo in rgb (the emission of an emissive material) */
    const uint16_t* rma;       /* R16G16B16A16_SFLOAT: roughness, metallic, ao */
    const uint16_t* mat_flags; /* R16_UINT: material flags (SNKPT_MAT_FLAG_*) */
    const uint16_t* motion;    /* R16G16_SFLOAT: history texcoord - texcoord (GBuffer.frag:40-42);
                                  NULL: reproject with the previous frame's camera */
} snkpt_gbuffer;

/* RayGen.rgen:531-555: the frame of snkpt_render with the primary vertex read
 * from a G-buffer instead of traced camera rays.  World position from depth
 * through inverse(view) * inverse(proj) (Util.glsl:1-7); rays leave it with the
 * reference's bias 0.01 + max(0.001 |cam - p|, 0.01) along the normal
 * (RayGen.rgen:552); an emissive texel's radiance is its albedo
 * (RayGen.rgen:539-542).  Every mode; in the cache mode the training paths
 * still start with camera rays.  Same output, accumulation and statistics
 * (camera_rays = 0). */
snkpt_status snkpt_render_gbuffer(snkpt_ctx* ctx, const snkpt_camera* cam, const snkpt_settings* settings,
                                  const snkpt_gbuffer* gbuffer, void* out_rgba, uint32_t w, uint32_t h,
                                  uint32_t out_format, snkpt_stream stream);

/* TAA_ResolvePass::RecordCommandBuffer (TAA_ResolvePass.cpp:39-56,
 * TAA_Resolve.comp:37-86): temporal resolve of a rendered frame.  History
 * texcoord = texcoord + velocity; the bilinear (repeat-addressed) history is
 * clipped to the 3x3 neighbourhood's mean +- 0.55 sigma box, blended 0.9 with
 * the current frame, and replaced by the neighbourhood mean where
 * |previous velocity + velocity| exceeds 0.001 (weight (len - 0.001) * 10).
 * The context keeps the history and previous-velocity images (the pass's two
 * blits).  current_rgba and out_rgba: device, w*h RGBA in `format`, may be the
 * same buffer (every pixel reads the frame before the resolve); velocity:
 * R16G16_SFLOAT, w*h.  reset != 0 (or a new size / format) starts a new
 * history: out = current. */
snkpt_status snkpt_taa_resolve(snkpt_ctx* ctx, const void* current_rgba, void* out_rgba, const uint16_t* velocity,
                               uint32_t w, uint32_t h, uint32_t format, uint32_t reset, snkpt_stream stream);

/* State generation of a context: changes whenever a call changes its scene data
 * or may move a device buffer that launched work reads (snkpt_upload_mesh,
 * _set_instances, _set_materials, _set_lights, _set_prev_transforms,
 * _build_accel, _cache_reset, and a render whose frame size differs from the
 * last one or grows the render buffers).  A frame captured into a HIP graph
 * (SNKPT_FLAG_NO_STATS) replays correctly only while the generation is the one
 * read right after the capture; a batch compares its scenes' generations and
 * rebuilds its tables and graph when one changed.  No reference counterpart
 * (the reference re-records its command buffer every frame). */
snkpt_status snkpt_get_generation(const snkpt_ctx* ctx, uint64_t* out);

/* Rays traced by the last snkpt_render, by kind. */
typedef struct snkpt_frame_stats {
    uint64_t camera_rays;
    uint64_t extension_rays; /* closest-hit rays after the camera ray */
    uint64_t shadow_rays;
    uint64_t cache_queries;
    uint64_t cache_train_samples;
    uint64_t restir_history_reused; /* SNKPT_MODE_RESTIR: pixels whose reprojected history was accepted */
} snkpt_frame_stats;
snkpt_status snkpt_get_frame_stats(snkpt_ctx* ctx, snkpt_frame_stats* out);

/* Converged-image oracle: adds `spp` samples per pixel of the plain path
 * tracer to accum_rgb (device, w*h*3 doubles, sums).  The only FP64 on the
 * device; tracing and shading stay FP32. */
snkpt_status snkpt_oracle_accumulate(snkpt_ctx* ctx, const snkpt_camera* cam,
                                     const snkpt_settings* settings, double* accum_rgb, uint32_t w,
                                     uint32_t h, uint32_t spp, snkpt_stream stream);

/* ---- MLP radiance cache (no reference counterpart) -------------------------
 * A training / query sample is 16 floats: position[3], direction[3] (towards the
 * previous path vertex), normal[3], roughness, albedo[3], target radiance[3]
 * (ignored for queries).  Sample and output pointers are device memory.  The
 * cache is a 16-level hash grid (2 FP16 features, 2^19 entries per level) feeding
 * a 64-wide MLP with 4 hidden ReLU layers and 3 outputs, run on MFMA in FP16
 * with FP32 accumulation; Adam (beta 0.9 / 0.99, eps 1e-15). */
#define SNKPT_CACHE_SAMPLE_FLOATS 16
/* (Re)initialises weights, grid and Adam state.  lo3/hi3: box the positions are
 * normalised by; NULL = the built scene's bounds. */
snkpt_status snkpt_cache_reset(snkpt_ctx* ctx, uint32_t seed, const float* lo3, const float* hi3,
                               snkpt_stream stream);
/* out_rgba: n x 4 floats (radiance rgb, 0). */
snkpt_status snkpt_cache_infer(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out_rgba,
                               snkpt_stream stream);
/* snkpt_cache_infer for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same kernels and bit-identical outputs, but the
 * encoded queries go to a static 8 MiB device workspace instead of buffers the
 * context grows: n <= 65,536 per call (larger n: SNKPT_ERR_CAPACITY), and no
 * hipMalloc / hipFree or allocating library call is made.  Every call on one GPU
 * shares that workspace, so calls that overlap on one GPU (from any context or
 * stream) race; serialise them.  The cache must have been set up with
 * snkpt_cache_reset (which does allocate). */
snkpt_status snkpt_cache_infer_noalloc(snkpt_ctx* ctx, const float* samples, uint32_t n, float* out_rgba,
                                       snkpt_stream stream);
/* One optimisation step on n samples (relative L2 loss).  apply_adam = 0 only
 * computes the gradients (readable through snkpt_cache_params 2 and 3). */
snkpt_status snkpt_cache_train(snkpt_ctx* ctx, const float* samples, uint32_t n, float lr, int apply_adam,
                               float* out_loss, snkpt_stream stream);
/* which: 0 MLP fp32 weights (17,408: W0..W3 64x64, W4 16x64 row-major [out][in], rows
 * 3..15 of W4 zero), 1 grid fp32 features (16 x 2^19 x 2), 2 MLP gradient of the last
 * train call, 3 grid gradient of the last train call made without Adam,
 * 4 encoded inputs of the last call (fp16, n x 64); of the last cache-mode render
 * pass: 5 the queries (16 floats each, target fields unused), 6 their records
 * (float4: throughput rgb at the query vertex, path index as uint32 bits),
 * 7 the training samples built from that pass's training paths. */
snkpt_status snkpt_cache_params(snkpt_ctx* ctx, uint32_t which, void* host_dst, size_t bytes,
                                size_t* out_bytes);
snkpt_status snkpt_cache_set_params(snkpt_ctx* ctx, uint32_t which, const void* host_src, size_t bytes);

/* ---- synthetic scenes (no reference counterpart: workload generators) ------
 * config 1..5 as described in the scene generator; variant picks one of the
 * scenes of a multi-scene config (config 5: 0..15) and is 0 otherwise.
 * All arrays are host memory owned by the scene object. */
typedef struct snkpt_scene {
    float* pos;
    float* nrm;
    float* uv;
    float* tan;
    uint32_t n_verts;
    uint32_t* idx;
    uint32_t n_idx;
    snkpt_instance* inst;
    uint32_t n_inst;
    float* mat4s;
    uint32_t n_mat4;
    snkpt_material* mats;
    uint32_t n_mats;
    uint32_t* emissive_instances;
    uint32_t n_emissive;
    snkpt_camera camera;
    uint32_t width, height;
    uint32_t max_depth, rr_start;
    uint32_t seed;
} snkpt_scene;
snkpt_status snkpt_scene_generate(uint32_t config, uint32_t variant, snkpt_scene** out);
void snkpt_scene_free(snkpt_scene* scene);
/* Camera of frame f of the config's camera path (config 3: seeded 64-frame orbit arc
 * in front of the room's open side). */
snkpt_status snkpt_scene_camera(uint32_t config, uint32_t variant, uint32_t frame, uint32_t w,
                                uint32_t h, snkpt_camera* out);
/* upload_mesh + set_instances + set_materials + set_lights for a generated scene. */
snkpt_status snkpt_scene_load(snkpt_ctx* ctx, const snkpt_scene* scene);
/* Batched workload (config 5): view v (0..63) shows scene variant v / 4 at a
 * w x h resolution (side log-uniform in [128, 1024]). */
snkpt_status snkpt_scene_view(uint32_t config, uint32_t view, uint32_t* out_variant, uint32_t* out_w,
                              uint32_t* out_h);

/* ---- batched rendering (no reference counterpart: RT::RecordRenderCmdBuf,
 * Raytracing.cpp:227-267, renders one view of one scene per command buffer) ----
 * Many views of many scenes share one set of wavefront queues; each stage is
 * one persistent launch per bounce over every view, and a frame can be
 * captured once as a hipGraph and replayed.  Every view's image is
 * bit-identical to snkpt_render of that view alone with the view's seed
 * (plain path tracing, 1 spp per frame).  Scenes are borrowed contexts with a
 * built acceleration structure; they and the output buffers must outlive the
 * batch.  A scene may be rebuilt, re-lit or re-uploaded between batch frames:
 * the next snkpt_batch_render sees its changed generation (snkpt_get_generation)
 * and rebuilds the batch's tables and graph first.  A batch is externally
 * synchronised like a context. */
typedef struct snkpt_batch snkpt_batch;
typedef struct snkpt_batch_tuning {
    uint32_t trace_block;         /* lanes per traversal block: 64 or 128 (extension and shadow rays) */
    uint32_t trace_blocks_per_cu; /* persistent traversal grid = CUs x this */
    uint32_t shade_blocks_per_cu; /* persistent shade grid (256-lane blocks) = CUs x this */
    uint32_t queue_capacity;      /* paths per group of views processed together; 0 = all views at once */
    float ms_per_frame;           /* set by snkpt_batch_autotune: best measured frame time */
    uint32_t tail_bounce;         /* bounce from which the paths still alive run one thread per path (same images
                                     bit for bit); 0 = the frame's rr_start, 0xFFFFFFFF (or >= max_depth) = never */
    uint32_t reserved[2];
} snkpt_batch_tuning;
snkpt_status snkpt_batch_create(int hip_device, snkpt_batch** out);
void snkpt_batch_destroy(snkpt_batch* batch);
const char* snkpt_batch_last_error(const snkpt_batch* batch);
snkpt_status snkpt_batch_add_scene(snkpt_batch* batch, snkpt_ctx* scene, uint32_t* out_scene);
/* out_rgba_device: w*h RGBA pixels in out_format, row 0 at the top. */
snkpt_status snkpt_batch_add_view(snkpt_batch* batch, uint32_t scene, const snkpt_camera* cam, uint32_t w,
                                  uint32_t h, uint32_t seed, void* out_rgba_device, uint32_t out_format,
                                  uint32_t* out_view);
/* One frame of every view.  Uses mode (SNKPT_MODE_PT or SNKPT_MODE_CACHE),
 * max_depth, rr_start, frame, the ACCUMULATE / RESET flags and the cache
 * fields and flags of `s`.  use_graph: capture the frame on the first call (or
 * after the views, tuning, depth, cache settings or stream change) and replay
 * it afterwards (the frame index and flags are read from device memory on every
 * replay); needs a created (non-null) stream.  In the cache form each scene's
 * own network answers its views' queries and takes its optimiser steps on
 * training paths through its first view; a captured cache-form frame runs its
 * training paths as a second graph on a side stream beside the render. */
snkpt_status snkpt_batch_render(snkpt_batch* batch, const snkpt_settings* s, int use_graph, snkpt_stream stream);
/* Times whole frames of this batch over traversal block size, persistent grid
 * sizes, queue capacity and tail bounce, keeps the fastest (images do not depend on it). */
snkpt_status snkpt_batch_autotune(snkpt_batch* batch, const snkpt_settings* s, snkpt_stream stream,
                                  snkpt_batch_tuning* out);
snkpt_status snkpt_batch_set_tuning(snkpt_batch* batch, const snkpt_batch_tuning* tuning);
snkpt_status snkpt_batch_get_tuning(const snkpt_batch* batch, snkpt_batch_tuning* out);
/* Ray counts of the last frame (waits for the stream). */
snkpt_status snkpt_batch_frame_stats(snkpt_batch* batch, snkpt_frame_stats* out, snkpt_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* SNKPT_H */
