/*
 * snkpt oracle — CPU restatement used ONLY as test infrastructure.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this library, and only as the checker or the timed CPU baseline.  The
 * product (snake-vk_amd/libsnkpt.so) never links, loads or calls it.
 *
 * What it restates and where each part is pinned (see the header comment of
 * each source file for the per-function pins):
 *   pcg32          O'Neill 2014 PCG32; pinned by the published pcg32 demo vectors.
 *   flatten        object->world transform of instances (TLAS.cpp:188-200 semantics,
 *                  RaytracingInstances.glsl:7-13 indexing); pinned by closed-form transforms.
 *   intersect      Woop, Benthin & Wald 2013 watertight test, both faces (Q5);
 *                  pinned by closed-form hits, double-precision t and shared-edge closure.
 *   build          LBVH (Karras 2012) + SAH collapse to BVH8 (Ylitie et al. 2017);
 *                  pinned by invariants (sortedness, range/split property, box containment,
 *                  every triangle in exactly one leaf) and by traversal == brute force.
 *   trace          BVH8 traversal and brute-force closest/any hit.
 *   shading / PT / ReSTIR / MLP: see snkpt_oracle_pt.c and snkpt_oracle_mlp.c.
 */
#ifndef SNKPT_ORACLE_H
#define SNKPT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PCG32 ---------------------------------------------------------------- */
typedef struct {
    uint64_t state, inc;
} o_pcg32;
void o_pcg32_seed(o_pcg32* r, uint64_t initstate, uint64_t initseq);
uint32_t o_pcg32_next(o_pcg32* r);
/* n outputs of pcg32(initstate, initseq) into out */
void o_pcg32_fill(uint64_t initstate, uint64_t initseq, uint32_t* out, uint32_t n);

/* ---- geometry ------------------------------------------------------------- */
/* World-space triangle records, 12 x 32-bit per triangle:
 * v0.xyz v1.xyz v2.xyz (float) then inst, prim, gid (u32). */
int o_flatten(const float* pos, uint32_t n_verts, const uint32_t* idx, uint32_t n_idx, const uint32_t* inst5,
              uint32_t n_inst, const float* mat4s, uint32_t n_mat4, float* out, uint32_t max_tris);

/* Watertight intersection of one ray (o, d) with one triangle; returns 1 on a
 * hit of the supporting triangle (no t range test) with t, u, v. */
int o_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                float* t, float* u, float* v);

/* ---- LBVH -> BVH8 build ----------------------------------------------------- */
typedef struct {
    uint32_t n;           /* triangles */
    uint32_t num_nodes;   /* BVH8 nodes */
    uint32_t depth;       /* BVH8 levels */
    uint32_t binary_depth;
    uint32_t* keys;       /* n, sorted Morton keys */
    uint32_t* ids;        /* n, sorted triangle ids */
    uint32_t* child;      /* 2*(n-1), tagged 0x80000000 | leaf */
    uint32_t* first;      /* n-1 */
    uint32_t* count;      /* n-1 */
    float* aabb;          /* 6*(2n-1): internal nodes then leaves */
    uint32_t* leafify;    /* n-1 */
    uint32_t* nodes;      /* 20 u32 per BVH8 node */
    float* leaf;          /* 12 x 32-bit per leaf record */
} o_bvh;

int o_build(const float* world, uint32_t n, o_bvh* out);
void o_bvh_free(o_bvh* b);

/* ---- tracing ---------------------------------------------------------------
 * rays: 8 floats per ray (ox oy oz tmin dx dy dz tmax).  Outputs as the
 * library's snkpt_hits (t = -1 and ids ~0 on a miss); u, v, nvis, tvis may be NULL. */
void o_trace_bvh8(const o_bvh* b, const float* rays, uint32_t n, int any_hit, float* t, uint32_t* inst,
                  uint32_t* prim, float* u, float* v, uint32_t* nvis, uint32_t* tvis, int threads);
void o_trace_brute(const float* world, uint32_t ntris, const float* rays, uint32_t n, int any_hit, float* t,
                   uint32_t* inst, uint32_t* prim, float* u, float* v, int threads);

void o_trace_one(const o_bvh* b, const float* ray, int any_hit, float* t, uint32_t* inst, uint32_t* prim, float* u,
                 float* v);

/* ---- shading and path tracing (snkpt_oracle_pt.c) -------------------------- */
uint64_t o_pass_state(uint32_t seed, uint32_t frame, uint32_t pass);
void o_camera_matrix(const float* view16, const float* proj16, float out16[16]);
void o_brdf_eval(const float albedo[3], float roughness, float metallic, uint32_t flags, const float n[3],
                 const float wo[3], const float wi[3], float out_f[3], float* out_pdf);
int o_brdf_sample(const float albedo[3], float roughness, float metallic, uint32_t flags, const float n[3],
                  const float wo[3], float ul, float u1, float u2, float out_wi[3]);
/* Emitter table + `n` draws of the emitter sampler (per emitter: triangle id, pdf, count). */
uint32_t o_emitter_sample(const float* world, uint32_t ntris, const uint32_t* inst5, uint32_t n_inst,
                          const uint32_t* mats16, const uint32_t* emissive, uint32_t n_emissive, uint64_t seed,
                          uint32_t n, uint32_t cap, uint32_t* out_gid, float* out_pdf, uint32_t* out_count);
/* The reference's own ReSTIR DI (SNKPT_MODE_REFERENCE_COMPAT's direct light) over `frames`
 * frames of a static camera; writes the last frame's emission + direct light. */
int o_compat_direct(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5, uint32_t n_inst,
                    const uint32_t* idx, const float* nrm, const float* mat4s, const uint32_t* mats16,
                    const uint32_t* emissive, uint32_t n_emissive, const uint8_t* light_block, size_t block_bytes,
                    const float* view16, const float* proj16, const float* proj_view16, const float* cam_pos,
                    uint32_t w, uint32_t h, uint32_t frames, uint32_t temporal, uint32_t spatial, float radius,
                    uint32_t seed, float* out_rgb, int threads);
/* Mean of `spp` samples per pixel (pass k uses o_pass_state(seed ^ seed_xor, frame, k)). */
int o_pt_render(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5, uint32_t n_inst,
                const uint32_t* idx, const float* nrm, const float* mat4s, const uint32_t* mats16,
                const uint32_t* emissive, uint32_t n_emissive, const float* view16, const float* proj16,
                const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth, uint32_t rr_start,
                uint32_t seed, uint32_t frame, uint32_t seed_xor, float* out_rgb, int threads);

/* G-buffer of the reference's formats (EditorLayer.cpp:227-302), host arrays. */
typedef struct o_gbuffer {
    const float* depth;     /* D32 clip depth, >= 1 = nothing rasterized */
    const uint16_t* normal; /* RGBA16F */
    const int8_t* albedo;   /* RGBA8 SNORM */
    const uint16_t* rma;    /* RGBA16F */
    const uint16_t* flags;  /* R16_UINT */
} o_gbuffer;
float o_half_to_float(uint16_t h);
/* o_pt_render with the primary vertex of each pixel read from the G-buffer. */
int o_pt_render_gbuffer(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5,
                        uint32_t n_inst, const uint32_t* idx, const float* nrm, const float* mat4s,
                        const uint32_t* mats16, const uint32_t* emissive, uint32_t n_emissive, const float* view16,
                        const float* proj16, const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp,
                        uint32_t max_depth, uint32_t rr_start, uint32_t seed, uint32_t frame, const o_gbuffer* gb,
                        float* out_rgb, int threads);
int o_compat_bounce(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5, uint32_t n_inst,
                    const uint32_t* idx, const float* nrm, const float* mat4s, const uint32_t* mats16,
                    const float* dir_lights, uint32_t n_dir, const float* view16, const float* proj16,
                    const float* cam_pos, uint32_t w, uint32_t h, uint32_t spp, uint32_t seed, uint32_t frame,
                    float* out_rgb, int threads);
void o_primary_surfaces(const float* world, uint32_t ntris, const o_bvh* bvh, const uint32_t* inst5,
                        uint32_t n_inst, const uint32_t* idx, const float* nrm, const float* mat4s,
                        const uint32_t* mats16, const float* view16, const float* proj16, const float* proj_view16,
                        const float* cam_pos, uint32_t w, uint32_t h, float* depth, float* normal3, float* albedo3,
                        float* rm2, uint32_t* flags);

/* ---- neural radiance cache (snkpt_oracle_mlp.c), FP64 ----------------------- */
void o_cache_geom(const float lo[3], const float hi[3], float out_lo_inv[4], uint32_t res[16], uint32_t dense[16]);
uint32_t o_grid_index(uint32_t x, uint32_t y, uint32_t z, uint32_t res, uint32_t dense);
void o_cache_encode(const float* lo_inv, const uint32_t* res, const uint32_t* dense, const float* grid,
                    const float* samples, uint32_t n, double* out);
void o_cache_infer(const float* lo_inv, const uint32_t* res, const uint32_t* dense, const float* grid, const float* W,
                   const float* samples, uint32_t n, double* out_y);
double o_cache_grads(const float* lo_inv, const uint32_t* res, const uint32_t* dense, const float* grid,
                     const float* W, const float* samples, uint32_t n, double* dW, double* dgrid,
                     const double* den_in, double* den_out);
void o_adam(float* p, float* m, float* v, const double* g, size_t n, uint32_t t, double lr, double b1, double b2,
            double eps);

#ifdef __cplusplus
}
#endif
#endif
