// Emitter table: every emissive triangle with its world-space vertices,
// emission, area, and a power-proportional alias table (Walker 1977 / Vose 1991).
//
// The reference samples an emissive *instance* uniformly, then a triangle of
// it uniformly (DoInitialWRS, RayGen.rgen:311-337, pdf at :337).  Sampling
// triangles by power instead keeps the estimator unbiased and lowers its
// variance for the many-light configuration; the emissive instance list
// (UpdateEmissiveIdxBuffer, RaytracingBufferSystem.cpp:105-130) still decides
// which instances are emitters.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "snkpt_host.h"

namespace snkpt {

snkpt_status build_emitters(snkpt_ctx* ctx, hipStream_t st) {
    Emitters& E = ctx->emit;
    E.count = 0;
    E.total_power = 0;
    const uint32_t n_inst = (uint32_t)ctx->inst.size();
    std::vector<uint32_t> off(n_inst + 1, 0);
    for (uint32_t i = 0; i < n_inst; ++i) off[i + 1] = off[i] + ctx->inst[i].num_mesh_indices / 3;
    const uint32_t n_tris = off[n_inst];

    std::vector<uint32_t> list;
    if (ctx->emissive_auto) {
        for (uint32_t i = 0; i < n_inst; ++i)
            if (ctx->mats[ctx->inst[i].material_idx].flags & SNKPT_MAT_FLAG_EMISSIVE) list.push_back(i);
    } else {
        for (uint32_t i : ctx->emissive_inst) {
            if (i >= n_inst) return fail(ctx, SNKPT_ERR_INVALID_ARG, "emissive instance index out of range");
            list.push_back(i);
        }
    }
    std::vector<float> tri;    // 12 floats per emitter
    std::vector<float> info;   // emission rgb, area
    std::vector<double> power;
    std::vector<uint32_t> gid_to_e(n_tris ? n_tris : 1, 0xFFFFFFFFu);
    for (uint32_t i : list) {
        const snkpt_instance& in = ctx->inst[i];
        const snkpt_material& mat = ctx->mats[in.material_idx];
        const float* m = &ctx->mat4s[16 * (size_t)in.transform_idx];
        double lum = 0.2126 * mat.albedo[0] + 0.7152 * mat.albedo[1] + 0.0722 * mat.albedo[2];
        for (uint32_t p = 0; p < in.num_mesh_indices / 3; ++p) {
            float w[9];
            for (int k = 0; k < 3; ++k) {
                uint32_t v = ctx->idx[in.mesh_buffer_index_offset + 3 * p + k] + in.mesh_buffer_vertex_offset;
                float x = ctx->pos[3 * (size_t)v], y = ctx->pos[3 * (size_t)v + 1], z = ctx->pos[3 * (size_t)v + 2];
                w[3 * k + 0] = ((m[0] * x + m[4] * y) + m[8] * z) + m[12];
                w[3 * k + 1] = ((m[1] * x + m[5] * y) + m[9] * z) + m[13];
                w[3 * k + 2] = ((m[2] * x + m[6] * y) + m[10] * z) + m[14];
            }
            double e1[3], e2[3];
            for (int a = 0; a < 3; ++a) {
                e1[a] = (double)w[3 + a] - w[a];
                e2[a] = (double)w[6 + a] - w[a];
            }
            double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                   cz = e1[0] * e2[1] - e1[1] * e2[0];
            double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            double pw = lum * area;
            if (!(pw > 0.0)) continue;  // zero-area or black emitters are never sampled
            for (int k = 0; k < 3; ++k) {
                tri.push_back(w[3 * k]);
                tri.push_back(w[3 * k + 1]);
                tri.push_back(w[3 * k + 2]);
                tri.push_back(0.0f);
            }
            info.push_back(mat.albedo[0]);
            info.push_back(mat.albedo[1]);
            info.push_back(mat.albedo[2]);
            info.push_back((float)area);
            power.push_back(pw);
            gid_to_e[off[i] + p] = (uint32_t)(power.size() - 1);
        }
    }
    const uint32_t ne = (uint32_t)power.size();
    double total = 0;
    for (double p : power) total += p;
    // Vose alias table
    std::vector<float> prob(ne ? ne : 1, 1.0f), pdf(ne ? ne : 1, 0.0f);
    std::vector<uint32_t> alias(ne ? ne : 1, 0);
    if (ne) {
        std::vector<double> scaled(ne);
        std::vector<uint32_t> small, large;
        for (uint32_t i = 0; i < ne; ++i) {
            scaled[i] = power[i] / total * ne;
            pdf[i] = (float)(power[i] / total);
            (scaled[i] < 1.0 ? small : large).push_back(i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t s = small.back(), l = large.back();
            small.pop_back();
            prob[s] = (float)scaled[s];
            alias[s] = l;
            scaled[l] = (scaled[l] + scaled[s]) - 1.0;
            if (scaled[l] < 1.0) {
                large.pop_back();
                small.push_back(l);
            }
        }
        for (uint32_t l : large) prob[l] = 1.0f, alias[l] = l;
        for (uint32_t s : small) prob[s] = 1.0f, alias[s] = s;
    }
    std::vector<float> rec(16 * (size_t)(ne ? ne : 1), 0.0f);
    for (uint32_t i = 0; i < ne; ++i) {
        float* r = &rec[16 * (size_t)i];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) r[4 * k + a] = tri[12 * (size_t)i + 4 * k + a];
        r[3] = prob[i];
        std::memcpy(&r[7], &alias[i], 4);
        r[11] = pdf[i];
        for (int a = 0; a < 4; ++a) r[12 + a] = info[4 * (size_t)i + a];
    }
    auto up = [&](DeviceBuffer& b, const void* src, size_t bytes) -> snkpt_status {
        hipError_t e = b.reserve(bytes ? bytes : 16);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        if (bytes) {
            e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync");
        }
        return SNKPT_OK;
    };
    // analytic lights of the light block (LightBuffers.glsl:1-63 layout): 4 float4 each
    std::vector<float> al;
    const std::vector<uint8_t>& lb = ctx->light_block;
    auto f32 = [&](size_t off) {
        float v;
        std::memcpy(&v, lb.data() + off, 4);
        return v;
    };
    auto push = [&](const float* w) { al.insert(al.end(), w, w + 16); };
    if (lb.size() >= 112) {
        const float dc[3] = {f32(0), f32(4), f32(8)};
        if (dc[0] > 0.0f || dc[1] > 0.0f || dc[2] > 0.0f) {
            double dx = f32(16), dy = f32(20), dz = f32(24), l = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (l > 0.0) {
                const float w[16] = {dc[0], dc[1], dc[2], 0.0f, (float)(dx / l), (float)(dy / l), (float)(dz / l), 0.0f,
                                     0, 0, 0, 0, 0, 0, 0, 0};
                push(w);
            }
        }
        uint32_t np, ns;
        std::memcpy(&np, lb.data() + 96, 4);
        std::memcpy(&ns, lb.data() + 100, 4);
        for (uint32_t i = 0; i < np; ++i) {
            const size_t o = 112 + 48 * (size_t)i;
            const float c[3] = {f32(o), f32(o + 4), f32(o + 8)};
            if (!(c[0] > 0.0f || c[1] > 0.0f || c[2] > 0.0f)) continue;
            const float w[16] = {c[0], c[1], c[2], 1.0f, f32(o + 16), f32(o + 20), f32(o + 24), 0.0f, 0, 0, 0, 0,
                                 f32(o + 40), f32(o + 36), f32(o + 32), 0.0f};
            push(w);
        }
        const size_t so = 112 + 48 * (size_t)SNKPT_MAX_POINTLIGHTS;
        for (uint32_t i = 0; i < ns && so + 64 * (size_t)(i + 1) <= lb.size(); ++i) {
            const size_t o = so + 64 * (size_t)i;
            const float c[3] = {f32(o), f32(o + 4), f32(o + 8)};
            if (!(c[0] > 0.0f || c[1] > 0.0f || c[2] > 0.0f)) continue;
            double dx = f32(o + 32), dy = f32(o + 36), dz = f32(o + 40), l = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (!(l > 0.0)) continue;
            const float w[16] = {c[0], c[1], c[2], 2.0f, f32(o + 16), f32(o + 20), f32(o + 24), f32(o + 60),
                                 (float)(dx / l), (float)(dy / l), (float)(dz / l), 0.0f,
                                 f32(o + 56), f32(o + 52), f32(o + 48), 0.0f};
            push(w);
        }
    }
    // the block as the reference's shaders read it: all point lights (black ones
    // too: num_pointlights is their count) and the directional light unchanged
    std::vector<float> rp;
    std::fill(E.ref_dir, E.ref_dir + 8, 0.0f);
    if (lb.size() >= 112) {
        for (int k = 0; k < 3; ++k) {
            E.ref_dir[k] = f32(4 * (size_t)k);
            E.ref_dir[4 + k] = f32(16 + 4 * (size_t)k);
        }
        uint32_t np;
        std::memcpy(&np, lb.data() + 96, 4);
        for (uint32_t i = 0; i < np; ++i) {
            const size_t o = 112 + 48 * (size_t)i;
            const float w[12] = {f32(o), f32(o + 4), f32(o + 8), 0.0f, f32(o + 16), f32(o + 20), f32(o + 24), 0.0f,
                                 f32(o + 40), f32(o + 36), f32(o + 32), 0.0f};
            rp.insert(rp.end(), w, w + 12);
        }
    }
    E.n_ref_pl = (uint32_t)(rp.size() / 12);
    E.n_em_list = (uint32_t)list.size();
    E.n_analytic = (uint32_t)(al.size() / 16);
    E.q_tri = ne == 0 ? (E.n_analytic ? 0.0f : 1.0f) : (E.n_analytic ? 0.5f : 1.0f);
    snkpt_status s;
    if ((s = up(E.analytic, al.data(), al.size() * 4)) != SNKPT_OK) return s;
    if ((s = up(E.rec, rec.data(), rec.size() * 4)) != SNKPT_OK) return s;
    if ((s = up(E.gid_to_emitter, gid_to_e.data(), gid_to_e.size() * 4)) != SNKPT_OK) return s;
    if ((s = up(E.ref_pl, rp.data(), rp.size() * 4)) != SNKPT_OK) return s;
    if ((s = up(E.em_list, list.data(), list.size() * 4)) != SNKPT_OK) return s;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    E.count = ne;
    E.total_power = total;
    return SNKPT_OK;
}

}  // namespace snkpt
