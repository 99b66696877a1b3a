// Temporal anti-aliasing resolve after the pass: TAA_ResolvePass::RecordCommandBuffer
// (Core/src/renderpasses/TAA_ResolvePass.cpp:39-56) dispatching TAA_Resolve.comp:37-86,
// then blitting the output to the history image and the velocity image to the
// previous-velocity image.
//
// One kernel does the resolve and both "blits": each 16x16 block stages its
// tile of the current frame with a one-texel halo in LDS (the 3x3
// neighbourhood statistics read LDS, not HBM), samples the history and the
// previous velocity bilinearly with repeat addressing (the reference's sampler,
// Images.cpp:14-33), and writes the output, the next history (in the output's
// format, as the blit copies the output image) and the next previous velocity
// into ping-pong buffers, so no texel that another block samples is
// overwritten during the launch.  Deviations, both deliberate: every pixel
// reads the frame as it was before the resolve (the shader reads neighbours
// from the image it is writing), and the grid covers the whole image (the
// reference dispatches size / 8 groups, truncated).
#include "snkpt_host.h"

namespace snkpt {

struct TaaState {
    uint32_t w = 0, h = 0, fmt = 0;
    int cur = 0;
    bool valid = false;
    DeviceBuffer hist[2], vel[2], scratch;
};
void taa_state_free(TaaState* t) { delete t; }

namespace {

constexpr int kTile = 16;
constexpr uint32_t kXcds = 8;  // MI355X: 8 accelerator dies, each with its own L2

__device__ __forceinline__ float h2f(uint16_t b) {
    _Float16 x;
    __builtin_memcpy(&x, &b, 2);
    return (float)x;
}
__device__ __forceinline__ uint16_t f2h(float f) {
    _Float16 x = (_Float16)f;
    uint16_t b;
    __builtin_memcpy(&b, &x, 2);
    return b;
}

template <int FMT>
__device__ __forceinline__ float4 load_px(const void* p, size_t i) {
    if (FMT == SNKPT_FORMAT_RGBA32F) return ((const float4*)p)[i];
    const uint2 b = ((const uint2*)p)[i];
    return make_float4(h2f((uint16_t)(b.x & 0xFFFFu)), h2f((uint16_t)(b.x >> 16)), h2f((uint16_t)(b.y & 0xFFFFu)),
                       h2f((uint16_t)(b.y >> 16)));
}
template <int FMT>
__device__ __forceinline__ void store_px(void* p, size_t i, float4 v) {
    if (FMT == SNKPT_FORMAT_RGBA32F) {
        ((float4*)p)[i] = v;
    } else {
        ((uint2*)p)[i] = make_uint2((uint32_t)f2h(v.x) | ((uint32_t)f2h(v.y) << 16),
                                    (uint32_t)f2h(v.z) | ((uint32_t)f2h(v.w) << 16));
    }
}
__device__ __forceinline__ float2 load_vel(const uint32_t* v, size_t i) {
    const uint32_t b = v[i];
    return make_float2(h2f((uint16_t)(b & 0xFFFFu)), h2f((uint16_t)(b >> 16)));
}

// Bilinear footprint of texcoord (u, v) with repeat addressing.
struct Tap {
    size_t i00, i10, i01, i11;
    float ax, ay;
};
__device__ __forceinline__ Tap bilinear(float u, float v, uint32_t w, uint32_t h) {
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    Tap t;
    t.ax = x - fx;
    t.ay = y - fy;
    int x0 = (int)fx % (int)w, y0 = (int)fy % (int)h;
    if (x0 < 0) x0 += (int)w;
    if (y0 < 0) y0 += (int)h;
    const int x1 = x0 + 1 == (int)w ? 0 : x0 + 1, y1 = y0 + 1 == (int)h ? 0 : y0 + 1;
    t.i00 = (size_t)y0 * w + x0;
    t.i10 = (size_t)y0 * w + x1;
    t.i01 = (size_t)y1 * w + x0;
    t.i11 = (size_t)y1 * w + x1;
    return t;
}
__device__ __forceinline__ float lerp2(float c00, float c10, float c01, float c11, float ax, float ay) {
    const float top = (1.0f - ax) * c00 + ax * c10;
    const float bot = (1.0f - ax) * c01 + ax * c11;
    return (1.0f - ay) * top + ay * bot;
}

// clip_aabb (TAA_Resolve.comp:11-35): pull q toward p until inside [lo, hi].
__device__ __forceinline__ void clip_aabb(const float lo[3], const float hi[3], const float p[3], float q[3]) {
    float r[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    const float eps = 0.00000001f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float rmax = hi[a] - p[a];
        if (r[a] > rmax + eps) {
            const float s = rmax / r[a];
            r[0] *= s;
            r[1] *= s;
            r[2] *= s;
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float rmin = lo[a] - p[a];
        if (r[a] < rmin - eps) {
            const float s = rmin / r[a];
            r[0] *= s;
            r[1] *= s;
            r[2] *= s;
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = p[a] + r[a];
}

template <int FMT>
__global__ __launch_bounds__(kTile* kTile) void k_taa_resolve(const void* __restrict__ cur, void* __restrict__ out,
                                                               const uint32_t* __restrict__ vel,
                                                               const void* __restrict__ hist,
                                                               const uint32_t* __restrict__ pvel,
                                                               void* __restrict__ hist_out,
                                                               uint32_t* __restrict__ pvel_out, uint32_t w,
                                                               uint32_t h, uint32_t reset) {
    __shared__ float tile[kTile + 2][kTile + 2][3];
    // The hardware deals consecutive blocks to the XCDs in turn; tile t of the
    // row-major tile order goes to block (t % per) * kXcds + t / per, so each XCD
    // walks a contiguous run of tiles and the halo columns and bilinear taps a
    // tile shares with its neighbours are still in that XCD's L2.
    const uint32_t nbx = (w + kTile - 1) / kTile, ntiles = nbx * ((h + kTile - 1) / kTile);
    const uint32_t per = (ntiles + kXcds - 1) / kXcds;
    const uint32_t tile_id = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (tile_id >= ntiles) return;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int bx = (int)(tile_id % nbx) * kTile, by = (int)(tile_id / nbx) * kTile;
    // tile + halo, edge-clamped texel coordinates, colours max(c, 0)
    for (int k = threadIdx.x; k < (kTile + 2) * (kTile + 2); k += kTile * kTile) {
        const int lx = k % (kTile + 2), ly = k / (kTile + 2);
        const int gx = min(max(bx + lx - 1, 0), (int)w - 1), gy = min(max(by + ly - 1, 0), (int)h - 1);
        const float4 c = load_px<FMT>(cur, (size_t)gy * w + gx);
        tile[ly][lx][0] = fmaxf(c.x, 0.0f);
        tile[ly][lx][1] = fmaxf(c.y, 0.0f);
        tile[ly][lx][2] = fmaxf(c.z, 0.0f);
    }
    __syncthreads();
    const uint32_t x = bx + tx, y = by + ty;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    const float4 c = load_px<FMT>(cur, i);
    const uint32_t vb = vel[i];
    pvel_out[i] = vb;
    if (reset) {
        store_px<FMT>(out, i, c);
        store_px<FMT>(hist_out, i, c);
        return;
    }
    const float2 v = load_vel(vel, i);
    const float hu = fminf(fmaxf(((float)x + 0.5f) / (float)w + v.x, 0.0f), 1.0f);
    const float hv = fminf(fmaxf(((float)y + 0.5f) / (float)h + v.y, 0.0f), 1.0f);
    float mn[3] = {9999.0f, 9999.0f, 9999.0f}, mx[3] = {-9999.0f, -9999.0f, -9999.0f};
    float m[3] = {0.0f, 0.0f, 0.0f}, m2[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float col = tile[ty + 1 + dy][tx + 1 + dx][a];
                mn[a] = fminf(mn[a], col);
                mx[a] = fmaxf(mx[a], col);
                m[a] = m[a] + col;
                m2[a] = m2[a] + col * col;
            }
    const float inv9 = 1.0f / 9.0f, gamma = 0.55f;
    float mu[3], lo[3], hi[3], pc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mu[a] = m[a] * inv9;
        const float sigma = sqrtf(fabsf(m2[a] * inv9 - mu[a] * mu[a]));
        lo[a] = mu[a] - gamma * sigma;
        hi[a] = mu[a] + gamma * sigma;
        pc[a] = fminf(fmaxf(mu[a], mn[a]), mx[a]);
    }
    const Tap t = bilinear(hu, hv, w, h);
    const float4 h00 = load_px<FMT>(hist, t.i00), h10 = load_px<FMT>(hist, t.i10);
    const float4 h01 = load_px<FMT>(hist, t.i01), h11 = load_px<FMT>(hist, t.i11);
    float hc[3] = {fmaxf(lerp2(h00.x, h10.x, h01.x, h11.x, t.ax, t.ay), 0.0f),
                   fmaxf(lerp2(h00.y, h10.y, h01.y, h11.y, t.ax, t.ay), 0.0f),
                   fmaxf(lerp2(h00.z, h10.z, h01.z, h11.z, t.ax, t.ay), 0.0f)};
    clip_aabb(lo, hi, pc, hc);
    const float2 p00 = load_vel(pvel, t.i00), p10 = load_vel(pvel, t.i10);
    const float2 p01 = load_vel(pvel, t.i01), p11 = load_vel(pvel, t.i11);
    const float pvx = lerp2(p00.x, p10.x, p01.x, p11.x, t.ax, t.ay) + v.x;
    const float pvy = lerp2(p00.y, p10.y, p01.y, p11.y, t.ax, t.ay) + v.y;
    const float vlen = sqrtf(pvx * pvx + pvy * pvy);
    const float dis = fminf(fmaxf((vlen - 0.001f) * 10.0f, 0.0f), 1.0f);
    const float cc[3] = {c.x, c.y, c.z};
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float b = cc[a] * (1.0f - 0.9f) + hc[a] * 0.9f;
        f[a] = b * (1.0f - dis) + mu[a] * dis;
    }
    const float4 o = make_float4(f[0], f[1], f[2], 1.0f);
    store_px<FMT>(out, i, o);
    store_px<FMT>(hist_out, i, o);
}

}  // namespace

#define HIPT(expr)                                                \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr);    \
    } while (0)

snkpt_status taa_resolve(snkpt_ctx* ctx, const void* current, void* out, const uint16_t* velocity, uint32_t w,
                         uint32_t h, uint32_t fmt, uint32_t reset, hipStream_t st) {
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    if (!ctx->taa) ctx->taa = new TaaState();
    TaaState& T = *ctx->taa;
    const size_t n = (size_t)w * h, px_bytes = fmt == SNKPT_FORMAT_RGBA32F ? 16 : 8;
    if (T.w != w || T.h != h || T.fmt != fmt) {
        for (int k = 0; k < 2; ++k) {
            HIPT(T.hist[k].reserve(n * px_bytes));
            HIPT(T.vel[k].reserve(n * 4));
        }
        T.w = w;
        T.h = h;
        T.fmt = fmt;
        T.valid = false;
    }
    // in place (out == current, as the reference resolves its output image):
    // read a copy so every pixel sees the frame before the resolve
    const void* src = current;
    if (current == out) {
        HIPT(T.scratch.reserve(n * px_bytes));
        HIPT(hipMemcpyAsync(T.scratch.ptr, current, n * px_bytes, hipMemcpyDeviceToDevice, st));
        src = T.scratch.ptr;
    }
    const uint32_t rs = (reset || !T.valid) ? 1u : 0u;
    const int nxt = T.cur ^ 1;
    const uint32_t ntiles = ((w + kTile - 1) / kTile) * ((h + kTile - 1) / kTile);
    const uint32_t grid = ((ntiles + kXcds - 1) / kXcds) * kXcds;
    if (fmt == SNKPT_FORMAT_RGBA32F)
        k_taa_resolve<SNKPT_FORMAT_RGBA32F><<<grid, kTile * kTile, 0, st>>>(
            src, out, (const uint32_t*)velocity, T.hist[T.cur].ptr, T.vel[T.cur].as<uint32_t>(), T.hist[nxt].ptr,
            T.vel[nxt].as<uint32_t>(), w, h, rs);
    else
        k_taa_resolve<SNKPT_FORMAT_RGBA16F><<<grid, kTile * kTile, 0, st>>>(
            src, out, (const uint32_t*)velocity, T.hist[T.cur].ptr, T.vel[T.cur].as<uint32_t>(), T.hist[nxt].ptr,
            T.vel[nxt].as<uint32_t>(), w, h, rs);
    HIPT(hipGetLastError());
    T.cur = nxt;
    T.valid = true;
    return SNKPT_OK;
}

}  // namespace snkpt
