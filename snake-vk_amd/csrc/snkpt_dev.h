// Device-side layouts and inline routines shared by the snkpt kernels.
//
// BVH8 node (80 B, five 16 B words), after Ylitie, Karras & Laine 2017
// ("Efficient incoherent ray traversal on GPUs through compressed wide BVHs"):
//   w0 = p.x, p.y, p.z, (e.x | e.y << 8 | e.z << 16 | imask << 24)
//   w1 = child_base, tri_base, meta[0..3], meta[4..7]
//   w2 = qlo_x[0..3], qlo_x[4..7], qlo_y[0..3], qlo_y[4..7]
//   w3 = qlo_z[0..3], qlo_z[4..7], qhi_x[0..3], qhi_x[4..7]
//   w4 = qhi_y[0..3], qhi_y[4..7], qhi_z[0..3], qhi_z[4..7]
// Child box of slot s on axis a: [p_a + qlo_a[s] * 2^(e_a-127), p_a + qhi_a[s] * 2^(e_a-127)].
// meta[s] == 0: empty slot.  Internal child (imask bit s): meta = 0x20 | (24 + s);
// its node index is child_base + popcount(imask & ((1 << s) - 1)).  Leaf child:
// meta = (unary triangle count 1/3/7) << 5 | offset, its triangles are leaf
// records tri_base + offset ... + count - 1.
// Slot s holds the child lying towards octant s (bit a set = positive side of
// axis a), so a ray visits slots in order k = 0, 1, ... with s = k ^ octant(ray).
//
// Leaf triangle record (48 B, three 16 B words):
//   r0 = v0.x, v0.y, v0.z, v1.x   r1 = v1.y, v1.z, v2.x, v2.y   r2 = v2.z, inst, prim, gid
// gid is the global triangle id (instances in record order, triangles in
// submesh order); equal-t hits are resolved to the smaller gid so the closest
// hit does not depend on traversal order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace snkpt {

constexpr int kNodeWords = 5;  // uint4 per BVH8 node
constexpr int kTriWords = 3;   // float4 per leaf triangle
// 1 + 2^-16: box exits (and the current closest t they are clipped to) are
// widened by this factor so rounding never culls a box holding a closer hit.
constexpr float kBoxRobust = 1.0000152587890625f;

struct Ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};

// Per-ray constants of the watertight test (Woop, Benthin & Wald 2013) and of
// the box test.
struct RayPre {
    float ox, oy, oz;
    float ix, iy, iz;  // 1/d, NaN on a parallel axis
    float Sx, Sy, Sz;
    // kx | ky << 2 | kz << 4 | oct << 6 | par << 9 in one register: the traversal
    // loop keeps all of RayPre live, and as five registers they made the 64-VGPR
    // closest-hit loop spill one register to scratch on every step
    uint32_t bits;
    __device__ __forceinline__ int kx() const { return (int)(bits & 3u); }
    __device__ __forceinline__ int ky() const { return (int)((bits >> 2) & 3u); }
    __device__ __forceinline__ int kz() const { return (int)((bits >> 4) & 3u); }
    // bit a set: the ray runs towards -a
    __device__ __forceinline__ uint32_t oct() const { return (bits >> 6) & 7u; }
    // bit a set: axis a is parallel
    __device__ __forceinline__ uint32_t par() const { return (bits >> 9) & 7u; }
};

__device__ __forceinline__ float sel3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }
__device__ __forceinline__ float ubyte_f(uint32_t w, int b) { return (float)((w >> (8 * b)) & 0xFFu); }

// 0xFF in byte i for every set bit i of a 4-bit mask.
__device__ __forceinline__ uint32_t byte_mask4(uint32_t m4) {
    return ((m4 & 1u) | (m4 & 2u) << 7 | (m4 & 4u) << 14 | (m4 & 8u) << 21) * 0xFFu;
}

// 1/d; an axis with |d| < 1e-20 is "parallel" and gets NaN: every fused slab
// term of that axis is then NaN, which fminf/fmaxf drop, and the axis is
// instead tested exactly (origin inside the child's decoded [lo, hi]).  A huge
// or infinite inverse would make the fused terms cancel or turn into NaN in
// the wrong places (a vertical ray would then visit most of the tree).
constexpr float kParallelEps = 1e-20f;
__device__ __forceinline__ float safe_inv(float d) {
    return fabsf(d) < kParallelEps ? __builtin_nanf("") : 1.0f / d;
}

__device__ __forceinline__ RayPre ray_setup(const Ray& r) {
    RayPre p;
    p.ox = r.ox;
    p.oy = r.oy;
    p.oz = r.oz;
    p.ix = safe_inv(r.dx);
    p.iy = safe_inv(r.dy);
    p.iz = safe_inv(r.dz);
    float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = sel3(r.dx, r.dy, r.dz, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p.Sx = sel3(r.dx, r.dy, r.dz, kx) / dkz;
    p.Sy = sel3(r.dx, r.dy, r.dz, ky) / dkz;
    p.Sz = 1.0f / dkz;
    const uint32_t oct = (p.ix < 0.0f ? 1u : 0u) | (p.iy < 0.0f ? 2u : 0u) | (p.iz < 0.0f ? 4u : 0u);
    const uint32_t par = (fabsf(r.dx) < kParallelEps ? 1u : 0u) | (fabsf(r.dy) < kParallelEps ? 2u : 0u) |
                         (fabsf(r.dz) < kParallelEps ? 4u : 0u);
    p.bits = (uint32_t)kx | (uint32_t)ky << 2 | (uint32_t)kz << 4 | oct << 6 | par << 9;
    return p;
}

// Watertight ray/triangle test, both faces, FP32 only.  On a hit writes
// t = T/det and returns true; U/V/W/det are kept for the barycentrics.
// The operation order is fixed (the library is built without FP contraction)
// so the CPU restatement reproduces every bit.
__device__ __forceinline__ bool intersect_tri(const RayPre& p, float v0x, float v0y, float v0z, float v1x, float v1y,
                                              float v1z, float v2x, float v2y, float v2z, float& t, float& U,
                                              float& V, float& W, float& det) {
    float Ax_ = v0x - p.ox, Ay_ = v0y - p.oy, Az_ = v0z - p.oz;
    float Bx_ = v1x - p.ox, By_ = v1y - p.oy, Bz_ = v1z - p.oz;
    float Cx_ = v2x - p.ox, Cy_ = v2y - p.oy, Cz_ = v2z - p.oz;
    float Akz = sel3(Ax_, Ay_, Az_, p.kz()), Bkz = sel3(Bx_, By_, Bz_, p.kz()), Ckz = sel3(Cx_, Cy_, Cz_, p.kz());
    float Ax = sel3(Ax_, Ay_, Az_, p.kx()) - p.Sx * Akz;
    float Ay = sel3(Ax_, Ay_, Az_, p.ky()) - p.Sy * Akz;
    float Bx = sel3(Bx_, By_, Bz_, p.kx()) - p.Sx * Bkz;
    float By = sel3(Bx_, By_, Bz_, p.ky()) - p.Sy * Bkz;
    float Cx = sel3(Cx_, Cy_, Cz_, p.kx()) - p.Sx * Ckz;
    float Cy = sel3(Cx_, Cy_, Cz_, p.ky()) - p.Sy * Ckz;
    U = Cx * By - Cy * Bx;
    V = Ax * Cy - Ay * Cx;
    W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    det = U + V + W;
    if (det == 0.0f) return false;
    float Az = p.Sz * Akz, Bz = p.Sz * Bkz, Cz = p.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    return true;
}

struct TraceResult {
    float t;
    uint32_t gid;
    uint32_t rec;  // leaf record of the closest hit, ~0u if none
};

// Children (bit s) whose decoded box does not contain the ray origin on some
// parallel axis.  fmaf rounds once and monotonically, so the decoded planes
// still enclose the child.
__device__ __forceinline__ uint32_t parallel_miss_axis(float o, float pa, uint32_t ebyte, uint32_t lo0, uint32_t lo1,
                                                       uint32_t hi0, uint32_t hi1) {
    float sc = __uint_as_float(ebyte << 23);
    uint32_t miss = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        float lo = fmaf(ubyte_f(s < 4 ? lo0 : lo1, s & 3), sc, pa);
        float hi = fmaf(ubyte_f(s < 4 ? hi0 : hi1, s & 3), sc, pa);
        if (!(lo <= o && o <= hi)) miss |= 1u << s;
    }
    return miss;
}
__device__ __forceinline__ uint32_t parallel_miss(const RayPre& p, uint4 w0, uint4 w2, uint4 w3, uint4 w4) {
    uint32_t miss = 0;
    if (p.par() & 1u)
        miss |= parallel_miss_axis(p.ox, __uint_as_float(w0.x), w0.w & 0xFFu, w2.x, w2.y, w3.z, w3.w);
    if (p.par() & 2u)
        miss |= parallel_miss_axis(p.oy, __uint_as_float(w0.y), (w0.w >> 8) & 0xFFu, w2.z, w2.w, w4.x, w4.y);
    if (p.par() & 4u)
        miss |= parallel_miss_axis(p.oz, __uint_as_float(w0.z), (w0.w >> 16) & 0xFFu, w3.x, w3.y, w4.z, w4.w);
    return miss;
}

// BVH8 traversal state of one ray.  Stack entries are node groups:
// (child_base, hit bits << 24 | imask); the stack never holds more than depth-1
// entries, so STACK >= depth-1 is checked by the launcher.
struct TravState {
    RayPre p;
    float tmin;
    uint32_t ng_base, ng_hits;
    uint32_t tg_base, tg_hits;  // leaf triangles of the last node still to test
    int sp;
    TraceResult res;
};

// Sets up the traversal of `ray`; false when the ray is invalid (it misses):
// non-finite rays, a zero direction or an empty [tmin, tmax] miss at once, as a
// NaN would otherwise pass every box test and walk the whole tree.
__device__ __forceinline__ bool trav_init(const Ray& ray, TravState& s) {
    s.res = TraceResult{ray.tmax, 0xFFFFFFFFu, 0xFFFFFFFFu};
    const bool valid = __builtin_isfinite(ray.ox) && __builtin_isfinite(ray.oy) && __builtin_isfinite(ray.oz) &&
                       __builtin_isfinite(ray.dx) && __builtin_isfinite(ray.dy) && __builtin_isfinite(ray.dz) &&
                       (ray.dx != 0.0f || ray.dy != 0.0f || ray.dz != 0.0f) && ray.tmin <= ray.tmax &&
                       !__builtin_isnan(ray.tmin) && !__builtin_isnan(ray.tmax);
    if (!valid) return false;
    s.p = ray_setup(ray);
    s.tmin = ray.tmin;
    s.sp = 0;
    // start with a node group holding the root in slot 0
    s.ng_base = 0;
    s.ng_hits = (1u << (24 + (0u ^ s.p.oct()))) | 1u;
    s.tg_base = 0;
    s.tg_hits = 0;
    return true;
}

// Tests the next pending leaf triangle of the ray (st.tg_hits must be non-zero);
// true when ANY and it is a hit.
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_tri(const float4* __restrict__ tris, TravState& st, uint32_t& tvis) {
    const uint32_t j = __builtin_ctz(st.tg_hits);
    st.tg_hits &= st.tg_hits - 1u;
    const uint32_t rec = st.tg_base + j;
    const float4* tp = tris + (size_t)rec * kTriWords;
    float4 r0 = tp[0], r1 = tp[1], r2 = tp[2];
    if (COUNT) ++tvis;
    float t, U, V, W, det;
    if (intersect_tri(st.p, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, t, U, V, W, det)) {
        const uint32_t gid = __float_as_uint(r2.w);
        if (t >= st.tmin && (t < st.res.t || (t == st.res.t && gid < st.res.gid))) {
            st.res.t = t;
            st.res.gid = gid;
            st.res.rec = rec;
            if (ANY) return true;
        }
    }
    return false;
}

// Leaf triangles a lane may test in one traversal step before it goes on to the
// next node.  Every triangle of a node is still tested before the next node, so
// a ray's sequence of tests (and its culling) does not depend on this; it only
// bounds how long the other lanes of the wave wait for a lane with a full leaf.
// 1.29M incoherent rays, Grays/s: with leaves of up to three triangles (SAH
// triangle cost 0.3) all of a node's triangles per step 2.82, one 3.09, two
// 3.22, three 3.19, four 3.09; with leaves of mostly one triangle (cost 2.0)
// two 3.57, one 3.72.
constexpr int kTrisPerStep = 1;

// One traversal step: up to kTrisPerStep pending leaf triangles, then, when
// none is left, the next node of the current group (its 8 child boxes, which
// sets the pending triangles of its hit leaves); pops the stack when the group
// is empty.  Returns true when the ray is finished.  `stk` points to this lane's
// first stack entry in LDS; consecutive entries are `stride` uint2 apart.
template <bool ANY, bool COUNT, int STACK>
__device__ __forceinline__ bool trav_step(const uint4* __restrict__ nodes, const float4* __restrict__ tris,
                                          TravState& st, uint2* stk, int stride, uint32_t& nvis, uint32_t& tvis) {
    const RayPre& p = st.p;
    const float tmin = st.tmin;
#pragma unroll
    for (int i = 0; i < kTrisPerStep; ++i)
        if (st.tg_hits && trav_tri<ANY, COUNT>(tris, st, tvis)) return true;
    if (!st.tg_hits && (st.ng_hits & 0xFF000000u)) {
        const bool nx = p.ix < 0.0f, ny = p.iy < 0.0f, nz = p.iz < 0.0f;
        uint32_t k = __builtin_ctz(st.ng_hits >> 24);
        st.ng_hits &= ~(1u << (24 + k));
        uint32_t s = k ^ p.oct();
        uint32_t node = st.ng_base + __builtin_popcount(st.ng_hits & 0xFFu & ((1u << s) - 1u));
        if (st.ng_hits & 0xFF000000u) {
            stk[st.sp * stride] = make_uint2(st.ng_base, st.ng_hits);
            ++st.sp;
        }
        const uint4* np = nodes + (size_t)node * kNodeWords;
        uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
        if (COUNT) ++nvis;
        float sx = __uint_as_float((w0.w & 0xFFu) << 23) * p.ix;
        float sy = __uint_as_float(((w0.w >> 8) & 0xFFu) << 23) * p.iy;
        float sz = __uint_as_float(((w0.w >> 16) & 0xFFu) << 23) * p.iz;
        float ax = (__uint_as_float(w0.x) - p.ox) * p.ix;
        float ay = (__uint_as_float(w0.y) - p.oy) * p.iy;
        float az = (__uint_as_float(w0.z) - p.oz) * p.iz;
        uint32_t imask = w0.w >> 24;
        // near / far plane bytes per axis (qlo first for a positive direction)
        uint32_t nx0 = nx ? w3.z : w2.x, nx1 = nx ? w3.w : w2.y;
        uint32_t fx0 = nx ? w2.x : w3.z, fx1 = nx ? w2.y : w3.w;
        uint32_t ny0 = ny ? w4.x : w2.z, ny1 = ny ? w4.y : w2.w;
        uint32_t fy0 = ny ? w2.z : w4.x, fy1 = ny ? w2.w : w4.y;
        uint32_t nz0 = nz ? w4.z : w3.x, nz1 = nz ? w4.w : w3.y;
        uint32_t fz0 = nz ? w3.x : w4.z, fz1 = nz ? w3.y : w4.w;
        // meta bytes of slots 0-3 and 4-7; a parallel-axis miss clears the slot's
        // byte, so it reads as empty below
        uint32_t m0 = w1.z, m1 = w1.w;
        if (p.par()) {
            const uint32_t pmiss = parallel_miss(p, w0, w2, w3, w4);
            m0 &= ~byte_mask4(pmiss & 15u);
            m1 &= ~byte_mask4(pmiss >> 4);
        }
        // One hit word for both kinds of child, four slots at a time (Ylitie et al.
        // 2017, sec. 4): an internal child's meta byte is 0x20 | (24 + s), the only
        // bytes with bits 4 and 3 both set (leaf offsets stay below 24), and XOR-ing
        // its low three bits with the octant turns its bit index into 24 + (s ^ oct),
        // the visiting-order bit of the node group.  A leaf child keeps its index,
        // the offset of its first triangle, and its unary count.  An empty slot's
        // byte is 0 and sets nothing.
        const uint32_t oct4 = __builtin_amdgcn_perm(0u, p.oct(), 0u);  // octant in every byte
        const uint32_t in0 = (m0 & (m0 << 1)) & 0x10101010u, in1 = (m1 & (m1 << 1)) & 0x10101010u;
        // (in >> 1) - (in >> 4): 0x07 in the bytes of internal children, no borrow between bytes
        const uint32_t bi0 = (m0 ^ (oct4 & ((in0 >> 1) - (in0 >> 4)))) & 0x1F1F1F1Fu;
        const uint32_t bi1 = (m1 ^ (oct4 & ((in1 >> 1) - (in1 >> 4)))) & 0x1F1F1F1Fu;
        uint32_t hits = 0;
#pragma unroll
        for (int s2 = 0; s2 < 8; ++s2) {
            const int b = s2 & 3;
            uint32_t nxw = s2 < 4 ? nx0 : nx1, fxw = s2 < 4 ? fx0 : fx1;
            uint32_t nyw = s2 < 4 ? ny0 : ny1, fyw = s2 < 4 ? fy0 : fy1;
            uint32_t nzw = s2 < 4 ? nz0 : nz1, fzw = s2 < 4 ? fz0 : fz1;
            float tnx = fmaf(ubyte_f(nxw, b), sx, ax);
            float tny = fmaf(ubyte_f(nyw, b), sy, ay);
            float tnz = fmaf(ubyte_f(nzw, b), sz, az);
            float tfx = fmaf(ubyte_f(fxw, b), sx, ax);
            float tfy = fmaf(ubyte_f(fyw, b), sy, ay);
            float tfz = fmaf(ubyte_f(fzw, b), sz, az);
            float tn = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
            float tf = fminf(fminf(fminf(tfx, tfy), tfz), st.res.t) * kBoxRobust;
            const uint32_t bits = (((s2 < 4 ? m0 : m1) >> (8 * b + 5)) & 7u) << (((s2 < 4 ? bi0 : bi1) >> (8 * b)) & 31u);
            hits |= tn <= tf ? bits : 0u;
        }
        const uint32_t int_hits = hits & 0xFF000000u, tri_hits = hits & 0x00FFFFFFu;
        st.ng_base = w1.x;
        st.ng_hits = int_hits | imask;
        st.tg_base = w1.y;
        st.tg_hits = tri_hits;
    }
    if (!(st.ng_hits & 0xFF000000u)) {
        if (st.sp == 0) return st.tg_hits == 0u;
        --st.sp;
        uint2 e = stk[st.sp * stride];
        st.ng_base = e.x;
        st.ng_hits = e.y;
    }
    return false;
}

// BVH8 traversal of one ray to completion (closest hit, or any hit with ANY).
template <bool ANY, bool COUNT, int STACK>
__device__ __forceinline__ TraceResult traverse(const uint4* __restrict__ nodes, const float4* __restrict__ tris,
                                                const Ray& ray, uint2* stk, int stride, uint32_t& nvis,
                                                uint32_t& tvis) {
    TravState st;
    if (!trav_init(ray, st)) return st.res;
    while (!trav_step<ANY, COUNT, STACK>(nodes, tris, st, stk, stride, nvis, tvis)) {
    }
    return st.res;
}

// Persistent traversal with ray refill (after Aila & Laine 2009, "Understanding
// the efficiency of ray traversal on GPUs"), for the calling wave.  The rays
// [0, n) are cut into bands of nwaves x 64 rays.  In every whole band wave w
// owns the 64-ray chunk w (so the waves in flight trace one band of
// neighbouring rays at a time and share its BVH nodes in L2); the last, partial
// band is split evenly, wave w taking its [w rest / nwaves, (w+1) rest / nwaves)
// slice, so no wave gets a whole chunk more than another.  Its rays are handed
// out in order to idle lanes whenever at least REFILL lanes
// are idle (or all are), so a wave does not wait for its slowest ray with the
// other lanes masked off; between two such checks a lane takes up to
// kStepsPerRefill traversal steps.  load(id, ray, nodes, tris) fetches ray `id` and the
// BVH it is traced against; done(id, state, nvis, tvis) is called once per ray
// with its finished state, also for invalid rays (which miss).  Every ray's
// traversal is the one of traverse(), so results are identical.  All 64 lanes
// of the wave must call it; no atomics (a wave's cursor is uniform and private).
// Traversal steps between two refill checks: the check (a ballot and the ray
// hand-out) costs as much as a light step.  C2 incoherent rays, Grays/s: 1 step
// 3.21, 2: 3.33, 3: 3.37, 4: 3.35, 6: 3.31; the config-5 batch frame 15.6, 14.4,
// 13.9, 13.6, 13.7 ms.
constexpr int kStepsPerRefill = 4;

// Up to N traversal steps; true as soon as one finishes the ray.
template <bool ANY, bool COUNT, int STACK, int N>
__device__ __forceinline__ bool trav_steps(const uint4* __restrict__ nodes, const float4* __restrict__ tris,
                                           TravState& st, uint2* stk, int stride, uint32_t& nvis, uint32_t& tvis) {
    if (trav_step<ANY, COUNT, STACK>(nodes, tris, st, stk, stride, nvis, tvis)) return true;
    if constexpr (N > 1) return trav_steps<ANY, COUNT, STACK, N - 1>(nodes, tris, st, stk, stride, nvis, tvis);
    return false;
}

// refill_trace with the refill threshold and the steps per check chosen at run
// time / compile time (STEPS); refill_trace below fixes them.
template <bool ANY, bool COUNT, int STACK, int STEPS, class Load, class Done>
__device__ __forceinline__ void refill_trace_tuned(uint32_t n, uint32_t wave, uint32_t nwaves, uint2* stk, int stride,
                                                   uint32_t refill, Load load, Done done) {
    const uint32_t lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t band = nwaves * 64u;
    const uint32_t full = n / band;  // whole bands
    const uint32_t base = full * band, rest = n - base;
    const uint32_t lo = (uint32_t)(((uint64_t)wave * rest) / nwaves);
    // this wave's rays, counted in its own order: 64 per whole band, then its slice
    const uint32_t end = full * 64u + (uint32_t)(((uint64_t)(wave + 1u) * rest) / nwaves) - lo;
    uint32_t cursor = 0;
    bool active = false;
    uint32_t id = 0, nv = 0, tv = 0;
    const uint4* nodes = nullptr;
    const float4* tris = nullptr;
    TravState st;
    while (true) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle == 64u && cursor >= end) break;
        if (cursor < end && (nidle >= refill || nidle == 64u)) {
            if (!active) {
                const uint32_t k = cursor + (uint32_t)__popcll(idle & below);
                if (k < end) {
                    id = k < full * 64u ? (wave + (k >> 6) * nwaves) * 64u + (k & 63u) : base + lo + (k - full * 64u);
                    Ray ray;
                    load(id, ray, nodes, tris);
                    nv = tv = 0;
                    active = trav_init(ray, st);
                    if (!active) done(id, st, nv, tv);
                }
            }
            cursor += nidle;
        }
        if (active && trav_steps<ANY, COUNT, STACK, STEPS>(nodes, tris, st, stk, stride, nv, tv)) {
            active = false;
            done(id, st, nv, tv);
        }
    }
}

template <bool ANY, bool COUNT, int STACK, int REFILL, class Load, class Done>
__device__ __forceinline__ void refill_trace(uint32_t n, uint32_t wave, uint32_t nwaves, uint2* stk, int stride,
                                             Load load, Done done) {
    refill_trace_tuned<ANY, COUNT, STACK, kStepsPerRefill>(n, wave, nwaves, stk, stride, (uint32_t)REFILL, load, done);
}

// Barycentrics (u, v) of a leaf record hit; bary = (1-u-v, u, v).
__device__ __forceinline__ void hit_uv(const float4* __restrict__ tris, uint32_t rec, const Ray& ray, float& u,
                                       float& v) {
    RayPre p = ray_setup(ray);
    const float4* tp = tris + (size_t)rec * kTriWords;
    float4 r0 = tp[0], r1 = tp[1], r2 = tp[2];
    float t, U, V, W, det;
    intersect_tri(p, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, t, U, V, W, det);
    u = V / det;
    v = W / det;
}

// Block-level queue append: every thread of the block calls it (convergent:
// no thread may have returned before); threads with `want` get consecutive
// slots from one atomicAdd per block, in wave-then-lane order.  Same-address
// atomics are serialised by the memory system, so one per block instead of one
// per wave matters for the big early bounces.  `lds` holds BLOCK / 64 + 1 words
// and may be reused right after the call.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_append(bool want, uint32_t* counter, uint32_t* lds) {
    constexpr int kWaves = BLOCK / 64;
    const uint64_t m = __ballot(want);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (lane == 0) lds[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < kWaves; ++w) {
            uint32_t c = lds[w];
            lds[w] = run;
            run += c;
        }
        lds[kWaves] = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t slot = lds[kWaves] + lds[wave] + (uint32_t)__popcll(m & lt);
    __syncthreads();
    return want ? slot : 0xFFFFFFFFu;
}

}  // namespace snkpt
