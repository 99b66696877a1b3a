/*
 * snkpt oracle: PCG32, instance flattening, watertight intersection, LBVH ->
 * BVH8 build and BVH8 / brute-force tracing.  TEST INFRASTRUCTURE ONLY: the
 * product never links or calls this file.
 *
 * The reference builds no BVH of its own: BLAS/TLAS builds, traversal and the
 * ray/triangle test run inside the Vulkan driver (TLAS.cpp:74,178,
 * RayGen.rgen:106-118,164-176).  What is restated here is therefore the
 * algorithm this build specifies for that path, written independently of the
 * HIP code:
 *   - o_flatten: instance i, triangle p has vertices
 *     idx[mesh_buffer_index_offset + 3p + k] + mesh_buffer_vertex_offset
 *     (RayClosestHit.rchit:91, RaytracingInstances.glsl:7-13) transformed by
 *     transforms[transform_idx] (RayClosestHit.rchit:109-110).
 *   - o_intersect: Woop, Benthin & Wald 2013, both faces (no cull flags, TLAS.cpp:27,114,192).
 *   - o_build: top-down recursive radix-tree construction (split at the highest
 *     differing bit of (key, index)), numbered with Karras 2012's rule, then a
 *     post-order SAH pass and a BFS collapse to BVH8 nodes in the Ylitie et al.
 *     2017 layout.  The GPU builds the same tree bottom-up in parallel; the two
 *     are compared bit for bit.
 *   - o_trace_bvh8 replays the GPU traversal order (so node/triangle visit
 *     counts match); o_trace_brute tests every triangle.
 * Floating point: built with -ffp-contract=off; fmaf() only where the GPU
 * code uses an explicit fma.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "snkpt_oracle.h"

#define LEAF_TAG 0x80000000u
#define NONE 0xFFFFFFFFu
#define COST_NODE 1.0f
#define COST_TRI 2.0f /* a leaf triangle test costs two node visits (accel.hip kCostTri) */
#define MAX_LEAF_TRIS 3u
#define BOX_ROBUST 1.0000152587890625f

static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* ---------------------------------------------------------------- PCG32 --- */
void o_pcg32_seed(o_pcg32* r, uint64_t initstate, uint64_t initseq) {
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    o_pcg32_next(r);
    r->state += initstate;
    o_pcg32_next(r);
}

uint32_t o_pcg32_next(o_pcg32* r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

void o_pcg32_fill(uint64_t initstate, uint64_t initseq, uint32_t* out, uint32_t n) {
    o_pcg32 r;
    o_pcg32_seed(&r, initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = o_pcg32_next(&r);
}

/* -------------------------------------------------------------- flatten --- */
int o_flatten(const float* pos, uint32_t n_verts, const uint32_t* idx, uint32_t n_idx, const uint32_t* inst5,
              uint32_t n_inst, const float* mat4s, uint32_t n_mat4, float* out, uint32_t max_tris) {
    uint32_t g = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
        const uint32_t* in = inst5 + 5 * i;
        uint32_t xf = in[0], io = in[1], vo = in[2], ni = in[4];
        if (xf >= n_mat4 || io + ni > n_idx) return -1;
        const float* m = mat4s + 16 * (size_t)xf;
        for (uint32_t p = 0; p < ni / 3; ++p) {
            if (g >= max_tris) return -1;
            float* o = out + 12 * (size_t)g;
            for (int k = 0; k < 3; ++k) {
                uint32_t v = idx[io + 3 * p + k] + vo;
                if (v >= n_verts) return -1;
                float x = pos[3 * (size_t)v], y = pos[3 * (size_t)v + 1], z = pos[3 * (size_t)v + 2];
                /* column-major: x' = m0 x + m4 y + m8 z + m12, summed left to right */
                float wx = m[0] * x, wy = m[1] * x, wz = m[2] * x;
                wx = wx + m[4] * y;
                wy = wy + m[5] * y;
                wz = wz + m[6] * y;
                wx = wx + m[8] * z;
                wy = wy + m[9] * z;
                wz = wz + m[10] * z;
                o[3 * k + 0] = wx + m[12];
                o[3 * k + 1] = wy + m[13];
                o[3 * k + 2] = wz + m[14];
            }
            o[9] = u2f(i);
            o[10] = u2f(p);
            o[11] = u2f(g);
            ++g;
        }
    }
    return (int)g;
}

/* ------------------------------------------------------------ intersect --- */
typedef struct {
    float o[3], inv[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
    uint32_t oct, par;
} o_raypre;

/* 1/d; |d| < 1e-20 marks a parallel axis: its inverse is NaN (the fused slab
 * terms become NaN and fminf/fmaxf drop them) and the axis is tested exactly
 * instead: the origin must lie inside the child's decoded [lo, hi]. */
#define PARALLEL_EPS 1e-20f
static float safe_inv(float d) { return fabsf(d) < PARALLEL_EPS ? NAN : 1.0f / d; }

static void ray_pre(const float o[3], const float d[3], o_raypre* p) {
    for (int a = 0; a < 3; ++a) {
        p->o[a] = o[a];
        p->inv[a] = safe_inv(d[a]);
    }
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz;
    if (ax >= ay && ax >= az)
        kz = 0;
    else if (ay >= az)
        kz = 1;
    else
        kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (d[kz] < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p->kx = kx;
    p->ky = ky;
    p->kz = kz;
    p->Sx = d[kx] / d[kz];
    p->Sy = d[ky] / d[kz];
    p->Sz = 1.0f / d[kz];
    p->oct = (p->inv[0] < 0.0f ? 1u : 0u) | (p->inv[1] < 0.0f ? 2u : 0u) | (p->inv[2] < 0.0f ? 4u : 0u);
    p->par = 0;
    for (int a = 0; a < 3; ++a)
        if (fabsf(d[a]) < PARALLEL_EPS) p->par |= 1u << a;
}

/* returns 1 if the ray's line crosses the triangle (edge functions agree) */
static int tri_test(const o_raypre* p, const float* v0, const float* v1, const float* v2, float* t, float* U,
                    float* V, float* W, float* det) {
    float A[3], B[3], C[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = v0[a] - p->o[a];
        B[a] = v1[a] - p->o[a];
        C[a] = v2[a] - p->o[a];
    }
    float Ax = A[p->kx] - p->Sx * A[p->kz];
    float Ay = A[p->ky] - p->Sy * A[p->kz];
    float Bx = B[p->kx] - p->Sx * B[p->kz];
    float By = B[p->ky] - p->Sy * B[p->kz];
    float Cx = C[p->kx] - p->Sx * C[p->kz];
    float Cy = C[p->ky] - p->Sy * C[p->kz];
    float u = Cx * By - Cy * Bx;
    float v = Ax * Cy - Ay * Cx;
    float w = Bx * Ay - By * Ax;
    if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) return 0;
    float dt = u + v + w;
    if (dt == 0.0f) return 0;
    float Az = p->Sz * A[p->kz], Bz = p->Sz * B[p->kz], Cz = p->Sz * C[p->kz];
    float T = u * Az + v * Bz + w * Cz;
    *t = T / dt;
    *U = u;
    *V = v;
    *W = w;
    *det = dt;
    return 1;
}

int o_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                float* t, float* u, float* v) {
    o_raypre p;
    ray_pre(o, d, &p);
    float tt, U, V, W, det;
    if (!tri_test(&p, v0, v1, v2, &tt, &U, &V, &W, &det)) return 0;
    *t = tt;
    *u = V / det;
    *v = W / det;
    return 1;
}

/* ---------------------------------------------------------------- build --- */
static inline void rec_box(const float* r, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(fminf(r[a], r[3 + a]), r[6 + a]);
        hi[a] = fmaxf(fmaxf(r[a], r[3 + a]), r[6 + a]);
    }
}

static inline uint32_t spread10(uint32_t x) {
    /* bit i of x goes to bit 3i */
    uint32_t r = 0;
    for (int i = 0; i < 10; ++i) r |= ((x >> i) & 1u) << (3 * i);
    return r;
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static inline float area6(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

typedef struct {
    const uint64_t* ak; /* augmented keys (key << 32 | index) */
    uint32_t* child;
    uint32_t* first;
    uint32_t* count;
} karras_ctx;

/* largest split s in [f, l-1] with (ak[f] ^ ak[s]) sharing more leading bits than (ak[f] ^ ak[l]) */
static uint32_t find_split(const uint64_t* ak, uint32_t f, uint32_t l) {
    int common = clz64(ak[f] ^ ak[l]);
    uint32_t split = f, step = l - f;
    do {
        step = (step + 1) >> 1;
        uint32_t ns = split + step;
        if (ns < l && clz64(ak[f] ^ ak[ns]) > common) split = ns;
    } while (step > 1);
    return split;
}

/* Karras numbering: the internal node over [f, l] split at s has children
 * (f == s ? leaf s : internal s) and (l == s+1 ? leaf s+1 : internal s+1). */
static void build_range(karras_ctx* k, uint32_t node, uint32_t f, uint32_t l) {
    uint32_t s = find_split(k->ak, f, l);
    k->first[node] = f;
    k->count[node] = l - f + 1;
    uint32_t left = (f == s) ? (LEAF_TAG | s) : s;
    uint32_t right = (l == s + 1) ? (LEAF_TAG | (s + 1)) : (s + 1);
    k->child[2 * node] = left;
    k->child[2 * node + 1] = right;
    if (!(left & LEAF_TAG)) build_range(k, left, f, s);
    if (!(right & LEAF_TAG)) build_range(k, right, s + 1, l);
}

typedef struct {
    uint32_t n_int;
    const uint32_t* child;
    const uint32_t* count;
    float* aabb;
    float* cost;
    uint32_t* height;
    uint32_t* leafify;
    float* dpc;     /* C(x, i), i = 1..8, of the BVH8 collapse (8 per internal node) */
    uint32_t* dpk;  /* split choices, 3 bits per i */
} sah_ctx;

/* SAH cost of a binary subtree spread over i slots of a BVH8 node (Ylitie, Karras
   & Laine 2017, sec. 3.1), for the leaves the SAH pass fixed:
     leaf (a triangle or a leaf cluster of N):  C(x, i) = c_tri N A(x)
     internal x:  D(x, j) = min_{1 <= k < j} C(l, k) + C(r, j - k)
                  C(x, 1) = c_node A(x) + D(x, 8);  C(x, i) = min(C(x, i - 1), D(x, i))
   k ascending, strict comparisons (accel.hip collapse_costs makes the same choices);
   K: the k of D(x, 8) in bits 0-2, for i >= 2 the k of D(x, i) or 0 when C(x, i) = C(x, i - 1) */
static void collapse_costs(const float Cl[8], const float Cr[8], float area, float C[8], uint32_t* K) {
    float D[9];
    uint32_t Kd[9];
    for (int j = 2; j <= 8; ++j) {
        D[j] = INFINITY;
        Kd[j] = 1u;
        for (int k = 1; k < j; ++k) {
            float v = Cl[k - 1] + Cr[j - k - 1];
            if (v < D[j]) {
                D[j] = v;
                Kd[j] = (uint32_t)k;
            }
        }
    }
    C[0] = COST_NODE * area + D[8];
    *K = Kd[8];
    for (int i = 2; i <= 8; ++i) {
        if (D[i] < C[i - 2]) {
            C[i - 1] = D[i];
            *K |= Kd[i] << (3 * (i - 1));
        } else {
            C[i - 1] = C[i - 2];
        }
    }
}

static const float* node_box(const float* aabb, uint32_t n_int, uint32_t tagged) {
    uint32_t k = (tagged & LEAF_TAG) ? n_int + (tagged & ~LEAF_TAG) : tagged;
    return aabb + 6 * (size_t)k;
}

static void sah_post(sah_ctx* s, uint32_t node) {
    uint32_t c[2] = {s->child[2 * node], s->child[2 * node + 1]};
    float cc[2];
    uint32_t hh[2];
    for (int i = 0; i < 2; ++i) {
        if (c[i] & LEAF_TAG) {
            cc[i] = COST_TRI * area6(node_box(s->aabb, s->n_int, c[i]));
            hh[i] = 1;
        } else {
            sah_post(s, c[i]);
            cc[i] = s->cost[c[i]];
            hh[i] = s->height[c[i]];
        }
    }
    const float* bl = node_box(s->aabb, s->n_int, c[0]);
    const float* br = node_box(s->aabb, s->n_int, c[1]);
    float* b = s->aabb + 6 * (size_t)node;
    for (int a = 0; a < 3; ++a) {
        b[a] = fminf(bl[a], br[a]);
        b[3 + a] = fmaxf(bl[3 + a], br[3 + a]);
    }
    float area = area6(b);
    float c_split = COST_NODE * area + (cc[0] + cc[1]);
    uint32_t lf = 0;
    float cost = c_split;
    if (s->count[node] <= MAX_LEAF_TRIS) {
        float c_leaf = (COST_TRI * (float)s->count[node]) * area;
        if (c_leaf <= c_split) {
            lf = 1;
            cost = c_leaf;
        }
    }
    s->cost[node] = cost;
    s->leafify[node] = lf;
    s->height[node] = (hh[0] > hh[1] ? hh[0] : hh[1]) + 1;
    float Cl[8], Cr[8], C[8];
    for (int i = 0; i < 8; ++i) {
        Cl[i] = (c[0] & LEAF_TAG) ? cc[0] : s->dpc[8 * (size_t)c[0] + i];
        Cr[i] = (c[1] & LEAF_TAG) ? cc[1] : s->dpc[8 * (size_t)c[1] + i];
    }
    uint32_t K = 0;
    if (lf) {
        for (int i = 0; i < 8; ++i) C[i] = cost;
    } else {
        collapse_costs(Cl, Cr, area, C, &K);
    }
    for (int i = 0; i < 8; ++i) s->dpc[8 * (size_t)node + i] = C[i];
    s->dpk[node] = K;
}

static int quant_exp(float lo, float hi) {
    double ext = (double)hi - (double)lo;
    int e = -126;
    if (ext > 0.0) {
        int k;
        (void)frexp(ext / 255.0, &k);
        e = k;
        while (e > -126 && ldexp(255.0, e - 1) >= ext) --e;
        while (ldexp(255.0, e) < ext) ++e;
    }
    if (e < -126) e = -126;
    if (e > 127) e = 127;
    return e;
}

static uint32_t quant(float c, float p, int e, int up) {
    double x = ldexp((double)c - (double)p, -e);
    double q = up ? ceil(x) : floor(x);
    if (q < 0.0) q = 0.0;
    if (q > 255.0) q = 255.0;
    return (uint32_t)q;
}

int o_build(const float* world, uint32_t n, o_bvh* B) {
    memset(B, 0, sizeof(*B));
    B->n = n;
    if (n == 0) return 0;
    const uint32_t n_int = n - 1;
    B->keys = (uint32_t*)malloc(sizeof(uint32_t) * n);
    B->ids = (uint32_t*)malloc(sizeof(uint32_t) * n);
    B->child = (uint32_t*)calloc(2 * (size_t)(n_int ? n_int : 1), sizeof(uint32_t));
    B->first = (uint32_t*)calloc(n_int ? n_int : 1, sizeof(uint32_t));
    B->count = (uint32_t*)calloc(n_int ? n_int : 1, sizeof(uint32_t));
    B->aabb = (float*)calloc(6 * (size_t)(n_int + n), sizeof(float));
    B->leafify = (uint32_t*)calloc(n_int ? n_int : 1, sizeof(uint32_t));
    B->nodes = (uint32_t*)calloc(20 * (size_t)(n + 1), sizeof(uint32_t));
    B->leaf = (float*)calloc(12 * (size_t)n, sizeof(float));
    float* cost = (float*)calloc(n_int ? n_int : 1, sizeof(float));
    float* dpc = (float*)calloc(8 * (size_t)(n_int ? n_int : 1), sizeof(float));
    uint32_t* dpk = (uint32_t*)calloc(n_int ? n_int : 1, sizeof(uint32_t));
    uint32_t* height = (uint32_t*)calloc(n_int ? n_int : 1, sizeof(uint32_t));
    uint64_t* ak = (uint64_t*)malloc(sizeof(uint64_t) * n);

    /* centroid bounds */
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t g = 0; g < n; ++g) {
        float lo[3], hi[3];
        rec_box(world + 12 * (size_t)g, lo, hi);
        for (int a = 0; a < 3; ++a) {
            float c = (lo[a] + hi[a]) * 0.5f;
            clo[a] = fminf(clo[a], c);
            chi[a] = fmaxf(chi[a], c);
        }
    }
    /* Morton keys, sorted by (key, triangle id) */
    for (uint32_t g = 0; g < n; ++g) {
        float lo[3], hi[3];
        rec_box(world + 12 * (size_t)g, lo, hi);
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            float ext = chi[a] - clo[a];
            float inv = ext > 0.0f ? 1024.0f / ext : 0.0f;
            float c = (lo[a] + hi[a]) * 0.5f;
            float f = (c - clo[a]) * inv;
            if (f < 0.0f) f = 0.0f;
            if (f > 1023.0f) f = 1023.0f;
            q[a] = (uint32_t)f;
        }
        uint32_t key = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
        ak[g] = ((uint64_t)key << 32) | g;
    }
    qsort(ak, n, sizeof(uint64_t), cmp_u64);
    for (uint32_t i = 0; i < n; ++i) {
        B->keys[i] = (uint32_t)(ak[i] >> 32);
        B->ids[i] = (uint32_t)ak[i];
        ak[i] = ((uint64_t)B->keys[i] << 32) | i; /* augment with the sorted position */
    }
    /* radix tree */
    if (n >= 2) {
        karras_ctx k = {ak, B->child, B->first, B->count};
        build_range(&k, 0, 0, n - 1);
    }
    for (uint32_t j = 0; j < n; ++j) {
        float lo[3], hi[3];
        rec_box(world + 12 * (size_t)B->ids[j], lo, hi);
        float* o = B->aabb + 6 * (size_t)(n_int + j);
        for (int a = 0; a < 3; ++a) {
            o[a] = lo[a];
            o[3 + a] = hi[a];
        }
    }
    if (n >= 2) {
        sah_ctx s = {n_int, B->child, B->count, B->aabb, cost, height, B->leafify, dpc, dpk};
        sah_post(&s, 0);
        B->binary_depth = height[0];
    } else {
        B->binary_depth = 1;
    }
    /* BFS collapse */
    uint32_t* work = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
    work[0] = n >= 2 ? 0u : LEAF_TAG;
    uint32_t head = 0, tail = 1, tri_cursor = 0, level_end = 1;
    B->depth = 0;
#define EXPANDABLE(t) (!((t)&LEAF_TAG) && B->leafify[(t)] == 0u)
    while (head < tail) {
        uint32_t b = work[head];
        uint32_t S[8];
        int ns = 0;
        if (EXPANDABLE(b)) {
            /* the SAH collapse: the 8 slots split between b's children as D(b, 8)
               chose; then the first candidate the costs open (slot count i > 1 after
               dropping the i whose cost equals i - 1's) is replaced by its left child
               and its right child appended */
            uint32_t I[8], done = 0;
            uint32_t kb = dpk[b] & 7u;
            S[0] = B->child[2 * b];
            S[1] = B->child[2 * b + 1];
            I[0] = kb;
            I[1] = 8u - kb;
            ns = 2;
            for (;;) {
                int cur = -1;
                for (int k2 = 0; k2 < ns; ++k2)
                    if (!((done >> k2) & 1u)) {
                        cur = k2;
                        break;
                    }
                if (cur < 0) break;
                uint32_t x = S[cur], i = I[cur];
                if (!EXPANDABLE(x)) {
                    done |= 1u << cur;
                    continue;
                }
                uint32_t kx = dpk[x];
                while (i > 1u && ((kx >> (3u * (i - 1u))) & 7u) == 0u) --i;
                if (i == 1u) {
                    done |= 1u << cur;
                    continue;
                }
                uint32_t kk = (kx >> (3u * (i - 1u))) & 7u;
                S[cur] = B->child[2 * x];
                I[cur] = kk;
                S[ns] = B->child[2 * x + 1];
                I[ns] = i - kk;
                ++ns;
            }
        } else {
            S[0] = b;
            ns = 1;
        }
        const float* nb = node_box(B->aabb, n_int, b);
        float nc[3] = {(nb[0] + nb[3]) * 0.5f, (nb[1] + nb[4]) * 0.5f, (nb[2] + nb[5]) * 0.5f};
        float cst[8][8];
        for (int c = 0; c < ns; ++c) {
            const float* cb = node_box(B->aabb, n_int, S[c]);
            float dx = (cb[0] + cb[3]) * 0.5f - nc[0];
            float dy = (cb[1] + cb[4]) * 0.5f - nc[1];
            float dz = (cb[2] + cb[5]) * 0.5f - nc[2];
            for (int s = 0; s < 8; ++s) cst[c][s] = ((s & 1 ? dx : -dx) + (s & 2 ? dy : -dy)) + (s & 4 ? dz : -dz);
        }
        uint32_t slots[8];
        for (int s = 0; s < 8; ++s) slots[s] = NONE;
        uint32_t used_c = 0, used_s = 0;
        for (int r = 0; r < ns; ++r) {
            int bc = -1, bs = -1;
            float bvv = 0.0f;
            for (int c = 0; c < ns; ++c) {
                if (used_c & (1u << c)) continue;
                for (int s = 0; s < 8; ++s) {
                    if (used_s & (1u << s)) continue;
                    if (bc < 0 || cst[c][s] > bvv) {
                        bvv = cst[c][s];
                        bc = c;
                        bs = s;
                    }
                }
            }
            used_c |= 1u << bc;
            used_s |= 1u << bs;
            slots[bs] = S[bc];
        }
        /* emit node `head` */
        uint32_t* nd = B->nodes + 20 * (size_t)head;
        float p[3] = {nb[0], nb[1], nb[2]};
        int e[3];
        for (int a = 0; a < 3; ++a) e[a] = quant_exp(nb[a], nb[3 + a]);
        uint32_t child_base = tail, tri_base = tri_cursor, imask = 0, toff = 0;
        uint32_t meta[2] = {0, 0}, ql[3][2] = {{0}}, qh[3][2] = {{0}};
        for (int s = 0; s < 8; ++s) {
            uint32_t c = slots[s];
            if (c == NONE) continue;
            const float* cb = node_box(B->aabb, n_int, c);
            int w = s >> 2, sh = 8 * (s & 3);
            for (int a = 0; a < 3; ++a) {
                ql[a][w] |= quant(cb[a], p[a], e[a], 0) << sh;
                qh[a][w] |= quant(cb[3 + a], p[a], e[a], 1) << sh;
            }
            uint32_t m;
            if (EXPANDABLE(c)) {
                imask |= 1u << s;
                m = 0x20u | (24u + (uint32_t)s);
                work[tail++] = c;
            } else {
                uint32_t f = (c & LEAF_TAG) ? (c & ~LEAF_TAG) : B->first[c];
                uint32_t cnt = (c & LEAF_TAG) ? 1u : B->count[c];
                m = (((1u << cnt) - 1u) << 5) | toff;
                for (uint32_t t = 0; t < cnt; ++t)
                    memcpy(B->leaf + 12 * (size_t)(tri_base + toff + t), world + 12 * (size_t)B->ids[f + t],
                           48);
                toff += cnt;
            }
            meta[w] |= m << sh;
        }
        tri_cursor += toff;
        nd[0] = f2u(p[0]);
        nd[1] = f2u(p[1]);
        nd[2] = f2u(p[2]);
        nd[3] = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16) |
                (imask << 24);
        nd[4] = child_base;
        nd[5] = tri_base;
        nd[6] = meta[0];
        nd[7] = meta[1];
        nd[8] = ql[0][0];
        nd[9] = ql[0][1];
        nd[10] = ql[1][0];
        nd[11] = ql[1][1];
        nd[12] = ql[2][0];
        nd[13] = ql[2][1];
        nd[14] = qh[0][0];
        nd[15] = qh[0][1];
        nd[16] = qh[1][0];
        nd[17] = qh[1][1];
        nd[18] = qh[2][0];
        nd[19] = qh[2][1];
        ++head;
        if (head == level_end) {
            ++B->depth;
            level_end = tail;
        }
    }
#undef EXPANDABLE
    B->num_nodes = tail;
    free(work);
    free(ak);
    free(cost);
    free(dpc);
    free(dpk);
    free(height);
    return 0;
}

void o_bvh_free(o_bvh* b) {
    free(b->keys);
    free(b->ids);
    free(b->child);
    free(b->first);
    free(b->count);
    free(b->aabb);
    free(b->leafify);
    free(b->nodes);
    free(b->leaf);
    memset(b, 0, sizeof(*b));
}

/* ---------------------------------------------------------------- trace --- */
typedef struct {
    float t;
    uint32_t gid, rec;
} o_res;

/* finite origin and direction, non-zero direction, tmin <= tmax */
static int ray_valid(const float* r) {
    for (int k = 0; k < 8; ++k)
        if (k != 7 && !isfinite(r[k])) return 0;
    if (isnan(r[7])) return 0;
    if (r[4] == 0.0f && r[5] == 0.0f && r[6] == 0.0f) return 0;
    return r[3] <= r[7];
}

static o_res trace_one(const o_bvh* B, const float* ray, int any, uint32_t* nvis, uint32_t* tvis) {
    o_res res = {ray[7], NONE, NONE};
    if (!ray_valid(ray)) return res;
    o_raypre p;
    float o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
    ray_pre(o, d, &p);
    const float tmin = ray[3];
    const int nx = p.inv[0] < 0.0f, ny = p.inv[1] < 0.0f, nz = p.inv[2] < 0.0f;
    uint32_t stk[2 * 128];
    int sp = 0;
    uint32_t ng_base = 0, ng_hits = (1u << (24 + p.oct)) | 1u;
    for (;;) {
        uint32_t tg_base = 0, tg_hits = 0;
        if (ng_hits & 0xFF000000u) {
            uint32_t k = (uint32_t)__builtin_ctz(ng_hits >> 24);
            ng_hits &= ~(1u << (24 + k));
            uint32_t s = k ^ p.oct;
            uint32_t node = ng_base + (uint32_t)__builtin_popcount(ng_hits & 0xFFu & ((1u << s) - 1u));
            if (ng_hits & 0xFF000000u) {
                stk[2 * sp] = ng_base;
                stk[2 * sp + 1] = ng_hits;
                ++sp;
            }
            const uint32_t* w = B->nodes + 20 * (size_t)node;
            ++*nvis;
            float sx = u2f((w[3] & 0xFFu) << 23) * p.inv[0];
            float sy = u2f(((w[3] >> 8) & 0xFFu) << 23) * p.inv[1];
            float sz = u2f(((w[3] >> 16) & 0xFFu) << 23) * p.inv[2];
            float ax = (u2f(w[0]) - p.o[0]) * p.inv[0];
            float ay = (u2f(w[1]) - p.o[1]) * p.inv[1];
            float az = (u2f(w[2]) - p.o[2]) * p.inv[2];
            uint32_t imask = w[3] >> 24;
            /* qlo_x = w[8..9], qlo_y = w[10..11], qlo_z = w[12..13], qhi_x = w[14..15], qhi_y = 16..17, qhi_z = 18..19 */
            const uint32_t* nxw = nx ? w + 14 : w + 8;
            const uint32_t* fxw = nx ? w + 8 : w + 14;
            const uint32_t* nyw = ny ? w + 16 : w + 10;
            const uint32_t* fyw = ny ? w + 10 : w + 16;
            const uint32_t* nzw = nz ? w + 18 : w + 12;
            const uint32_t* fzw = nz ? w + 12 : w + 18;
            uint32_t pmiss = 0;
            for (int a = 0; a < 3 && p.par; ++a) {
                if (!((p.par >> a) & 1u)) continue;
                float sc = u2f(((w[3] >> (8 * a)) & 0xFFu) << 23), pa = u2f(w[a]);
                for (int s = 0; s < 8; ++s) {
                    float lo = fmaf((float)((w[8 + 2 * a + (s >> 2)] >> (8 * (s & 3))) & 0xFFu), sc, pa);
                    float hi = fmaf((float)((w[14 + 2 * a + (s >> 2)] >> (8 * (s & 3))) & 0xFFu), sc, pa);
                    if (!(lo <= p.o[a] && p.o[a] <= hi)) pmiss |= 1u << s;
                }
            }
            uint32_t int_hits = 0, tri_hits = 0;
            for (int s2 = 0; s2 < 8; ++s2) {
                int wi = s2 >> 2, sh = 8 * (s2 & 3);
                float tnx = fmaf((float)((nxw[wi] >> sh) & 0xFFu), sx, ax);
                float tny = fmaf((float)((nyw[wi] >> sh) & 0xFFu), sy, ay);
                float tnz = fmaf((float)((nzw[wi] >> sh) & 0xFFu), sz, az);
                float tfx = fmaf((float)((fxw[wi] >> sh) & 0xFFu), sx, ax);
                float tfy = fmaf((float)((fyw[wi] >> sh) & 0xFFu), sy, ay);
                float tfz = fmaf((float)((fzw[wi] >> sh) & 0xFFu), sz, az);
                float tn = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
                float tf = fminf(fminf(fminf(tfx, tfy), tfz), res.t) * BOX_ROBUST;
                uint32_t meta = (w[6 + wi] >> sh) & 0xFFu;
                if (meta != 0u && tn <= tf && !((pmiss >> s2) & 1u)) {
                    if ((imask >> s2) & 1u)
                        int_hits |= 1u << (24 + ((uint32_t)s2 ^ p.oct));
                    else
                        tri_hits |= (meta >> 5) << (meta & 31u);
                }
            }
            ng_base = w[4];
            ng_hits = int_hits | imask;
            tg_base = w[5];
            tg_hits = tri_hits;
        }
        while (tg_hits) {
            uint32_t j = (uint32_t)__builtin_ctz(tg_hits);
            tg_hits &= tg_hits - 1u;
            uint32_t rec = tg_base + j;
            const float* r = B->leaf + 12 * (size_t)rec;
            ++*tvis;
            float t, U, V, W, det;
            if (tri_test(&p, r, r + 3, r + 6, &t, &U, &V, &W, &det)) {
                uint32_t gid = f2u(r[11]);
                if (t >= tmin && (t < res.t || (t == res.t && gid < res.gid))) {
                    res.t = t;
                    res.gid = gid;
                    res.rec = rec;
                    if (any) return res;
                }
            }
        }
        if (!(ng_hits & 0xFF000000u)) {
            if (sp == 0) break;
            --sp;
            ng_base = stk[2 * sp];
            ng_hits = stk[2 * sp + 1];
        }
    }
    return res;
}

typedef struct {
    const o_bvh* B;
    const float* world;
    uint32_t ntris;
    const float* rays;
    uint32_t n, begin, end;
    int any;
    float *t, *u, *v;
    uint32_t *inst, *prim, *nvis, *tvis;
} o_job;

static void finish_hit(o_job* j, uint32_t i, const float* ray, const float* rec, float t) {
    if (rec) {
        j->t[i] = t;
        if (j->inst) j->inst[i] = f2u(rec[9]);
        if (j->prim) j->prim[i] = f2u(rec[10]);
        if (j->u || j->v) {
            o_raypre p;
            float o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
            ray_pre(o, d, &p);
            float tt, U, V, W, det;
            tri_test(&p, rec, rec + 3, rec + 6, &tt, &U, &V, &W, &det);
            if (j->u) j->u[i] = V / det;
            if (j->v) j->v[i] = W / det;
        }
    } else {
        j->t[i] = -1.0f;
        if (j->inst) j->inst[i] = NONE;
        if (j->prim) j->prim[i] = NONE;
        if (j->u) j->u[i] = 0.0f;
        if (j->v) j->v[i] = 0.0f;
    }
}

static void* bvh_worker(void* arg) {
    o_job* j = (o_job*)arg;
    for (uint32_t i = j->begin; i < j->end; ++i) {
        const float* ray = j->rays + 8 * (size_t)i;
        uint32_t nv = 0, tv = 0;
        o_res r = trace_one(j->B, ray, j->any, &nv, &tv);
        finish_hit(j, i, ray, r.rec != NONE ? j->B->leaf + 12 * (size_t)r.rec : NULL, r.t);
        if (j->nvis) j->nvis[i] = nv;
        if (j->tvis) j->tvis[i] = tv;
    }
    return NULL;
}

static void* brute_worker(void* arg) {
    o_job* j = (o_job*)arg;
    for (uint32_t i = j->begin; i < j->end; ++i) {
        const float* ray = j->rays + 8 * (size_t)i;
        o_raypre p;
        float o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
        ray_pre(o, d, &p);
        float best = ray[7];
        uint32_t best_gid = NONE;
        const float* best_rec = NULL;
        const uint32_t ntris = ray_valid(ray) ? j->ntris : 0;
        for (uint32_t g = 0; g < ntris; ++g) {
            const float* r = j->world + 12 * (size_t)g;
            float t, U, V, W, det;
            if (!tri_test(&p, r, r + 3, r + 6, &t, &U, &V, &W, &det)) continue;
            uint32_t gid = f2u(r[11]);
            if (t >= ray[3] && (t < best || (t == best && gid < best_gid))) {
                best = t;
                best_gid = gid;
                best_rec = r;
                if (j->any) break;
            }
        }
        finish_hit(j, i, ray, best_rec, best);
    }
    return NULL;
}

static void run_jobs(o_job* proto, int threads, void* (*fn)(void*)) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    o_job jobs[256];
    uint32_t n = proto->n, per = (n + (uint32_t)threads - 1) / (uint32_t)threads;
    int started = 0;
    for (int k = 0; k < threads; ++k) {
        jobs[k] = *proto;
        jobs[k].begin = (uint32_t)k * per < n ? (uint32_t)k * per : n;
        jobs[k].end = jobs[k].begin + per < n ? jobs[k].begin + per : n;
        if (threads == 1) {
            fn(&jobs[k]);
        } else if (pthread_create(&th[k], NULL, fn, &jobs[k]) == 0) {
            ++started;
        } else {
            fn(&jobs[k]);
            th[k] = 0;
        }
    }
    if (threads > 1)
        for (int k = 0; k < threads; ++k)
            if (th[k]) pthread_join(th[k], NULL);
    (void)started;
}

void o_trace_bvh8(const o_bvh* b, const float* rays, uint32_t n, int any_hit, float* t, uint32_t* inst,
                  uint32_t* prim, float* u, float* v, uint32_t* nvis, uint32_t* tvis, int threads) {
    if (b->n == 0) {
        o_job j = {0};
        j.t = t, j.inst = inst, j.prim = prim, j.u = u, j.v = v;
        for (uint32_t i = 0; i < n; ++i) {
            finish_hit(&j, i, rays + 8 * (size_t)i, NULL, 0.0f);
            if (nvis) nvis[i] = 0;
            if (tvis) tvis[i] = 0;
        }
        return;
    }
    o_job j = {b, NULL, 0, rays, n, 0, 0, any_hit, t, u, v, inst, prim, nvis, tvis};
    run_jobs(&j, threads, bvh_worker);
}

void o_trace_brute(const float* world, uint32_t ntris, const float* rays, uint32_t n, int any_hit, float* t,
                   uint32_t* inst, uint32_t* prim, float* u, float* v, int threads) {
    o_job j = {NULL, world, ntris, rays, n, 0, 0, any_hit, t, u, v, inst, prim, NULL, NULL};
    run_jobs(&j, threads, brute_worker);
}

void o_trace_one(const o_bvh* b, const float* ray, int any_hit, float* t, uint32_t* inst, uint32_t* prim, float* u,
                 float* v) {
    o_job j = {0};
    j.t = t, j.inst = inst, j.prim = prim, j.u = u, j.v = v;
    if (b->n == 0) {
        finish_hit(&j, 0, ray, NULL, 0.0f);
        return;
    }
    uint32_t nv = 0, tv = 0;
    o_res r = trace_one(b, ray, any_hit, &nv, &tv);
    finish_hit(&j, 0, ray, r.rec != NONE ? b->leaf + 12 * (size_t)r.rec : NULL, r.t);
}
