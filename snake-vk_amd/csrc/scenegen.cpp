// Synthetic scene generators for the five benchmark configurations.
//
// There is no reference counterpart: the reference loads assets through
// assimp (AssetLoader.cpp:337-466).  The generators produce exactly the inputs
// the reference's pass consumes: a global SoA vertex/index buffer
// (MeshBufferManager.cpp:22-88), 20 B instance records
// (RaytracingBufferSystem.cpp:69-103), column-major mat4 transforms, 64 B
// materials (ResourceBufferManagers.cpp:64-101, emission = albedo when
// flags & 1) and a CommonUBO camera (SceneInfoBufferSystem.cpp:66-91).
//
// Randomness: PCG32 (O'Neill 2014) seeded with 0x534E4B00 + config, one stream
// per purpose, so every scene is bit-identical from run to run.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "snkpt.h"
#include "snkpt_host.h"

namespace snkpt {

namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t initstate, uint64_t initseq) {
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
    }
    double uniform() { return (next() >> 8) * (1.0 / 16777216.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct V3 {
    double x, y, z;
};
static V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 normalize(V3 a) {
    double l = std::sqrt(dot(a, a));
    return l > 0 ? a * (1.0 / l) : a;
}

// One mesh under construction, vertices in the mesh's local space.
struct Mesh {
    std::vector<V3> p, n;
    std::vector<double> uv;
    std::vector<uint32_t> idx;
    uint32_t add_vertex(V3 pos, V3 nrm, double u = 0, double v = 0) {
        p.push_back(pos);
        n.push_back(nrm);
        uv.push_back(u);
        uv.push_back(v);
        return (uint32_t)p.size() - 1;
    }
    void tri(uint32_t a, uint32_t b, uint32_t c) {
        idx.push_back(a);
        idx.push_back(b);
        idx.push_back(c);
    }
};

struct Builder {
    std::vector<float> pos, nrm, uv, tan;
    std::vector<uint32_t> idx;
    std::vector<snkpt_instance> inst;
    std::vector<float> mat4s;
    std::vector<snkpt_material> mats;
    std::vector<uint32_t> emissive;

    // Appends the mesh; returns (vertex offset, index offset).
    std::pair<uint32_t, uint32_t> add_mesh(const Mesh& m) {
        uint32_t vo = (uint32_t)(pos.size() / 3), io = (uint32_t)idx.size();
        for (size_t i = 0; i < m.p.size(); ++i) {
            pos.push_back((float)m.p[i].x);
            pos.push_back((float)m.p[i].y);
            pos.push_back((float)m.p[i].z);
            nrm.push_back((float)m.n[i].x);
            nrm.push_back((float)m.n[i].y);
            nrm.push_back((float)m.n[i].z);
            uv.push_back((float)m.uv[2 * i]);
            uv.push_back((float)m.uv[2 * i + 1]);
            // tangent: any unit vector orthogonal to n
            V3 nn = m.n[i];
            V3 t = std::fabs(nn.x) > 0.9 ? V3{0, 1, 0} : V3{1, 0, 0};
            t = normalize(cross(t, nn));
            tan.push_back((float)t.x);
            tan.push_back((float)t.y);
            tan.push_back((float)t.z);
        }
        for (uint32_t v : m.idx) idx.push_back(v);
        return {vo, io};
    }
    uint32_t add_transform(const double m[16]) {
        for (int i = 0; i < 16; ++i) mat4s.push_back((float)m[i]);
        return (uint32_t)(mat4s.size() / 16) - 1;
    }
    uint32_t add_material(double r, double g, double b, double rough, double metal, uint32_t flags) {
        snkpt_material m;
        std::memset(&m, 0, sizeof(m));
        m.albedo_tex_idx = m.normal_tex_idx = m.roughness_tex_idx = m.metallic_tex_idx = m.ao_tex_idx =
            SNKPT_INVALID_INDEX;
        m.roughness = (float)rough;
        m.metallic = (float)metal;
        m.ao = 1.0f;
        m.albedo[0] = (float)r;
        m.albedo[1] = (float)g;
        m.albedo[2] = (float)b;
        m.albedo[3] = 1.0f;
        m.flags = flags;
        mats.push_back(m);
        return (uint32_t)mats.size() - 1;
    }
    uint32_t add_instance(uint32_t transform, uint32_t index_offset, uint32_t vertex_offset, uint32_t material,
                          uint32_t num_indices) {
        snkpt_instance r{transform, index_offset, vertex_offset, material, num_indices};
        inst.push_back(r);
        uint32_t id = (uint32_t)inst.size() - 1;
        if (mats[material].flags & SNKPT_MAT_FLAG_EMISSIVE) emissive.push_back(id);
        return id;
    }
};

static void mat_identity(double m[16]) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
}
// column-major T(c) * S(s)
static void mat_translate_scale(double m[16], V3 c, V3 s) {
    mat_identity(m);
    m[0] = s.x;
    m[5] = s.y;
    m[10] = s.z;
    m[12] = c.x;
    m[13] = c.y;
    m[14] = c.z;
}

// Axis-aligned quad from corner o spanned by edges a and b, normal = normalize(a x b).
static Mesh quad(V3 o, V3 a, V3 b) {
    Mesh m;
    V3 n = normalize(cross(a, b));
    uint32_t v0 = m.add_vertex(o, n, 0, 0), v1 = m.add_vertex(o + a, n, 1, 0);
    uint32_t v2 = m.add_vertex(o + a + b, n, 1, 1), v3 = m.add_vertex(o + b, n, 0, 1);
    m.tri(v0, v1, v2);
    m.tri(v0, v2, v3);
    return m;
}

// Unit cube [-0.5, 0.5]^3, 24 vertices, 12 triangles, outward normals.
static Mesh unit_box() {
    Mesh m;
    const V3 faces[6][3] = {
        {{0.5, -0.5, -0.5}, {0, 0, 1}, {0, 1, 0}},   // +x
        {{-0.5, -0.5, 0.5}, {0, 0, -1}, {0, 1, 0}},  // -x
        {{-0.5, 0.5, -0.5}, {0, 0, 1}, {1, 0, 0}},   // +y
        {{-0.5, -0.5, -0.5}, {1, 0, 0}, {0, 0, 1}},  // -y
        {{-0.5, -0.5, 0.5}, {0, 1, 0}, {1, 0, 0}},   // +z
        {{-0.5, -0.5, -0.5}, {1, 0, 0}, {0, 1, 0}},  // -z
    };
    for (auto& f : faces) {
        Mesh q = quad(f[0], f[1], f[2]);
        uint32_t base = (uint32_t)m.p.size();
        for (size_t i = 0; i < q.p.size(); ++i) m.add_vertex(q.p[i], q.n[i], q.uv[2 * i], q.uv[2 * i + 1]);
        for (uint32_t v : q.idx) m.idx.push_back(base + v);
    }
    return m;
}

// Unit UV sphere with nu segments and nv rings; pole rings are triangle fans,
// so it has 2*nu*(nv-1) triangles (64x32 -> 3968).
static Mesh uv_sphere(uint32_t nu, uint32_t nv) {
    Mesh m;
    for (uint32_t j = 0; j <= nv; ++j) {
        double theta = M_PI * (double)j / nv;
        for (uint32_t i = 0; i <= nu; ++i) {
            double phi = 2.0 * M_PI * (double)i / nu;
            V3 p{std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
            m.add_vertex(p, p, (double)i / nu, (double)j / nv);
        }
    }
    auto at = [&](uint32_t j, uint32_t i) { return j * (nu + 1) + i; };
    for (uint32_t j = 0; j < nv; ++j)
        for (uint32_t i = 0; i < nu; ++i) {
            uint32_t a = at(j, i), b = at(j + 1, i), c = at(j + 1, i + 1), d = at(j, i + 1);
            if (j != 0) m.tri(a, c, d);
            if (j != nv - 1) m.tri(a, b, c);
        }
    return m;
}

// Unit icosphere subdivided `level` times: 20*4^level triangles.
static Mesh icosphere(int level) {
    Mesh m;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const V3 v[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) {
        V3 q = normalize(p);
        m.add_vertex(q, q);
    }
    std::vector<uint32_t> f = {0, 11, 5, 0, 5,  1,  0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                               4, 11, 10, 2, 10, 7,  6, 7, 1, 8, 3,  9,  4, 3,  4,  2, 3, 2, 6, 3,
                               6, 8,  3,  8, 9,  4,  9, 5, 2, 4, 11, 6,  2, 10, 8,  6, 7, 9, 8, 1};
    for (int l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V3 q = normalize((m.p[a] + m.p[b]) * 0.5);
            uint32_t id = m.add_vertex(q, q);
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> g;
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t nt[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            g.insert(g.end(), nt, nt + 12);
        }
        f.swap(g);
    }
    m.idx = f;
    return m;
}

static void recompute_normals(Mesh& m) {
    std::vector<V3> acc(m.p.size(), V3{0, 0, 0});
    for (size_t i = 0; i < m.idx.size(); i += 3) {
        V3 a = m.p[m.idx[i]], b = m.p[m.idx[i + 1]], c = m.p[m.idx[i + 2]];
        V3 fn = cross(b - a, c - a);
        for (int k = 0; k < 3; ++k) acc[m.idx[i + k]] = acc[m.idx[i + k]] + fn;
    }
    for (size_t i = 0; i < m.p.size(); ++i) m.n[i] = normalize(acc[i]);
}

static uint32_t config_seed(uint32_t config) { return 0x534E4B00u + config; }

// ---- cameras -----------------------------------------------------------------

// glm::lookAt (right-handed) and glm::perspective with GLM_FORCE_DEPTH_ZERO_TO_ONE,
// Y flipped for Vulkan (CameraComponent.h:15-19, SceneInfoBufferSystem.cpp:79).
static void make_camera(V3 eye, V3 target, double fov_deg, double aspect, uint32_t frame, snkpt_camera* cam) {
    std::memset(cam, 0, sizeof(*cam));
    V3 up{0, 1, 0};
    V3 f = normalize(target - eye);
    V3 s = normalize(cross(f, up));
    V3 u = cross(s, f);
    double view[16];
    mat_identity(view);
    view[0] = s.x;
    view[4] = s.y;
    view[8] = s.z;
    view[1] = u.x;
    view[5] = u.y;
    view[9] = u.z;
    view[2] = -f.x;
    view[6] = -f.y;
    view[10] = -f.z;
    view[12] = -dot(s, eye);
    view[13] = -dot(u, eye);
    view[14] = dot(f, eye);
    const double znear = 0.1, zfar = 2500.0;
    double fovy = fov_deg * 0.5 * M_PI / 180.0;
    double th = std::tan(fovy / 2.0);
    double proj[16] = {0};
    proj[0] = 1.0 / (aspect * th);
    proj[5] = -1.0 / th;  // Vulkan Y flip
    proj[10] = zfar / (znear - zfar);
    proj[11] = -1.0;
    proj[14] = -(zfar * znear) / (zfar - znear);
    double pv[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc += proj[k * 4 + r] * view[c * 4 + k];
            pv[c * 4 + r] = acc;
        }
    for (int i = 0; i < 16; ++i) {
        cam->view[i] = (float)view[i];
        cam->proj[i] = (float)proj[i];
        cam->proj_view[i] = (float)pv[i];
    }
    cam->cam_pos[0] = (float)eye.x;
    cam->cam_pos[1] = (float)eye.y;
    cam->cam_pos[2] = (float)eye.z;
    cam->cam_pos[3] = 1.0f;
    cam->cam_forward[0] = (float)f.x;
    cam->cam_forward[1] = (float)f.y;
    cam->cam_forward[2] = (float)f.z;
    cam->cam_forward[3] = 1.0f;
    cam->frame_idx = frame;
}

// ---- configurations ----------------------------------------------------------

// Config 1 (and 3, 4, 5): open-front box room [-1,1]x[0,2]x[-1,1], two boxes, a
// UV sphere and a 0.25 x 0.25 emissive ceiling quad of radiance 15.
static void build_room(Builder& B, uint32_t config, uint32_t sphere_nu, uint32_t sphere_nv,
                       uint32_t* sphere_transform, uint32_t* sphere_vo, uint32_t* sphere_io,
                       uint32_t* sphere_nidx) {
    Pcg32 mat_rng(config_seed(config), 1);
    double id[16];
    mat_identity(id);
    uint32_t t_id = B.add_transform(id);
    auto albedo_mat = [&]() {
        double r = mat_rng.uniform(0.2, 0.8), g = mat_rng.uniform(0.2, 0.8), b = mat_rng.uniform(0.2, 0.8);
        return B.add_material(r, g, b, 1.0, 0.0, SNKPT_MAT_FLAG_LAMBERT);
    };
    struct Q {
        V3 o, a, b;
    };
    const Q walls[5] = {
        {{-1, 0, -1}, {0, 0, 2}, {2, 0, 0}},  // floor, normal +y
        {{-1, 2, -1}, {2, 0, 0}, {0, 0, 2}},  // ceiling, normal -y
        {{-1, 0, -1}, {2, 0, 0}, {0, 2, 0}},  // back, normal +z
        {{-1, 0, -1}, {0, 2, 0}, {0, 0, 2}},  // left, normal +x
        {{1, 0, -1}, {0, 0, 2}, {0, 2, 0}},   // right, normal -x
    };
    for (auto& w : walls) {
        Mesh q = quad(w.o, w.a, w.b);
        auto off = B.add_mesh(q);
        B.add_instance(t_id, off.second, off.first, albedo_mat(), (uint32_t)q.idx.size());
    }
    Mesh box = unit_box();
    auto box_off = B.add_mesh(box);
    const V3 box_c[2] = {{-0.35, 0.6, -0.3}, {0.4, 0.3, 0.3}};
    const V3 box_s[2] = {{0.5, 1.2, 0.5}, {0.5, 0.6, 0.5}};
    for (int i = 0; i < 2; ++i) {
        double m[16];
        mat_translate_scale(m, box_c[i], box_s[i]);
        uint32_t t = B.add_transform(m);
        B.add_instance(t, box_off.second, box_off.first, albedo_mat(), (uint32_t)box.idx.size());
    }
    Mesh sph = uv_sphere(sphere_nu, sphere_nv);
    auto sph_off = B.add_mesh(sph);
    double m[16];
    mat_translate_scale(m, {0.4, 1.0, 0.3}, {0.4, 0.4, 0.4});
    uint32_t t_sph = B.add_transform(m);
    B.add_instance(t_sph, sph_off.second, sph_off.first, albedo_mat(), (uint32_t)sph.idx.size());
    if (sphere_transform) *sphere_transform = t_sph;
    if (sphere_vo) *sphere_vo = sph_off.first;
    if (sphere_io) *sphere_io = sph_off.second;
    if (sphere_nidx) *sphere_nidx = (uint32_t)sph.idx.size();
    // emissive ceiling quad, facing down, just under the ceiling
    Mesh light = quad({-0.125, 2.0 - 1e-4, -0.125}, {0.25, 0, 0}, {0, 0, 0.25});
    auto l_off = B.add_mesh(light);
    uint32_t lm = B.add_material(15.0, 15.0, 15.0, 1.0, 0.0, SNKPT_MAT_FLAG_EMISSIVE);
    B.add_instance(t_id, l_off.second, l_off.first, lm, (uint32_t)light.idx.size());
}

// Config 2: ground plane + 200 level-4 icospheres (~1.02M triangles), GGX
// materials, 16 emissive sphere faces.
static void build_spheres(Builder& B, uint32_t config) {
    Pcg32 place(config_seed(config), 2), mat_rng(config_seed(config), 1), em_rng(config_seed(config), 4);
    double id[16];
    mat_identity(id);
    uint32_t t_id = B.add_transform(id);
    auto ggx = [&]() {
        double rough = mat_rng.uniform(0.05, 1.0);
        double metal = mat_rng.uniform() < 0.3 ? 1.0 : 0.0;
        double r = mat_rng.uniform(0.1, 0.9), g = mat_rng.uniform(0.1, 0.9), b = mat_rng.uniform(0.1, 0.9);
        return B.add_material(r, g, b, rough, metal, 0);
    };
    Mesh ground = quad({-100, 0, -100}, {0, 0, 200}, {200, 0, 0});
    auto g_off = B.add_mesh(ground);
    B.add_instance(t_id, g_off.second, g_off.first, ggx(), 6);

    const int n_spheres = 200, n_emit = 16;
    Mesh base = icosphere(4);
    const uint32_t n_faces = (uint32_t)base.idx.size() / 3;  // 5120
    // choose the emissive faces first (distinct (sphere, face) pairs)
    std::vector<std::vector<uint32_t>> emit_faces(n_spheres);
    std::vector<double> emit_L;
    for (int e = 0; e < n_emit;) {
        uint32_t s = em_rng.next() % n_spheres, f = em_rng.next() % n_faces;
        bool dup = false;
        for (uint32_t x : emit_faces[s]) dup |= (x == f);
        if (dup) continue;
        emit_faces[s].push_back(f);
        ++e;
    }
    for (int s = 0; s < n_spheres; ++s) {
        Pcg32 disp(config_seed(config), 1000 + (uint64_t)s);
        V3 c{place.uniform(-50, 50), place.uniform(0, 10), place.uniform(-50, 50)};
        double r = place.uniform(0.5, 3.0);
        Mesh m = base;
        for (auto& p : m.p) p = p * (1.0 + 0.02 * disp.normal());
        recompute_normals(m);
        // move emissive faces to the end of the index list
        std::vector<uint32_t> keep, tail;
        std::vector<uint32_t>& ef = emit_faces[s];
        for (uint32_t f = 0; f < n_faces; ++f) {
            bool is_e = false;
            for (uint32_t x : ef) is_e |= (x == f);
            if (!is_e)
                keep.insert(keep.end(), m.idx.begin() + 3 * f, m.idx.begin() + 3 * f + 3);
        }
        for (uint32_t x : ef) tail.insert(tail.end(), m.idx.begin() + 3 * x, m.idx.begin() + 3 * x + 3);
        m.idx = keep;
        m.idx.insert(m.idx.end(), tail.begin(), tail.end());
        auto off = B.add_mesh(m);
        double tm[16];
        mat_translate_scale(tm, c, {r, r, r});
        uint32_t t = B.add_transform(tm);
        B.add_instance(t, off.second, off.first, ggx(), (uint32_t)keep.size());
        for (size_t k = 0; k < ef.size(); ++k) {
            double L = em_rng.uniform(5.0, 50.0);
            uint32_t em = B.add_material(L, L, L, 1.0, 0.0, SNKPT_MAT_FLAG_EMISSIVE);
            B.add_instance(t, off.second + (uint32_t)keep.size() + 3 * (uint32_t)k, off.first, em, 3);
        }
    }
}

// Config 3: config-1 room + 10,000 emissive triangles, each a random sphere
// face shrunk x0.5 about its centroid and lifted along its normal, grouped into
// 16 emissive submeshes with log-uniform radiance in [1, 100].
static void build_many_lights(Builder& B, uint32_t config) {
    uint32_t t_sph = 0, vo = 0, io = 0, nidx = 0;
    build_room(B, config, 64, 32, &t_sph, &vo, &io, &nidx);
    Pcg32 rng(config_seed(config), 5);
    const uint32_t n_tris = 10000, n_groups = 16;
    const uint32_t n_faces = nidx / 3;
    const double lift = 1e-4 / 0.4;  // local units: the sphere transform scales by 0.4
    for (uint32_t g = 0; g < n_groups; ++g) {
        uint32_t count = n_tris / n_groups + (g < n_tris % n_groups ? 1u : 0u);
        Mesh m;
        for (uint32_t k = 0; k < count; ++k) {
            uint32_t f = rng.next() % n_faces;
            V3 p[3];
            for (int i = 0; i < 3; ++i) {
                uint32_t v = B.idx[io + 3 * f + i] + vo;
                p[i] = {B.pos[3 * v], B.pos[3 * v + 1], B.pos[3 * v + 2]};
            }
            V3 c = (p[0] + p[1] + p[2]) * (1.0 / 3.0);
            V3 n = normalize(cross(p[1] - p[0], p[2] - p[0]));
            if (dot(n, c) < 0) n = n * -1.0;  // outward
            uint32_t a[3];
            for (int i = 0; i < 3; ++i) a[i] = m.add_vertex(c + (p[i] - c) * 0.5 + n * lift, n);
            m.tri(a[0], a[1], a[2]);
        }
        auto off = B.add_mesh(m);
        double L = std::exp(rng.uniform(0.0, std::log(100.0)));
        uint32_t em = B.add_material(L, L, L, 1.0, 0.0, SNKPT_MAT_FLAG_EMISSIVE);
        B.add_instance(t_sph, off.second, off.first, em, (uint32_t)m.idx.size());
    }
}

static uint32_t c5_sphere_rings(uint32_t variant) {
    Pcg32 rng(config_seed(5), 100 + (uint64_t)variant);
    double target = std::exp(rng.uniform(std::log(1e4), std::log(2e6)));
    // sphere triangles 2*nu*(nv-1) with nu = 2*nv
    double nv = std::sqrt(target / 4.0);
    uint32_t r = (uint32_t)std::lround(nv);
    return r < 4 ? 4 : r;
}

}  // namespace

snkpt_status scene_camera(uint32_t config, uint32_t variant, uint32_t frame, uint32_t w, uint32_t h,
                          snkpt_camera* out) {
    if (!out || w == 0 || h == 0) return SNKPT_ERR_INVALID_ARG;
    double aspect = (double)w / (double)h;
    switch (config) {
        case 1:
        case 4:
            make_camera({0, 1, 3.4}, {0, 1, 0}, 90.0, aspect, frame, out);
            return SNKPT_OK;
        case 2:
            make_camera({0, 15, 75}, {0, 3, 0}, 90.0, aspect, frame, out);
            return SNKPT_OK;
        case 3: {
            // 64-frame orbit arc of radius 3.4 at y = 1 around the room's open side
            // (+z): the angle swings +-0.6 rad about the config 1 view, with a small
            // seeded jitter, so every frame looks into the lit room (the walls'
            // outsides are unlit)
            Pcg32 rng(config_seed(3), 200 + (uint64_t)frame);
            double th = 0.6 * std::sin(2.0 * M_PI * (double)(frame % 64) / 64.0) + rng.uniform(-0.01, 0.01);
            make_camera({3.4 * std::sin(th), 1, 3.4 * std::cos(th)}, {0, 1, 0}, 90.0, aspect, frame, out);
            return SNKPT_OK;
        }
        case 5: {
            // views inside/around the room; view = frame
            Pcg32 rng(config_seed(5), 300 + 64ull * variant + frame);
            double th = rng.uniform(-0.6, 0.6);
            double d = rng.uniform(2.6, 3.6);
            make_camera({d * std::sin(th), rng.uniform(0.6, 1.4), d * std::cos(th)},
                        {rng.uniform(-0.2, 0.2), 1, 0}, 90.0, aspect, frame, out);
            return SNKPT_OK;
        }
        default:
            return SNKPT_ERR_INVALID_ARG;
    }
}

// Resolution of view `view` of config 5 (side log-uniform in [128, 1024]).
uint32_t c5_view_side(uint32_t view) {
    Pcg32 rng(config_seed(5), 400 + (uint64_t)view);
    double s = std::exp(rng.uniform(std::log(128.0), std::log(1024.0)));
    uint32_t side = (uint32_t)std::lround(s);
    return side < 16 ? 16 : side;
}

template <typename T>
static T* dup_array(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = (T*)std::malloc(v.size() * sizeof(T));
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

snkpt_status scene_generate(uint32_t config, uint32_t variant, snkpt_scene** out) {
    if (!out) return SNKPT_ERR_INVALID_ARG;
    *out = nullptr;
    Builder B;
    uint32_t w = 256, h = 256, depth = 4, rr = 0;
    switch (config) {
        case 1:
            if (variant) return SNKPT_ERR_INVALID_ARG;
            build_room(B, 1, 64, 32, nullptr, nullptr, nullptr, nullptr);
            break;
        case 2:
            if (variant) return SNKPT_ERR_INVALID_ARG;
            build_spheres(B, 2);
            w = 1920, h = 1080, depth = 8, rr = 3;
            break;
        case 3:
            if (variant) return SNKPT_ERR_INVALID_ARG;
            build_many_lights(B, 3);
            w = 1920, h = 1080, depth = 1, rr = 0;
            break;
        case 4:
            if (variant) return SNKPT_ERR_INVALID_ARG;
            build_room(B, 4, 64, 32, nullptr, nullptr, nullptr, nullptr);
            w = 1920, h = 1080, depth = 8, rr = 3;
            break;
        case 5: {
            if (variant >= 16) return SNKPT_ERR_INVALID_ARG;
            uint32_t nv = c5_sphere_rings(variant);
            build_room(B, 5, 2 * nv, nv, nullptr, nullptr, nullptr, nullptr);
            w = h = c5_view_side(4 * variant);
            depth = 6, rr = 3;
            break;
        }
        default:
            return SNKPT_ERR_INVALID_ARG;
    }
    snkpt_scene* s = (snkpt_scene*)std::calloc(1, sizeof(snkpt_scene));
    s->pos = dup_array(B.pos);
    s->nrm = dup_array(B.nrm);
    s->uv = dup_array(B.uv);
    s->tan = dup_array(B.tan);
    s->n_verts = (uint32_t)(B.pos.size() / 3);
    s->idx = dup_array(B.idx);
    s->n_idx = (uint32_t)B.idx.size();
    s->inst = dup_array(B.inst);
    s->n_inst = (uint32_t)B.inst.size();
    s->mat4s = dup_array(B.mat4s);
    s->n_mat4 = (uint32_t)(B.mat4s.size() / 16);
    s->mats = dup_array(B.mats);
    s->n_mats = (uint32_t)B.mats.size();
    s->emissive_instances = dup_array(B.emissive);
    s->n_emissive = (uint32_t)B.emissive.size();
    s->width = w;
    s->height = h;
    s->max_depth = depth;
    s->rr_start = rr;
    s->seed = config_seed(config);
    scene_camera(config, variant, 0, w, h, &s->camera);
    *out = s;
    return SNKPT_OK;
}

void scene_free(snkpt_scene* s) {
    if (!s) return;
    std::free(s->pos);
    std::free(s->nrm);
    std::free(s->uv);
    std::free(s->tan);
    std::free(s->idx);
    std::free(s->inst);
    std::free(s->mat4s);
    std::free(s->mats);
    std::free(s->emissive_instances);
    std::free(s);
}

}  // namespace snkpt
