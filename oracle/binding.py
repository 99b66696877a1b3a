"""numpy binding of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg;
never by the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SNKPT_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of `make -C oracle sanitize`)
LIB_PATH = os.environ.get("SNKPT_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    VP, U32, I = C.c_void_p, C.c_uint32, C.c_int
    lib.o_pcg32_fill.argtypes = [C.c_uint64, C.c_uint64, VP, U32]
    lib.o_flatten.argtypes = [VP, U32, VP, U32, VP, U32, VP, U32, VP, U32]
    lib.o_flatten.restype = I
    lib.o_intersect.argtypes = [VP, VP, VP, VP, VP, VP, VP, VP]
    lib.o_intersect.restype = I
    lib.o_build.argtypes = [VP, U32, C.POINTER(OBvh)]
    lib.o_build.restype = I
    lib.o_bvh_free.argtypes = [C.POINTER(OBvh)]
    lib.o_trace_bvh8.argtypes = [C.POINTER(OBvh), VP, U32, I, VP, VP, VP, VP, VP, VP, VP, I]
    lib.o_trace_brute.argtypes = [VP, U32, VP, U32, I, VP, VP, VP, VP, VP, I]
    lib.o_pass_state.argtypes = [U32, U32, U32]
    lib.o_pass_state.restype = C.c_uint64
    lib.o_camera_matrix.argtypes = [VP, VP, VP]
    F = C.c_float
    lib.o_brdf_eval.argtypes = [VP, F, F, U32, VP, VP, VP, VP, VP]
    lib.o_brdf_sample.argtypes = [VP, F, F, U32, VP, VP, F, F, F, VP]
    lib.o_brdf_sample.restype = I
    lib.o_pt_render.argtypes = [VP, U32, C.POINTER(OBvh), VP, U32, VP, VP, VP, VP, VP, U32, VP, VP, VP, U32, U32,
                                U32, U32, U32, U32, U32, U32, VP, I]
    lib.o_pt_render.restype = I
    lib.o_pt_render_gbuffer.argtypes = [VP, U32, C.POINTER(OBvh), VP, U32, VP, VP, VP, VP, VP, U32, VP, VP, VP, U32,
                                        U32, U32, U32, U32, U32, U32, C.POINTER(OGbuffer), VP, I]
    lib.o_pt_render_gbuffer.restype = I
    lib.o_compat_bounce.argtypes = [VP, U32, C.POINTER(OBvh), VP, U32, VP, VP, VP, VP, VP, U32, VP, VP, VP, U32, U32,
                                    U32, U32, U32, VP, I]
    lib.o_compat_bounce.restype = I
    lib.o_compat_direct.argtypes = [VP, U32, C.POINTER(OBvh), VP, U32, VP, VP, VP, VP, VP, U32, VP, C.c_size_t, VP,
                                    VP, VP, VP, U32, U32, U32, U32, U32, F, U32, VP, I]
    lib.o_compat_direct.restype = I
    lib.o_emitter_sample.argtypes = [VP, U32, VP, U32, VP, VP, U32, C.c_uint64, U32, U32, VP, VP, VP]
    lib.o_emitter_sample.restype = U32
    lib.o_half_to_float.argtypes = [C.c_uint16]
    lib.o_half_to_float.restype = F
    lib.o_primary_surfaces.argtypes = [VP, U32, C.POINTER(OBvh), VP, U32, VP, VP, VP, VP, VP, VP, VP, VP, U32, U32,
                                       VP, VP, VP, VP, VP]
    lib.o_cache_geom.argtypes = [VP, VP, VP, VP, VP]
    lib.o_grid_index.argtypes = [U32, U32, U32, U32, U32]
    lib.o_grid_index.restype = U32
    lib.o_cache_encode.argtypes = [VP, VP, VP, VP, VP, U32, VP]
    lib.o_cache_infer.argtypes = [VP, VP, VP, VP, VP, VP, U32, VP]
    lib.o_cache_grads.argtypes = [VP, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP]
    lib.o_cache_grads.restype = C.c_double
    lib.o_adam.argtypes = [VP, VP, VP, VP, C.c_size_t, U32, C.c_double, C.c_double, C.c_double, C.c_double]
    _lib = lib
    return lib


class OGbuffer(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("normal", C.c_void_p), ("albedo", C.c_void_p), ("rma", C.c_void_p),
                ("flags", C.c_void_p)]


class OBvh(C.Structure):
    _fields_ = [("n", C.c_uint32), ("num_nodes", C.c_uint32), ("depth", C.c_uint32),
                ("binary_depth", C.c_uint32), ("keys", C.POINTER(C.c_uint32)), ("ids", C.POINTER(C.c_uint32)),
                ("child", C.POINTER(C.c_uint32)), ("first", C.POINTER(C.c_uint32)),
                ("count", C.POINTER(C.c_uint32)), ("aabb", C.POINTER(C.c_float)),
                ("leafify", C.POINTER(C.c_uint32)), ("nodes", C.POINTER(C.c_uint32)),
                ("leaf", C.POINTER(C.c_float))]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def pcg32(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    load().o_pcg32_fill(initstate, initseq, _p(out), n)
    return out


def flatten(pos, idx, inst5, mat4s) -> np.ndarray:
    """World triangle records (n, 12) float32 view; columns 9..11 hold inst, prim, gid bits."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, np.uint32)
    inst5 = np.ascontiguousarray(inst5, np.uint32).reshape(-1)
    mat4s = np.ascontiguousarray(mat4s, np.float32).reshape(-1)
    ntri = int(sum(int(inst5[5 * i + 4]) // 3 for i in range(inst5.size // 5)))
    out = np.zeros((max(ntri, 1), 12), np.float32)
    r = load().o_flatten(_p(pos), pos.size // 3, _p(idx), idx.size, _p(inst5), inst5.size // 5, _p(mat4s),
                         mat4s.size // 16, _p(out), ntri)
    if r < 0:
        raise ValueError("o_flatten: invalid scene")
    return out[:r]


def intersect(o, d, v0, v1, v2):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (o, d, v0, v1, v2)]
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = load().o_intersect(*[_p(a) for a in arrs], C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


class Bvh:
    """CPU twin of the GPU build; arrays are numpy views into oracle memory."""

    def __init__(self, world: np.ndarray):
        self.world = np.ascontiguousarray(world, np.float32)
        self.b = OBvh()
        n = self.world.shape[0]
        if load().o_build(_p(self.world), n, C.byref(self.b)) != 0:
            raise RuntimeError("o_build failed")
        b = self.b
        as_u32 = lambda p, k: np.ctypeslib.as_array(p, (k,)).copy() if k else np.zeros(0, np.uint32)
        as_f32 = lambda p, k: np.ctypeslib.as_array(p, (k,)).copy() if k else np.zeros(0, np.float32)
        n_int = max(n - 1, 0)
        self.n = n
        self.num_nodes = b.num_nodes
        self.depth = b.depth
        self.binary_depth = b.binary_depth
        self.keys = as_u32(b.keys, n)
        self.ids = as_u32(b.ids, n)
        self.child = as_u32(b.child, 2 * n_int)
        self.first = as_u32(b.first, n_int)
        self.count = as_u32(b.count, n_int)
        self.aabb = as_f32(b.aabb, 6 * (n_int + n))
        self.leafify = as_u32(b.leafify, n_int)
        self.nodes = as_u32(b.nodes, 20 * b.num_nodes)
        self.leaf = as_f32(b.leaf, 12 * n)

    def trace(self, rays: np.ndarray, any_hit=False, threads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        nv = np.zeros(n, np.uint32)
        tv = np.zeros(n, np.uint32)
        load().o_trace_bvh8(C.byref(self.b), _p(rays), n, int(any_hit), _p(t), _p(inst), _p(prim), _p(u), _p(v),
                            _p(nv), _p(tv), threads)
        return t, inst, prim, u, v, nv, tv

    def __del__(self):
        if _lib is not None and getattr(self, "b", None) is not None:
            _lib.o_bvh_free(C.byref(self.b))
            self.b = None


def trace_brute(world: np.ndarray, rays: np.ndarray, any_hit=False, threads=8):
    world = np.ascontiguousarray(world, np.float32)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.uint32)
    prim = np.zeros(n, np.uint32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    load().o_trace_brute(_p(world), world.shape[0], _p(rays), n, int(any_hit), _p(t), _p(inst), _p(prim), _p(u),
                         _p(v), threads)
    return t, inst, prim, u, v


def brdf_eval(albedo, roughness, metallic, flags, n, wo, wi):
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    a = [np.ascontiguousarray(x, np.float32) for x in (albedo, n, wo, wi)]
    load().o_brdf_eval(_p(a[0]), roughness, metallic, flags, _p(a[1]), _p(a[2]), _p(a[3]), _p(f), C.byref(pdf))
    return f, pdf.value


def brdf_sample(albedo, roughness, metallic, flags, n, wo, ul, u1, u2):
    wi = np.zeros(3, np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (albedo, n, wo)]
    ok = load().o_brdf_sample(_p(a[0]), roughness, metallic, flags, _p(a[1]), _p(a[2]), ul, u1, u2, _p(wi))
    return bool(ok), wi


def camera_matrix(cam):
    v = np.array(cam.view, np.float32)
    p = np.array(cam.proj, np.float32)
    out = np.zeros(16, np.float32)
    load().o_camera_matrix(_p(v), _p(p), _p(out))
    return out


def pt_render(arrs, bvh: "Bvh", cam, w, h, spp, max_depth, rr_start, seed, frame=0, seed_xor=0, threads=8):
    """CPU path tracer.  arrs = (pos, idx, inst5, mat4s, mats16, nrm, emissive_instances);
    returns (h, w, 3) float32 mean radiance."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    inst = np.ascontiguousarray(inst, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    m4 = np.ascontiguousarray(m4, np.float32)
    mats = np.ascontiguousarray(mats, np.uint32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    em = np.ascontiguousarray(em, np.uint32)
    view = np.array(cam.view, np.float32)
    proj = np.array(cam.proj, np.float32)
    cpos = np.array(cam.cam_pos[:3], np.float32)
    out = np.zeros((h, w, 3), np.float32)
    load().o_pt_render(_p(bvh.world), bvh.n, C.byref(bvh.b), _p(inst), inst.shape[0], _p(idx), _p(nrm), _p(m4),
                       _p(mats), _p(em), em.size, _p(view), _p(proj), _p(cpos), w, h, spp, max_depth, rr_start,
                       seed, frame, seed_xor, _p(out), threads)
    return out


def compat_bounce(arrs, bvh: "Bvh", cam, w, h, seed, frame=0, spp=1, dir_lights=None, threads=8):
    """The reference's BounceDiffuse term (RayGen.rgen:139-213) per pixel, as
    SNKPT_MODE_REFERENCE_COMPAT adds it to the direct light; dir_lights: rows of
    (colour rgb, direction towards the light).  Returns (h, w, 3) float32."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    inst = np.ascontiguousarray(inst, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    m4 = np.ascontiguousarray(m4, np.float32)
    mats = np.ascontiguousarray(mats, np.uint32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    dl = np.ascontiguousarray(np.zeros((0, 6)) if dir_lights is None else dir_lights, np.float32).reshape(-1, 6)
    view = np.array(cam.view, np.float32)
    proj = np.array(cam.proj, np.float32)
    cpos = np.array(cam.cam_pos[:3], np.float32)
    out = np.zeros((h, w, 3), np.float32)
    load().o_compat_bounce(_p(bvh.world), bvh.n, C.byref(bvh.b), _p(inst), inst.shape[0], _p(idx), _p(nrm), _p(m4),
                           _p(mats), _p(dl) if dl.size else None, dl.shape[0], _p(view), _p(proj), _p(cpos), w, h,
                           spp, seed, frame, _p(out), threads)
    return out


def emitter_sample(arrs, n, seed=1):
    """The oracle's emitter table for a scene (pt_render's arrs tuple) and n
    draws of its sampler: per emitter (world triangle id, pdf, draw count)."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    world = flatten(pos, idx, inst, m4)
    inst = np.ascontiguousarray(inst, np.uint32)
    mats = np.ascontiguousarray(mats, np.uint32)
    em = np.ascontiguousarray(em, np.uint32)
    cap = int(inst[em, 4].sum() // 3) if em.size else 0
    gid = np.zeros(max(cap, 1), np.uint32)
    pdf = np.zeros(max(cap, 1), np.float32)
    cnt = np.zeros(max(cap, 1), np.uint32)
    ne = load().o_emitter_sample(_p(world), world.shape[0], _p(inst), inst.shape[0], _p(mats), _p(em), em.size, seed,
                                 n, cap, _p(gid), _p(pdf), _p(cnt))
    return gid[:ne], pdf[:ne], cnt[:ne]


def compat_direct(arrs, bvh: "Bvh", cam, w, h, frames, light_block=None, temporal=1, spatial=3, radius=30.0,
                  seed=1, threads=8):
    """The reference's own ReSTIR DI (SNKPT_MODE_REFERENCE_COMPAT with
    SNKPT_FLAG_NO_BOUNCE) over `frames` frames of the static camera `cam`;
    returns the last frame (h, w, 3): emission of directly seen emitters plus
    the direct light."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    inst = np.ascontiguousarray(inst, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    m4 = np.ascontiguousarray(m4, np.float32)
    mats = np.ascontiguousarray(mats, np.uint32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    em = np.ascontiguousarray(em, np.uint32)
    lb = np.frombuffer(light_block, np.uint8) if light_block else np.zeros(0, np.uint8)
    view = np.array(cam.view, np.float32)
    proj = np.array(cam.proj, np.float32)
    pv = np.array(cam.proj_view, np.float32)
    cpos = np.array(cam.cam_pos[:3], np.float32)
    out = np.zeros((h, w, 3), np.float32)
    load().o_compat_direct(_p(bvh.world), bvh.n, C.byref(bvh.b), _p(inst), inst.shape[0], _p(idx), _p(nrm), _p(m4),
                           _p(mats), _p(em), em.size, _p(lb) if lb.size else None, lb.size, _p(view), _p(proj),
                           _p(pv), _p(cpos), w, h, frames, temporal, spatial, radius, seed, _p(out), threads)
    return out


def half_to_float(h: int) -> float:
    return load().o_half_to_float(h)


def primary_surfaces(arrs, bvh: "Bvh", cam, w, h):
    """Surfaces of the camera rays: depth (h, w) clip depth (1 = miss), normal
    (h, w, 3), albedo (h, w, 3), roughness/metallic (h, w, 2), flags (h, w)."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    inst = np.ascontiguousarray(inst, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    m4 = np.ascontiguousarray(m4, np.float32)
    mats = np.ascontiguousarray(mats, np.uint32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    view = np.array(cam.view, np.float32)
    proj = np.array(cam.proj, np.float32)
    pv = np.array(cam.proj_view, np.float32)
    cpos = np.array(cam.cam_pos[:3], np.float32)
    depth = np.zeros((h, w), np.float32)
    n3 = np.zeros((h, w, 3), np.float32)
    a3 = np.zeros((h, w, 3), np.float32)
    rm = np.zeros((h, w, 2), np.float32)
    fl = np.zeros((h, w), np.uint32)
    load().o_primary_surfaces(_p(bvh.world), bvh.n, C.byref(bvh.b), _p(inst), inst.shape[0], _p(idx), _p(nrm),
                              _p(m4), _p(mats), _p(view), _p(proj), _p(pv), _p(cpos), w, h, _p(depth), _p(n3),
                              _p(a3), _p(rm), _p(fl))
    return depth, n3, a3, rm, fl


def encode_gbuffer(depth, normal3, albedo3, rm2, flags, ao=1.0):
    """Host G-buffer in the reference's formats from float surfaces: depth f32,
    normal RGBA16F, albedo RGBA8 SNORM (round(c * 127)), rma RGBA16F, flags u16."""
    h, w = depth.shape
    nrm = np.ones((h, w, 4), np.float16)
    nrm[..., :3] = normal3.astype(np.float16)
    alb = np.full((h, w, 4), 127, np.int8)
    alb[..., :3] = np.clip(np.rint(albedo3 * 127.0), -127, 127).astype(np.int8)
    rma = np.ones((h, w, 4), np.float16)
    rma[..., :2] = rm2.astype(np.float16)
    rma[..., 2] = ao
    return {"depth": np.ascontiguousarray(depth, np.float32), "normal": nrm, "albedo": alb, "rma": rma,
            "mat_flags": flags.astype(np.uint16)}


def pt_render_gbuffer(arrs, bvh: "Bvh", cam, gb, w, h, spp, max_depth, rr_start, seed, frame=0, threads=8):
    """CPU path tracer with each pixel's primary vertex read from the G-buffer
    `gb` (host arrays as encode_gbuffer returns them)."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    inst = np.ascontiguousarray(inst, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    m4 = np.ascontiguousarray(m4, np.float32)
    mats = np.ascontiguousarray(mats, np.uint32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    em = np.ascontiguousarray(em, np.uint32)
    view = np.array(cam.view, np.float32)
    proj = np.array(cam.proj, np.float32)
    cpos = np.array(cam.cam_pos[:3], np.float32)
    keep = {k: np.ascontiguousarray(gb[k]) for k in ("depth", "normal", "albedo", "rma", "mat_flags")}
    g = OGbuffer(_p(keep["depth"]), _p(keep["normal"]), _p(keep["albedo"]), _p(keep["rma"]), _p(keep["mat_flags"]))
    out = np.zeros((h, w, 3), np.float32)
    rc = load().o_pt_render_gbuffer(_p(bvh.world), bvh.n, C.byref(bvh.b), _p(inst), inst.shape[0], _p(idx), _p(nrm),
                                    _p(m4), _p(mats), _p(em), em.size, _p(view), _p(proj), _p(cpos), w, h, spp,
                                    max_depth, rr_start, seed, frame, C.byref(g), _p(out), threads)
    if rc != 0:
        raise ValueError("o_pt_render_gbuffer: incomplete G-buffer")
    return out


def scene_arrays(scene):
    """The pt_render input tuple of a generated snkpt.SceneData."""
    pos, idx, inst, m4, mats = scene.arrays()
    s = scene.s
    nrm = np.ctypeslib.as_array(s.nrm, (s.n_verts * 3,)).copy()
    em = np.ctypeslib.as_array(s.emissive_instances, (s.n_emissive,)).copy().astype(np.uint32)
    return pos, idx, inst, m4, mats, nrm, em


GRID_PARAMS = 16 * (1 << 19) * 2
MLP_PARAMS = 4 * 4096 + 16 * 64


class CacheGeom:
    def __init__(self, lo, hi):
        self.lo_inv = np.zeros(4, np.float32)
        self.res = np.zeros(16, np.uint32)
        self.dense = np.zeros(16, np.uint32)
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        load().o_cache_geom(_p(lo), _p(hi), _p(self.lo_inv), _p(self.res), _p(self.dense))

    def args(self):
        return _p(self.lo_inv), _p(self.res), _p(self.dense)


def grid_index(x, y, z, res, dense):
    return load().o_grid_index(x, y, z, res, dense)


def cache_encode(geom, grid, samples):
    samples = np.ascontiguousarray(samples, np.float32)
    n = samples.size // 16
    out = np.zeros((n, 64), np.float64)
    load().o_cache_encode(*geom.args(), _p(np.ascontiguousarray(grid, np.float32)), _p(samples), n, _p(out))
    return out


def cache_infer(geom, grid, W, samples):
    samples = np.ascontiguousarray(samples, np.float32)
    n = samples.size // 16
    out = np.zeros((n, 3), np.float64)
    load().o_cache_infer(*geom.args(), _p(np.ascontiguousarray(grid, np.float32)), _p(np.ascontiguousarray(W, np.float32)),
                         _p(samples), n, _p(out))
    return out


def cache_grads(geom, grid, W, samples, want_grid=True, den=None, want_den=False):
    """Mean relative-L2 loss and its gradients (stop-gradient through the
    denominators).  den: frozen denominators (n, 3) for finite differences."""
    samples = np.ascontiguousarray(samples, np.float32)
    n = samples.size // 16
    dW = np.zeros(MLP_PARAMS, np.float64)
    dg = np.zeros(GRID_PARAMS, np.float64) if want_grid else None
    dout = np.zeros((n, 3), np.float64) if want_den else None
    din = None if den is None else np.ascontiguousarray(den, np.float64)
    loss = load().o_cache_grads(*geom.args(), _p(np.ascontiguousarray(grid, np.float32)),
                                _p(np.ascontiguousarray(W, np.float32)), _p(samples), n, _p(dW), _p(dg), _p(din),
                                _p(dout))
    if want_den:
        return loss, dW, dg, dout
    return loss, dW, dg


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.99, eps=1e-15):
    p = np.ascontiguousarray(p, np.float32).copy()
    m = np.ascontiguousarray(m, np.float32).copy()
    v = np.ascontiguousarray(v, np.float32).copy()
    g = np.ascontiguousarray(g, np.float64)
    load().o_adam(_p(p), _p(m), _p(v), _p(g), p.size, t, lr, b1, b2, eps)
    return p, m, v
