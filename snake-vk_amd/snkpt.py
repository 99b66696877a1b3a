"""ctypes binding of libsnkpt.so (include/snkpt.h) for tests and the benchmark.

The library is the product: HIP kernels behind a C ABI.  This module only
declares the ABI for Python callers; PyTorch supplies device memory and the
stream handle (plumbing).  If libsnkpt.so is missing, `load()` raises: there
is no CPU fallback.

The `RT` class mirrors the reference pass's interface
(Core/headers/rendering/Raytracing.h:11-31): `Init` binds a scene and output
image, `RecordRenderCmdBuf` renders one frame with `RtSettings`.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsnkpt.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_OOM = -2
ERR_HIP = -3
ERR_CAPACITY = -4
ERR_NOT_BUILT = -5

MAT_FLAG_EMISSIVE = 1
MAT_FLAG_LAMBERT = 2
INVALID_INDEX = 0xFFFFFFFF

MODE_PT = 0
MODE_RESTIR = 1
MODE_CACHE = 2
MODE_REFERENCE_COMPAT = 3  # the reference's biased BounceDiffuse estimator (ReSTIR direct light + 8 + 16 rays)
FLAG_ACCUMULATE = 1
FLAG_RESET = 2
FLAG_NO_TRAIN = 4
FLAG_CACHE_INDIRECT = 8
FLAG_CACHE_MIS = 32  # with FLAG_CACHE_INDIRECT: NEE + a traced BRDF probe under MIS at the query vertex
FLAG_WHOLE_PATHS = 16
FLAG_NO_STATS = 64  # no end-of-frame ray-count readback: PT frames become capturable into a HIP graph
FLAG_NO_BOUNCE = 128  # reference-compat mode: the reference's ReSTIR direct light only, no BounceDiffuse
FLAG_CACHE_TRAIN_AT_QUERY = 256  # cache mode: train only on the training paths' vertices at the query depth
FORMAT_RGBA32F = 0
FORMAT_RGBA16F = 1

# Symbols declared in include/snkpt.h (checked by the CPU test suite).
EXPORTS = [
    "snkpt_create", "snkpt_destroy", "snkpt_last_error", "snkpt_upload_mesh", "snkpt_set_instances",
    "snkpt_set_materials", "snkpt_set_lights", "snkpt_build_accel", "snkpt_get_accel_info",
    "snkpt_accel_export", "snkpt_trace", "snkpt_trace_counted", "snkpt_render", "snkpt_get_frame_stats",
    "snkpt_oracle_accumulate", "snkpt_scene_generate", "snkpt_scene_free", "snkpt_scene_camera",
    "snkpt_scene_load", "snkpt_cache_reset", "snkpt_cache_infer", "snkpt_cache_train", "snkpt_cache_params",
    "snkpt_cache_set_params", "snkpt_scene_view", "snkpt_batch_create", "snkpt_batch_destroy",
    "snkpt_batch_last_error", "snkpt_batch_add_scene", "snkpt_batch_add_view", "snkpt_batch_render",
    "snkpt_batch_autotune", "snkpt_batch_set_tuning", "snkpt_batch_get_tuning", "snkpt_batch_frame_stats",
    "snkpt_render_gbuffer", "snkpt_taa_resolve", "snkpt_trace_contacts", "snkpt_set_prev_transforms",
    "snkpt_import_memory_fd", "snkpt_release_memory", "snkpt_import_semaphore_fd", "snkpt_semaphore_wait",
    "snkpt_semaphore_signal", "snkpt_release_semaphore", "snkpt_cache_infer_noalloc", "snkpt_get_generation",
    "snkpt_get_trace_tuning", "snkpt_set_trace_tuning",
]
CACHE_SAMPLE_FLOATS = 16
CACHE_MLP_PARAMS = 4 * 4096 + 16 * 64
CACHE_GRID_PARAMS = 16 * (1 << 19) * 2


class Instance(C.Structure):
    _fields_ = [("transform_idx", C.c_uint32), ("mesh_buffer_index_offset", C.c_uint32),
                ("mesh_buffer_vertex_offset", C.c_uint32), ("material_idx", C.c_uint32),
                ("num_mesh_indices", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("albedo_tex_idx", C.c_uint32), ("normal_tex_idx", C.c_uint32),
                ("roughness_tex_idx", C.c_uint32), ("metallic_tex_idx", C.c_uint32),
                ("ao_tex_idx", C.c_uint32), ("roughness", C.c_float), ("metallic", C.c_float),
                ("ao", C.c_float), ("albedo", C.c_float * 4), ("flags", C.c_uint32),
                ("pad0", C.c_uint32), ("pad1", C.c_uint32), ("pad2", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("proj_view", C.c_float * 16),
                ("cam_pos", C.c_float * 4), ("cam_forward", C.c_float * 4), ("frame_idx", C.c_uint32),
                ("time_elapsed", C.c_float), ("delta_time", C.c_float), ("prev_frame_invalid", C.c_uint32)]


class Settings(C.Structure):
    _fields_ = [("num_temporal_resamples", C.c_int32), ("num_spatial_resamples", C.c_int32),
                ("mode", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_start", C.c_uint32), ("seed", C.c_uint32), ("frame", C.c_uint32),
                ("restir_candidates", C.c_uint32), ("restir_m_cap", C.c_uint32),
                ("restir_radius", C.c_float), ("cache_query_depth", C.c_uint32),
                ("cache_train_paths", C.c_uint32), ("cache_lr", C.c_float), ("flags", C.c_uint32),
                ("cache_ema", C.c_float), ("cache_train_steps", C.c_uint32),
                ("accum_weight", C.c_float), ("tail_bounce", C.c_uint32),
                ("reserved", C.c_uint32 * 1)]


class Rays(C.Structure):
    _fields_ = [("org_tmin", C.c_void_p), ("dir_tmax", C.c_void_p)]


class Hits(C.Structure):
    _fields_ = [("t", C.c_void_p), ("inst", C.c_void_p), ("prim", C.c_void_p), ("u", C.c_void_p),
                ("v", C.c_void_p)]


class AccelInfo(C.Structure):
    _fields_ = [("num_triangles", C.c_uint32), ("num_nodes", C.c_uint32), ("num_leaf_tris", C.c_uint32),
                ("depth", C.c_uint32), ("binary_depth", C.c_uint32), ("num_emitters", C.c_uint32),
                ("scene_lo", C.c_float * 3), ("scene_hi", C.c_float * 3), ("build_ms", C.c_double)]


class FrameStats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("cache_queries", C.c_uint64), ("cache_train_samples", C.c_uint64),
                ("restir_history_reused", C.c_uint64)]


class Gbuffer(C.Structure):
    """snkpt_gbuffer: device pointers of the rasterized G-buffer images."""
    _fields_ = [("depth", C.c_void_p), ("normal", C.c_void_p), ("albedo", C.c_void_p), ("rma", C.c_void_p),
                ("mat_flags", C.c_void_p), ("motion", C.c_void_p)]


class Contacts(C.Structure):
    """snkpt_contacts: device arrays of the contact records."""
    _fields_ = [("t", C.c_void_p), ("normal", C.c_void_p), ("position", C.c_void_p), ("velocity", C.c_void_p),
                ("inst", C.c_void_p), ("prim", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("trace_block", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("queue_capacity", C.c_uint32), ("ms_per_frame", C.c_float),
                ("tail_bounce", C.c_uint32), ("reserved", C.c_uint32 * 2)]

    def as_dict(self):
        return {"trace_block": self.trace_block, "trace_blocks_per_cu": self.trace_blocks_per_cu,
                "shade_blocks_per_cu": self.shade_blocks_per_cu, "queue_capacity": self.queue_capacity,
                "tail_bounce": self.tail_bounce, "ms_per_frame": round(self.ms_per_frame, 4)}


class TraceTuning(C.Structure):
    _fields_ = [("refill_lanes", C.c_uint32 * 2), ("steps_per_check", C.c_uint32 * 2),
                ("blocks_per_cu", C.c_uint32 * 2), ("ms", C.c_float * 2), ("fixed", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]

    def as_dict(self):
        return {k: [getattr(self, k)[0], getattr(self, k)[1]] for k in ("refill_lanes", "steps_per_check",
                                                                         "blocks_per_cu")} | \
            {"ms": [round(self.ms[0], 4), round(self.ms[1], 4)], "fixed": int(self.fixed)}


class Scene(C.Structure):
    _fields_ = [("pos", C.POINTER(C.c_float)), ("nrm", C.POINTER(C.c_float)), ("uv", C.POINTER(C.c_float)),
                ("tan", C.POINTER(C.c_float)), ("n_verts", C.c_uint32), ("idx", C.POINTER(C.c_uint32)),
                ("n_idx", C.c_uint32), ("inst", C.POINTER(Instance)), ("n_inst", C.c_uint32),
                ("mat4s", C.POINTER(C.c_float)), ("n_mat4", C.c_uint32), ("mats", C.POINTER(Material)),
                ("n_mats", C.c_uint32), ("emissive_instances", C.POINTER(C.c_uint32)),
                ("n_emissive", C.c_uint32), ("camera", Camera), ("width", C.c_uint32),
                ("height", C.c_uint32), ("max_depth", C.c_uint32), ("rr_start", C.c_uint32),
                ("seed", C.c_uint32)]


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libsnkpt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(make -C snake-vk_amd/csrc); there is no CPU fallback")
    if "torch" in sys.modules:
        # libsnkpt shares torch's HIP runtime: let torch initialise it first (its
        # lazy initialisation raised "No HIP GPUs are available" when a snkpt
        # context had been created before the first torch.cuda call)
        sys.modules["torch"].cuda.is_available()
    lib = C.CDLL(path)
    P, U32, I32, VP, SZ = C.POINTER, C.c_uint32, C.c_int32, C.c_void_p, C.c_size_t
    sig = {
        "snkpt_create": (I32, [C.c_int, P(VP)]),
        "snkpt_destroy": (None, [VP]),
        "snkpt_last_error": (C.c_char_p, [VP]),
        "snkpt_upload_mesh": (I32, [VP, VP, VP, VP, VP, U32, VP, U32, P(U32), P(U32)]),
        "snkpt_set_instances": (I32, [VP, VP, U32, VP, U32]),
        "snkpt_set_materials": (I32, [VP, VP, U32]),
        "snkpt_set_lights": (I32, [VP, VP, SZ, VP, U32]),
        "snkpt_build_accel": (I32, [VP, VP]),
        "snkpt_get_accel_info": (I32, [VP, P(AccelInfo)]),
        "snkpt_accel_export": (I32, [VP, U32, VP, SZ, P(SZ)]),
        "snkpt_trace": (I32, [VP, P(Rays), U32, P(Hits), C.c_int, VP]),
        "snkpt_trace_counted": (I32, [VP, P(Rays), U32, P(Hits), C.c_int, VP, VP, VP]),
        "snkpt_render": (I32, [VP, P(Camera), P(Settings), VP, U32, U32, U32, VP]),
        "snkpt_get_frame_stats": (I32, [VP, P(FrameStats)]),
        "snkpt_oracle_accumulate": (I32, [VP, P(Camera), P(Settings), VP, U32, U32, U32, VP]),
        "snkpt_scene_generate": (I32, [U32, U32, P(P(Scene))]),
        "snkpt_scene_free": (None, [P(Scene)]),
        "snkpt_scene_camera": (I32, [U32, U32, U32, U32, U32, P(Camera)]),
        "snkpt_scene_load": (I32, [VP, P(Scene)]),
        "snkpt_cache_reset": (I32, [VP, U32, VP, VP, VP]),
        "snkpt_cache_infer": (I32, [VP, VP, U32, VP, VP]),
        "snkpt_cache_infer_noalloc": (I32, [VP, VP, U32, VP, VP]),
        "snkpt_get_generation": (I32, [VP, VP]),
        "snkpt_get_trace_tuning": (I32, [VP, VP]),
        "snkpt_set_trace_tuning": (I32, [VP, VP]),
        "snkpt_cache_train": (I32, [VP, VP, U32, C.c_float, C.c_int, P(C.c_float), VP]),
        "snkpt_cache_params": (I32, [VP, U32, VP, SZ, P(SZ)]),
        "snkpt_cache_set_params": (I32, [VP, U32, VP, SZ]),
        "snkpt_scene_view": (I32, [U32, U32, P(U32), P(U32), P(U32)]),
        "snkpt_batch_create": (I32, [C.c_int, P(VP)]),
        "snkpt_batch_destroy": (None, [VP]),
        "snkpt_batch_last_error": (C.c_char_p, [VP]),
        "snkpt_batch_add_scene": (I32, [VP, VP, P(U32)]),
        "snkpt_batch_add_view": (I32, [VP, U32, P(Camera), U32, U32, U32, VP, U32, P(U32)]),
        "snkpt_batch_render": (I32, [VP, P(Settings), C.c_int, VP]),
        "snkpt_batch_autotune": (I32, [VP, P(Settings), VP, P(BatchTuning)]),
        "snkpt_batch_set_tuning": (I32, [VP, P(BatchTuning)]),
        "snkpt_batch_get_tuning": (I32, [VP, P(BatchTuning)]),
        "snkpt_batch_frame_stats": (I32, [VP, P(FrameStats), VP]),
        "snkpt_render_gbuffer": (I32, [VP, P(Camera), P(Settings), P(Gbuffer), VP, U32, U32, U32, VP]),
        "snkpt_taa_resolve": (I32, [VP, VP, VP, VP, U32, U32, U32, U32, VP]),
        "snkpt_trace_contacts": (I32, [VP, P(Rays), U32, P(Contacts), C.c_float, VP]),
        "snkpt_set_prev_transforms": (I32, [VP, VP, U32]),
        "snkpt_import_memory_fd": (I32, [VP, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, P(VP), P(VP)]),
        "snkpt_release_memory": (I32, [VP, VP]),
        "snkpt_import_semaphore_fd": (I32, [VP, C.c_int, C.c_int, P(VP)]),
        "snkpt_semaphore_wait": (I32, [VP, VP, C.c_uint64, VP]),
        "snkpt_semaphore_signal": (I32, [VP, VP, C.c_uint64, VP]),
        "snkpt_release_semaphore": (I32, [VP, VP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SnkptError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"snkpt error {code}: {msg}")
        self.code = code


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class SceneData:
    """A generated synthetic scene (host arrays owned by the library)."""

    def __init__(self, config: int, variant: int = 0):
        lib = load()
        self._p = C.POINTER(Scene)()
        rc = lib.snkpt_scene_generate(config, variant, C.byref(self._p))
        if rc != OK:
            raise SnkptError(rc, f"scene_generate({config}, {variant})")
        self.config, self.variant = config, variant

    @property
    def s(self) -> Scene:
        return self._p.contents

    def arrays(self):
        """numpy views: pos (V,3), idx (I,), inst (N,5) u32, mat4s (T,16), mats (M,16) raw u32."""
        import numpy as np
        s = self.s
        pos = np.ctypeslib.as_array(s.pos, (s.n_verts * 3,)).reshape(-1, 3)
        idx = np.ctypeslib.as_array(s.idx, (s.n_idx,))
        inst = np.ctypeslib.as_array(C.cast(s.inst, C.POINTER(C.c_uint32)), (s.n_inst * 5,)).reshape(-1, 5)
        m4 = np.ctypeslib.as_array(s.mat4s, (s.n_mat4 * 16,)).reshape(-1, 16)
        mats = np.ctypeslib.as_array(C.cast(s.mats, C.POINTER(C.c_uint32)), (s.n_mats * 16,)).reshape(-1, 16)
        return pos, idx, inst, m4, mats

    def camera(self, frame: int = 0, w: int | None = None, h: int | None = None) -> Camera:
        cam = Camera()
        w = w or self.s.width
        h = h or self.s.height
        rc = load().snkpt_scene_camera(self.config, self.variant, frame, w, h, C.byref(cam))
        if rc != OK:
            raise SnkptError(rc, "scene_camera")
        return cam

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.snkpt_scene_free(self._p)
            self._p = None


def light_block(dir_colour=(0.0, 0.0, 0.0), dir_to_light=(0.0, 1.0, 0.0), points=(), spots=()) -> bytes:
    """The reference's LightData SSBO (LightBuffers.glsl:1-63, 114,800 B):
    points = [(colour3, pos3, (atten_exp, atten_linear, atten_constant))],
    spots = [(colour3, pos3, dir3, (atten_exp, atten_linear, atten_constant), aperture)]."""
    import struct
    b = bytearray(112 + 48 * 1024 + 64 * 1024)
    struct.pack_into("4f4f", b, 0, *dir_colour, 1.0, *dir_to_light, 0.0)
    struct.pack_into("2I", b, 96, len(points), len(spots))
    for i, (c, p, att) in enumerate(points):
        struct.pack_into("4f4f4f", b, 112 + 48 * i, *c, 1.0, *p, 1.0, *att, 0.0)
    for i, (c, p, d, att, ap) in enumerate(spots):
        struct.pack_into("4f4f4f4f", b, 112 + 48 * 1024 + 64 * i, *c, 1.0, *p, 1.0, *d, 0.0, *att, ap)
    return bytes(b)


def scene_view(config: int, view: int):
    """(variant, w, h) of view `view` of the batched workload (config 5)."""
    v, w, h = C.c_uint32(), C.c_uint32(), C.c_uint32()
    rc = load().snkpt_scene_view(config, view, C.byref(v), C.byref(w), C.byref(h))
    if rc != OK:
        raise SnkptError(rc, "scene_view")
    return v.value, w.value, h.value


class Batch:
    """Many views of many scenes rendered together (snkpt_batch_*).  Scenes
    are Contexts with a built acceleration structure; outputs are device
    tensors (h, w, 4) float32 or float16 that must stay alive."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self._h = C.c_void_p()
        rc = self.lib.snkpt_batch_create(device, C.byref(self._h))
        if rc != OK:
            raise SnkptError(rc, "snkpt_batch_create")
        self._keep = []

    def _check(self, rc: int, what: str):
        if rc != OK:
            msg = self.lib.snkpt_batch_last_error(self._h)
            raise SnkptError(rc, f"{what}: {msg.decode() if msg else ''}")

    def add_scene(self, ctx) -> int:
        out = C.c_uint32()
        self._check(self.lib.snkpt_batch_add_scene(self._h, ctx._h, C.byref(out)), "snkpt_batch_add_scene")
        self._keep.append(ctx)
        return out.value

    def add_view(self, scene: int, cam: Camera, w: int, h: int, seed: int, out, fmt: int = FORMAT_RGBA32F) -> int:
        v = C.c_uint32()
        self._check(self.lib.snkpt_batch_add_view(self._h, scene, C.byref(cam), w, h, seed, _ptr(out), fmt,
                                                  C.byref(v)), "snkpt_batch_add_view")
        self._keep.append(out)
        return v.value

    def render(self, settings: Settings, use_graph: bool = False, stream=None):
        self._check(self.lib.snkpt_batch_render(self._h, C.byref(settings), int(use_graph), _stream(stream)),
                    "snkpt_batch_render")

    def autotune(self, settings: Settings, stream=None) -> BatchTuning:
        t = BatchTuning()
        self._check(self.lib.snkpt_batch_autotune(self._h, C.byref(settings), _stream(stream), C.byref(t)),
                    "snkpt_batch_autotune")
        return t

    def set_tuning(self, t: BatchTuning):
        self._check(self.lib.snkpt_batch_set_tuning(self._h, C.byref(t)), "snkpt_batch_set_tuning")

    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        self._check(self.lib.snkpt_batch_get_tuning(self._h, C.byref(t)), "snkpt_batch_get_tuning")
        return t

    def frame_stats(self, stream=None) -> FrameStats:
        st = FrameStats()
        self._check(self.lib.snkpt_batch_frame_stats(self._h, C.byref(st), _stream(stream)),
                    "snkpt_batch_frame_stats")
        return st

    def close(self):
        if self._h:
            self.lib.snkpt_batch_destroy(self._h)
            self._h = C.c_void_p()
        self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        self._h = C.c_void_p()
        rc = self.lib.snkpt_create(device, C.byref(self._h))
        if rc != OK:
            raise SnkptError(rc, f"snkpt_create(device={device})")

    def close(self):
        if self._h:
            self.lib.snkpt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != OK:
            msg = self.lib.snkpt_last_error(self._h)
            raise SnkptError(rc, f"{what}: {msg.decode() if msg else ''}")

    # ---- scene input
    def load_scene(self, scene: SceneData):
        self._check(self.lib.snkpt_scene_load(self._h, scene._p), "snkpt_scene_load")

    def upload_mesh(self, pos, idx, nrm=None, uv=None, tan=None):
        import numpy as np
        pos = np.ascontiguousarray(pos, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (nrm, uv, tan)]
        vo, io = C.c_uint32(), C.c_uint32()
        p = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in arrs]
        rc = self.lib.snkpt_upload_mesh(self._h, pos.ctypes.data_as(C.c_void_p), p[0], p[1], p[2],
                                        pos.size // 3, idx.ctypes.data_as(C.c_void_p), idx.size,
                                        C.byref(vo), C.byref(io))
        self._check(rc, "snkpt_upload_mesh")
        return vo.value, io.value

    def set_instances(self, inst5, mat4s):
        import numpy as np
        inst5 = np.ascontiguousarray(inst5, np.uint32).reshape(-1, 5)
        mat4s = np.ascontiguousarray(mat4s, np.float32).reshape(-1, 16)
        rc = self.lib.snkpt_set_instances(self._h, inst5.ctypes.data_as(C.c_void_p), inst5.shape[0],
                                          mat4s.ctypes.data_as(C.c_void_p), mat4s.shape[0])
        self._check(rc, "snkpt_set_instances")

    def set_materials(self, mats):
        arr = (Material * len(mats))(*mats)
        self._check(self.lib.snkpt_set_materials(self._h, C.cast(arr, C.c_void_p), len(mats)),
                    "snkpt_set_materials")

    def set_lights(self, light_block: bytes | None = None, emissive=None):
        import numpy as np
        lb = None if light_block is None else C.create_string_buffer(light_block, len(light_block))
        em = None if emissive is None else np.ascontiguousarray(emissive, np.uint32)
        rc = self.lib.snkpt_set_lights(self._h, None if lb is None else C.cast(lb, C.c_void_p),
                                       0 if light_block is None else len(light_block),
                                       None if em is None else em.ctypes.data_as(C.c_void_p),
                                       0 if em is None else em.size)
        self._check(rc, "snkpt_set_lights")

    def build_accel(self, stream=None):
        self._check(self.lib.snkpt_build_accel(self._h, _stream(stream)), "snkpt_build_accel")

    def accel_info(self) -> AccelInfo:
        info = AccelInfo()
        self._check(self.lib.snkpt_get_accel_info(self._h, C.byref(info)), "snkpt_get_accel_info")
        return info

    def export(self, which: int):
        import numpy as np
        nb = C.c_size_t()
        self._check(self.lib.snkpt_accel_export(self._h, which, None, 0, C.byref(nb)), "snkpt_accel_export")
        buf = np.zeros(max(nb.value, 4) // 4, np.uint32)
        self._check(self.lib.snkpt_accel_export(self._h, which, buf.ctypes.data_as(C.c_void_p), buf.nbytes,
                                                C.byref(nb)), "snkpt_accel_export")
        return buf[: nb.value // 4]

    # ---- tracing (device tensors)
    def trace(self, org_tmin, dir_tmax, any_hit=False, stream=None, with_uv=True, counted=False):
        import torch
        n = org_tmin.shape[0]
        dev = org_tmin.device
        t = torch.empty(n, dtype=torch.float32, device=dev)
        inst = torch.empty(n, dtype=torch.int32, device=dev)
        prim = torch.empty(n, dtype=torch.int32, device=dev)
        u = torch.empty(n, dtype=torch.float32, device=dev) if with_uv else None
        v = torch.empty(n, dtype=torch.float32, device=dev) if with_uv else None
        rays = Rays(org_tmin.data_ptr(), dir_tmax.data_ptr())
        hits = Hits(t.data_ptr(), inst.data_ptr(), prim.data_ptr(), u.data_ptr() if with_uv else None,
                    v.data_ptr() if with_uv else None)
        if counted:
            nv = torch.empty(n, dtype=torch.int32, device=dev)
            tv = torch.empty(n, dtype=torch.int32, device=dev)
            rc = self.lib.snkpt_trace_counted(self._h, C.byref(rays), n, C.byref(hits), int(any_hit),
                                              _ptr(nv), _ptr(tv), _stream(stream))
            self._check(rc, "snkpt_trace_counted")
            return t, inst, prim, u, v, nv, tv
        rc = self.lib.snkpt_trace(self._h, C.byref(rays), n, C.byref(hits), int(any_hit), _stream(stream))
        self._check(rc, "snkpt_trace")
        return t, inst, prim, u, v

    def trace_into(self, org_tmin, dir_tmax, t, inst=None, prim=None, any_hit=False, stream=None):
        """Trace into caller-owned output tensors (no allocation)."""
        rays = Rays(org_tmin.data_ptr(), dir_tmax.data_ptr())
        hits = Hits(t.data_ptr(), None if inst is None else inst.data_ptr(),
                    None if prim is None else prim.data_ptr(), None, None)
        rc = self.lib.snkpt_trace(self._h, C.byref(rays), org_tmin.shape[0], C.byref(hits), int(any_hit),
                                  _stream(stream))
        self._check(rc, "snkpt_trace")

    # ---- the pass
    def render(self, cam: Camera, settings: Settings, out, w: int, h: int, fmt: int = FORMAT_RGBA32F,
               stream=None):
        rc = self.lib.snkpt_render(self._h, C.byref(cam), C.byref(settings), _ptr(out), w, h, fmt,
                                   _stream(stream))
        self._check(rc, "snkpt_render")

    def render_gbuffer(self, cam: Camera, settings: Settings, gbuffer: dict, out, w: int, h: int,
                       fmt: int = FORMAT_RGBA32F, stream=None):
        """One frame with the primary vertices read from a G-buffer: `gbuffer`
        maps depth (float32 h x w), normal (float16 h x w x 4), albedo (int8
        h x w x 4, SNORM), rma (float16 h x w x 4), mat_flags (int16/uint16
        h x w) and optionally motion (float16 h x w x 2) to device tensors."""
        g = Gbuffer()
        for k in ("depth", "normal", "albedo", "rma", "mat_flags", "motion"):
            t = gbuffer.get(k)
            if t is not None and not t.is_contiguous():
                raise ValueError(f"G-buffer image {k} must be contiguous")
            setattr(g, k, None if t is None else t.data_ptr())
        rc = self.lib.snkpt_render_gbuffer(self._h, C.byref(cam), C.byref(settings), C.byref(g), _ptr(out), w, h,
                                           fmt, _stream(stream))
        self._check(rc, "snkpt_render_gbuffer")

    def trace_contacts(self, org_tmin, dir_tmax, timestep: float = 1.0 / 120.0, stream=None):
        """Terminate-on-first-hit collision rays with ParticleUpdate.rchit's contact
        record.  org_tmin / dir_tmax: (n, 4) float32 device tensors.  Returns
        (t, normal, position, velocity, inst, prim) device tensors."""
        import torch
        n = org_tmin.shape[0]
        dev = org_tmin.device
        t = torch.empty(n, dtype=torch.float32, device=dev)
        nrm = torch.empty(n, 4, dtype=torch.float32, device=dev)
        pos = torch.empty(n, 4, dtype=torch.float32, device=dev)
        vel = torch.empty(n, 4, dtype=torch.float32, device=dev)
        inst = torch.empty(n, dtype=torch.int32, device=dev)
        prim = torch.empty(n, dtype=torch.int32, device=dev)
        rays = Rays(org_tmin.data_ptr(), dir_tmax.data_ptr())
        c = Contacts(t.data_ptr(), nrm.data_ptr(), pos.data_ptr(), vel.data_ptr(), inst.data_ptr(), prim.data_ptr())
        self._check(self.lib.snkpt_trace_contacts(self._h, C.byref(rays), n, C.byref(c), timestep, _stream(stream)),
                    "snkpt_trace_contacts")
        return t, nrm, pos, vel, inst, prim

    # ---- Vulkan interop
    def import_memory_fd(self, fd: int, allocation_size: int, offset: int, size: int):
        """Imports a range of an exported (OPAQUE_FD) allocation; returns
        (device pointer as int, handle).  The library owns fd on success."""
        ptr, h = C.c_void_p(), C.c_void_p()
        self._check(self.lib.snkpt_import_memory_fd(self._h, fd, allocation_size, offset, size, C.byref(ptr),
                                                    C.byref(h)), "snkpt_import_memory_fd")
        return ptr.value, h

    def release_memory(self, handle):
        self._check(self.lib.snkpt_release_memory(self._h, handle), "snkpt_release_memory")

    def import_semaphore_fd(self, fd: int, timeline: bool = False):
        h = C.c_void_p()
        self._check(self.lib.snkpt_import_semaphore_fd(self._h, fd, int(timeline), C.byref(h)),
                    "snkpt_import_semaphore_fd")
        return h

    def semaphore_wait(self, sem, value: int = 0, stream=None):
        self._check(self.lib.snkpt_semaphore_wait(self._h, sem, value, _stream(stream)), "snkpt_semaphore_wait")

    def semaphore_signal(self, sem, value: int = 0, stream=None):
        self._check(self.lib.snkpt_semaphore_signal(self._h, sem, value, _stream(stream)), "snkpt_semaphore_signal")

    def release_semaphore(self, sem):
        self._check(self.lib.snkpt_release_semaphore(self._h, sem), "snkpt_release_semaphore")

    def set_prev_transforms(self, mat4s):
        import numpy as np
        m = np.ascontiguousarray(mat4s, np.float32).reshape(-1)
        self._check(self.lib.snkpt_set_prev_transforms(self._h, m.ctypes.data_as(C.c_void_p), m.size // 16),
                    "snkpt_set_prev_transforms")

    def taa_resolve(self, current, out, velocity, w: int, h: int, fmt: int = FORMAT_RGBA32F, reset: bool = False,
                    stream=None):
        """Temporal resolve of `current` into `out` (may be the same tensor);
        velocity: float16 (h, w, 2) device tensor."""
        rc = self.lib.snkpt_taa_resolve(self._h, _ptr(current), _ptr(out), _ptr(velocity), w, h, fmt, int(reset),
                                        _stream(stream))
        self._check(rc, "snkpt_taa_resolve")

    def trace_tuning(self) -> TraceTuning:
        t = TraceTuning()
        self._check(self.lib.snkpt_get_trace_tuning(self._h, C.byref(t)), "snkpt_get_trace_tuning")
        return t

    def set_trace_tuning(self, t: TraceTuning | None):
        self._check(self.lib.snkpt_set_trace_tuning(self._h, None if t is None else C.byref(t)),
                    "snkpt_set_trace_tuning")

    def generation(self) -> int:
        """The context's state generation (changes whenever a captured frame
        or a batch would have to be recorded again)."""
        g = C.c_uint64()
        self._check(self.lib.snkpt_get_generation(self._h, C.byref(g)), "snkpt_get_generation")
        return g.value

    def frame_stats(self) -> FrameStats:
        st = FrameStats()
        self._check(self.lib.snkpt_get_frame_stats(self._h, C.byref(st)), "snkpt_get_frame_stats")
        return st

    # ---- MLP radiance cache
    def cache_reset(self, seed: int = 1, lo=None, hi=None, stream=None):
        import numpy as np
        lo_a = None if lo is None else np.ascontiguousarray(lo, np.float32)
        hi_a = None if hi is None else np.ascontiguousarray(hi, np.float32)
        rc = self.lib.snkpt_cache_reset(self._h, seed, None if lo_a is None else lo_a.ctypes.data_as(C.c_void_p),
                                        None if hi_a is None else hi_a.ctypes.data_as(C.c_void_p), _stream(stream))
        self._check(rc, "snkpt_cache_reset")

    def cache_infer(self, samples, stream=None):
        import torch
        n = samples.shape[0]
        out = torch.empty(n, 4, dtype=torch.float32, device=samples.device)
        self._check(self.lib.snkpt_cache_infer(self._h, _ptr(samples), n, _ptr(out), _stream(stream)),
                    "snkpt_cache_infer")
        return out

    def cache_infer_noalloc(self, samples, out, stream=None):
        """snkpt_cache_infer into a caller-given (n, 4) float32 device tensor,
        with the library's static workspace (n <= 65,536; no allocation)."""
        self._check(self.lib.snkpt_cache_infer_noalloc(self._h, _ptr(samples), samples.shape[0], _ptr(out),
                                                       _stream(stream)), "snkpt_cache_infer_noalloc")
        return out

    def cache_train(self, samples, lr: float = 1e-2, apply: bool = True, stream=None) -> float:
        loss = C.c_float()
        rc = self.lib.snkpt_cache_train(self._h, _ptr(samples), samples.shape[0], lr, int(apply), C.byref(loss),
                                        _stream(stream))
        self._check(rc, "snkpt_cache_train")
        return loss.value

    def cache_params(self, which: int):
        import numpy as np
        nb = C.c_size_t()
        self._check(self.lib.snkpt_cache_params(self._h, which, None, 0, C.byref(nb)), "snkpt_cache_params")
        if which == 4:
            buf = np.zeros(max(nb.value, 2) // 2, np.float16)
        else:
            buf = np.zeros(max(nb.value, 4) // 4, np.float32)
        self._check(self.lib.snkpt_cache_params(self._h, which, buf.ctypes.data_as(C.c_void_p), buf.nbytes,
                                                C.byref(nb)), "snkpt_cache_params")
        return buf

    def cache_set_params(self, which: int, arr):
        import numpy as np
        a = np.ascontiguousarray(arr, np.float32)
        self._check(self.lib.snkpt_cache_set_params(self._h, which, a.ctypes.data_as(C.c_void_p), a.nbytes),
                    "snkpt_cache_set_params")

    def oracle_accumulate(self, cam: Camera, settings: Settings, accum, w: int, h: int, spp: int, stream=None):
        rc = self.lib.snkpt_oracle_accumulate(self._h, C.byref(cam), C.byref(settings), _ptr(accum), w, h, spp,
                                              _stream(stream))
        self._check(rc, "snkpt_oracle_accumulate")


def default_settings(**kw) -> Settings:
    s = Settings()
    s.num_temporal_resamples = 1
    s.num_spatial_resamples = 5
    s.mode = MODE_PT
    s.spp = 1
    s.max_depth = 4
    s.rr_start = 0
    s.seed = 1
    s.frame = 0
    s.restir_candidates = 32
    s.restir_m_cap = 20
    s.restir_radius = 30.0
    s.cache_query_depth = 2
    s.cache_train_paths = 65536
    s.cache_lr = 1e-2
    s.cache_ema = 0.99
    s.flags = 0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class RT:
    """Mirror of SNAKE::RT (Raytracing.h:11-31) over the C ABI.

    Init(scene, output_image) binds a context to a loaded scene and an output
    image tensor (H, W, 4); RecordRenderCmdBuf(stream, output_image, scene,
    settings) renders one frame into it.  Where the reference logs and returns
    when the scene has no TLAS system (Raytracing.cpp:270-275), this raises.
    """

    def __init__(self):
        self.ctx = None

    def Init(self, scene: SceneData, output_image, device: int = 0):
        self.ctx = Context(device)
        self.ctx.load_scene(scene)
        self.ctx.build_accel()
        self.scene = scene
        self.output_image = output_image

    def RecordRenderCmdBuf(self, stream, output_image, scene: SceneData, settings: Settings, cam: Camera = None,
                           fmt: int = FORMAT_RGBA32F, resources: dict = None):
        """resources: the G-buffer images of RaytracingResources (RenderCommon.h:6-35)
        as in Context.render_gbuffer; None traces camera rays instead."""
        if self.ctx is None:
            raise SnkptError(ERR_NOT_BUILT, "RT::Init was not called")
        h, w = output_image.shape[0], output_image.shape[1]
        cam = cam or scene.camera(settings.frame, w, h)
        if resources is not None:
            self.ctx.render_gbuffer(cam, settings, resources, output_image, w, h, fmt, stream)
        else:
            self.ctx.render(cam, settings, output_image, w, h, fmt, stream)
