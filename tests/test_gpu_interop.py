"""Vulkan interop entry points (snkpt_import_memory_fd and the semaphore
imports).  No Vulkan driver is installed, so an exportable allocation is made
with HIP's virtual-memory API (hipMemCreate with a POSIX file-descriptor
handle, hipMemExportToShareableHandle), the same kind of fd
vkGetMemoryFdKHR hands over; the pass renders into the imported pointer and
the image is read back through the allocation's own mapping."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Loc(C.Structure):
    _fields_ = [("type", C.c_int), ("id", C.c_int)]


class _AllocFlags(C.Structure):
    _fields_ = [("compressionType", C.c_ubyte), ("gpuDirectRDMACapable", C.c_ubyte), ("usage", C.c_ushort)]


class _Prop(C.Structure):
    _fields_ = [("type", C.c_int), ("requestedHandleType", C.c_int), ("location", _Loc),
                ("win32HandleMetaData", C.c_void_p), ("allocFlags", _AllocFlags)]


class _Access(C.Structure):
    _fields_ = [("location", _Loc), ("flags", C.c_int)]


class _Exported:
    """A VMM allocation exported as a file descriptor and mapped in this process."""

    def __init__(self, nbytes, device=0):
        import torch
        torch.cuda.init()
        self.hip = C.CDLL("libamdhip64.so.7")  # the runtime torch and libsnkpt already share
        h = self.hip
        for name in ("hipMemGetAllocationGranularity", "hipMemCreate", "hipMemExportToShareableHandle",
                     "hipMemAddressReserve", "hipMemMap", "hipMemSetAccess", "hipMemcpy", "hipMemUnmap",
                     "hipMemAddressFree", "hipMemRelease"):
            getattr(h, name).restype = C.c_int
        prop = _Prop(1, 1, _Loc(1, device), None, _AllocFlags(0, 0, 0))  # pinned, POSIX fd, device
        gran = C.c_size_t()
        self._ok(h.hipMemGetAllocationGranularity(C.byref(gran), C.byref(prop), 0), "granularity")
        self.size = (nbytes + gran.value - 1) // gran.value * gran.value
        self.handle = C.c_void_p()
        self._ok(h.hipMemCreate(C.byref(self.handle), C.c_size_t(self.size), C.byref(prop), C.c_ulonglong(0)),
                 "hipMemCreate")
        fd = C.c_int(-1)
        self._ok(h.hipMemExportToShareableHandle(C.byref(fd), self.handle, 1, C.c_ulonglong(0)), "export")
        self.fd = fd.value
        self.ptr = C.c_void_p()
        self._ok(h.hipMemAddressReserve(C.byref(self.ptr), C.c_size_t(self.size), C.c_size_t(0), None,
                                        C.c_ulonglong(0)), "reserve")
        self._ok(h.hipMemMap(self.ptr, C.c_size_t(self.size), C.c_size_t(0), self.handle, C.c_ulonglong(0)), "map")
        acc = _Access(_Loc(1, device), 3)
        self._ok(h.hipMemSetAccess(self.ptr, C.c_size_t(self.size), C.byref(acc), C.c_size_t(1)), "access")

    def _ok(self, rc, what):
        if rc != 0:
            pytest.skip(f"no exportable allocation on this runtime ({what}: hipError {rc})")

    def read(self, nbytes):
        out = np.empty(nbytes, np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(nbytes), 2) == 0  # D2H
        return out

    def write(self, data):
        data = np.ascontiguousarray(data)
        assert self.hip.hipMemcpy(self.ptr, data.ctypes.data_as(C.c_void_p), C.c_size_t(data.nbytes), 1) == 0

    def close(self):
        h = self.hip
        h.hipMemUnmap(self.ptr, C.c_size_t(self.size))
        h.hipMemAddressFree(self.ptr, C.c_size_t(self.size))
        h.hipMemRelease(self.handle)


class _Raw:
    def __init__(self, ptr):
        self.ptr = ptr

    def data_ptr(self):
        return self.ptr


def test_render_into_imported_memory(gpu):
    import torch

    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w = h = 64
    nbytes = w * h * 16
    ex = _Exported(nbytes)
    _say(f"exported {ex.size} B as fd {ex.fd}")
    pattern = np.arange(nbytes, dtype=np.uint32).astype(np.uint8)
    ex.write(pattern)
    ptr, handle = ctx.import_memory_fd(ex.fd, ex.size, 0, nbytes)
    _say("imported; reading the pattern through the imported pointer")
    got = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMemcpy.restype = C.c_int
    assert rt.hipMemcpy(C.c_void_p(got.data_ptr()), C.c_void_p(ptr), C.c_size_t(nbytes), 3) == 0  # D2D
    assert np.array_equal(got.cpu().numpy(), pattern)
    _say("rendering into the imported pointer")
    cam = scene.camera(0, w, h)
    s = snkpt.default_settings(max_depth=3, seed=5)
    ctx.render(cam, s, _Raw(ptr), w, h)
    want = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, s, want, w, h)
    torch.cuda.synchronize()
    img = ex.read(nbytes).view(np.float32).reshape(h, w, 4)
    assert np.array_equal(img, want.cpu().numpy())
    ctx.release_memory(handle)
    ex.close()


def test_interop_rejects_bad_handles(gpu):
    import snkpt
    ctx = snkpt.Context(0)
    with pytest.raises(snkpt.SnkptError):
        ctx.import_memory_fd(-1, 4096, 0, 4096)
    with pytest.raises(snkpt.SnkptError):
        ctx.import_memory_fd(0, 4096, 4000, 200)  # range past the allocation
    with pytest.raises(snkpt.SnkptError):
        ctx.import_semaphore_fd(-1)
