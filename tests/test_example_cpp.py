"""The C++ host program over the C ABI (snake-vk_amd/examples/render_room.cpp):
it builds against include/snkpt.h and links libsnkpt.so (CPU), and on the GPU
its accumulated image equals the same frames rendered through the Python
binding bit for bit."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "snake-vk_amd", "examples", "render_room")


def test_example_is_built_and_links_the_library():
    assert os.path.exists(EXE), "build() compiles the example next to its source"
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True, check=True).stdout
    assert "libsnkpt.so" in out


def _read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        assert scale < 0  # little endian
        a = np.frombuffer(f.read(), "<f4").reshape(h, w, 3)
    return a[::-1]  # bottom row first


@pytest.mark.gpu
def test_example_image_equals_the_binding(gpu, tmp_path):
    import torch

    import snkpt
    pfm = str(tmp_path / "room.pfm")
    r = subprocess.run([EXE, "8", pfm], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _read_pfm(pfm)
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w, h = scene.s.width, scene.s.height
    out = torch.zeros(h, w, 4, device="cuda")
    for f in range(8):
        fl = snkpt.FLAG_ACCUMULATE | (snkpt.FLAG_RESET if f == 0 else 0)
        s = snkpt.default_settings(max_depth=scene.s.max_depth, rr_start=scene.s.rr_start, seed=7, frame=f,
                                   flags=fl)
        ctx.render(scene.camera(0, w, h), s, out, w, h)
    want = out[..., :3].cpu().numpy()
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert float(want.mean()) > 0.0
