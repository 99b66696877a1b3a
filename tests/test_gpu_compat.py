"""GPU parity of SNKPT_MODE_REFERENCE_COMPAT: the reference's own ReSTIR DI
(RayGen.rgen:215-528, biased combine) and BounceDiffuse term (RayGen.rgen:
139-213).  A compat frame minus the direct-only frame of the same seed
(SNKPT_FLAG_NO_BOUNCE, a second context with the same history) is the bounce
term; it must equal the oracle's restatement pixel by pixel and the closed forms
of the oracle's own tests.  The direct-only frames equal the oracle's
restatement of the reference's ReSTIR pixel by pixel."""
import math
import sys

import numpy as np
import pytest

from helpers import ArrayScene, look_at_camera, make_ctx
from oracle import binding as O

pytestmark = pytest.mark.gpu

REF_PI = 3.1415926


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _frame(ctx, cam, w, h, mode, seed, frame=0, flags=0, **kw):
    import torch

    import snkpt
    out = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, snkpt.default_settings(mode=mode, max_depth=1, seed=seed, frame=frame, flags=flags, **kw), out,
               w, h)
    torch.cuda.synchronize()
    return out[..., :3].cpu().numpy().astype(np.float64).reshape(-1, 3)


def _bounce_gpu(make, cam, w, h, seed=5):
    import snkpt
    a, b = make(), make()
    return (_frame(a, cam, w, h, snkpt.MODE_REFERENCE_COMPAT, seed) -
            _frame(b, cam, w, h, snkpt.MODE_REFERENCE_COMPAT, seed, flags=snkpt.FLAG_NO_BOUNCE))


def _point_lights(n, seed, lo=(-0.9, 0.1, -0.9), hi=(0.9, 1.9, 0.9)):
    rng = np.random.default_rng(seed)
    return [(tuple(rng.uniform(0.5, 3.0, 3)), tuple(rng.uniform(lo, hi)), (0.5, 0.2, 1.0)) for _ in range(n)]


@pytest.mark.parametrize("lights", ["points+dir", "emitters only"])
def test_compat_direct_light_matches_the_oracle_per_pixel(gpu, lights):
    """Config 3's 10,000 emissive triangles (16 emissive instances) with 24 point
    lights and a directional light (or the emitters alone): four frames of the
    static camera through the reference's candidate, temporal and spatial
    passes; the last frame equals the oracle's restatement pixel by pixel."""
    import snkpt
    scene = snkpt.SceneData(3)
    lb = snkpt.light_block(dir_colour=(0.6, 0.6, 0.5), dir_to_light=(0.3, 0.9, 0.4), points=_point_lights(24, 3)) \
        if lights == "points+dir" else None
    ctx = make_ctx(scene)
    arrs = O.scene_arrays(scene)
    ctx.set_lights(lb, arrs[6])
    ctx.build_accel()
    w, h = 64, 48
    cam = scene.camera(0, w, h)
    for f in range(4):
        got = _frame(ctx, cam, w, h, snkpt.MODE_REFERENCE_COMPAT, 17, frame=f, flags=snkpt.FLAG_NO_BOUNCE,
                     num_temporal_resamples=1, num_spatial_resamples=3)
    want = O.compat_direct(arrs, O.Bvh(O.flatten(*arrs[:4])), cam, w, h, 4, lb, 1, 3, 30.0, 17).reshape(-1, 3)
    ok = (np.abs(got - want) <= 1e-3 * np.abs(want) + 1e-5).all(1)
    _say(f"compat direct ({lights}): {ok.mean():.4f} of pixels agree, mean {got.mean():.5f} vs {want.mean():.5f}")
    assert np.isfinite(got).all() and want.mean() > 0.0
    assert ok.mean() >= 0.97
    assert abs(got.mean() - want.mean()) <= 0.01 * want.mean()


def _array_ctx(sc, lights=None):
    import snkpt

    def make():
        ctx = snkpt.Context(0)
        sc.load_into(ctx, lights)
        return ctx
    return make


def test_compat_bounce_matches_the_oracle_per_pixel_on_the_room(gpu):
    import snkpt
    _say("compat: room 64x64")
    scene = snkpt.SceneData(1)
    w, h = 64, 64
    cam = scene.camera(0, w, h)
    got = _bounce_gpu(lambda: make_ctx(scene), cam, w, h, seed=11)
    arrs = O.scene_arrays(scene)
    from helpers import scene_world
    b = O.Bvh(scene_world(scene))
    want = O.compat_bounce(arrs, b, cam, w, h, 11).reshape(-1, 3).astype(np.float64)
    assert want.max() > 0.0
    err = np.abs(got - want).max(axis=1)
    tol = 1e-4 * (1.0 + np.abs(want).max(axis=1))
    ok = err <= tol
    _say(f"compat room: {ok.mean():.4f} of pixels agree, mean {got.mean():.5f} vs {want.mean():.5f}")
    # sinf / cosf of the two libraries may differ in the last ulp, which can
    # move a grazing ray onto another triangle; the rest agree to float rounding
    assert ok.mean() >= 0.97
    assert abs(got.mean() - want.mean()) <= 0.01 * want.mean()


def test_compat_closed_form_all_emitter_hemisphere(gpu):
    a, le = np.array([0.8, 0.6, 0.4]), np.array([1.0, 0.5, 0.25])
    S, Hc = 5.0, 5.0
    sc = ArrayScene()
    sc.quad([-S, 0, -S], [0, 0, 2 * S], [2 * S, 0, 0], a, 2)
    sc.quad([-S, Hc, -S], [2 * S, 0, 0], [0, 0, 2 * S], le, 1)
    for o, u, v in (([-S, 0, -S], [2 * S, 0, 0], [0, Hc, 0]), ([-S, 0, S], [0, Hc, 0], [2 * S, 0, 0]),
                    ([-S, 0, -S], [0, Hc, 0], [0, 0, 2 * S]), ([S, 0, -S], [0, 0, 2 * S], [0, Hc, 0])):
        sc.quad(o, u, v, le, 1)
    w = h = 48
    cam = look_at_camera([0.0, 1.0, 0.0], [0.001, 0.0, 0.0], fov=20.0, w=w, h=h)
    got = _bounce_gpu(_array_ctx(sc), cam, w, h).mean(axis=0)
    want = 2.0 * REF_PI * a * le * math.sin(1.0)
    _say(f"compat closed form: {got} vs {want}")
    assert np.allclose(got, want, rtol=5e-3)


def test_compat_closed_form_directional_light_payload(gpu):
    import snkpt
    a, ac, col = np.array([0.9, 0.6, 0.3]), np.array([0.5, 0.4, 0.8]), np.array([1.5, 1.0, 2.0])
    S = 1000.0
    sc = ArrayScene()
    sc.quad([-S, 0, -S], [0, 0, 2 * S], [2 * S, 0, 0], a, 2)
    sc.quad([-S, 1.0, -S], [0, 0, 2 * S], [2 * S, 0, 0], ac, 2)
    lights = snkpt.light_block(dir_colour=tuple(col), dir_to_light=(0.0, 1.0, 0.0))
    w = h = 48
    cam = look_at_camera([0.0, 0.5, 0.0], [0.001, 0.0, 0.0], fov=20.0, w=w, h=h)
    got = _bounce_gpu(_array_ctx(sc, lights), cam, w, h).mean(axis=0)
    want = 2.0 * a * ac * col * math.sin(1.0)
    _say(f"compat directional payload: {got} vs {want}")
    assert np.allclose(got, want, rtol=5e-3)


def test_compat_accumulates_and_rejects_unknown_modes(gpu):
    import torch

    import snkpt
    scene = snkpt.SceneData(1)
    ctx = make_ctx(scene)
    w, h = 32, 32
    cam = scene.camera(0, w, h)
    out = torch.zeros(h, w, 4, device="cuda")
    frames = []
    for f in range(3):
        fl = snkpt.FLAG_ACCUMULATE | (snkpt.FLAG_RESET if f == 0 else 0)
        ctx.render(cam, snkpt.default_settings(mode=snkpt.MODE_REFERENCE_COMPAT, seed=3, frame=f, flags=fl), out, w, h)
        torch.cuda.synchronize()
        frames.append(out[..., :3].cpu().numpy().copy())
    assert np.all(np.isfinite(frames[-1])) and frames[-1].mean() > 0.0
    with pytest.raises(snkpt.SnkptError):
        ctx.render(cam, snkpt.default_settings(mode=4), out, w, h)


def test_context_before_the_first_torch_cuda_call(gpu):
    """A fresh process that creates a snkpt context before touching
    torch.cuda can still use torch's GPU tensors afterwards."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "import torch, snkpt\n"
            "ctx = snkpt.Context(0)\n"
            "x = torch.ones(4, device='cuda')\n"
            "print('ok', float(x.sum()))\n") % (root, os.path.join(root, "snake-vk_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok 4.0" in r.stdout, r.stdout + r.stderr
