"""TAA resolve (snkpt_taa_resolve; TAA_Resolve.comp:37-86) on the GPU against
the numpy restatement, over several frames so the history and previous
velocity the context keeps are exercised; in place == out of place; reset;
RGBA16F; argument checks."""
import sys

import numpy as np
import pytest

from oracle import taa

pytestmark = pytest.mark.gpu
F = np.float32


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _frames(h, w, n, seed):
    rng = np.random.default_rng(seed)
    frames, vels = [], []
    for _ in range(n):
        img = (rng.random((h, w, 4)) ** 3 * 4.0).astype(F)  # HDR-ish, heavy tail
        img[rng.random((h, w)) < 0.02, :3] *= -1.0            # a few negative texels (clamped by the resolve)
        v = ((rng.random((h, w, 2)) - 0.5) * 0.02).astype(np.float16)
        v[rng.random((h, w)) < 0.05] = np.float16(0.15)        # disoccluded texels
        frames.append(img)
        vels.append(v)
    return frames, vels


def _run(ctx, frames, vels, fmt, in_place=False):
    import torch

    import snkpt
    outs = []
    for k, (img, v) in enumerate(zip(frames, vels)):
        if fmt == snkpt.FORMAT_RGBA16F:
            cur = torch.from_numpy(img.astype(np.float16)).cuda()
        else:
            cur = torch.from_numpy(img).cuda()
        out = cur if in_place else torch.empty_like(cur)
        ctx.taa_resolve(cur, out, torch.from_numpy(v).cuda(), img.shape[1], img.shape[0], fmt, reset=(k == 0))
        torch.cuda.synchronize()
        outs.append(out.float().cpu().numpy())
    return outs


def _oracle(frames, vels, half=False):
    q = (lambda a: a.astype(np.float16).astype(F)) if half else (lambda a: a)
    outs = [q(frames[0])]
    hist, pvel = q(frames[0]), vels[0].astype(F)
    for img, v in zip(frames[1:], vels[1:]):
        o = q(taa.resolve(q(img), v.astype(F), hist, pvel))
        outs.append(o)
        hist, pvel = o, v.astype(F)
    return outs


@pytest.mark.parametrize("w,h", [(64, 48), (37, 21)])
def test_taa_matches_numpy_restatement(gpu, w, h):
    import snkpt
    frames, vels = _frames(h, w, 4, seed=w)
    got = _run(snkpt.Context(0), frames, vels, snkpt.FORMAT_RGBA32F)
    want = _oracle(frames, vels)
    assert np.array_equal(got[0], frames[0])  # reset frame: out = current
    for k in range(1, 4):
        err = np.abs(got[k] - want[k]) / (np.abs(want[k]) + 1e-3)
        _say(f"frame {k}: max rel err {err.max():.2e}")
        assert np.allclose(got[k], want[k], rtol=1e-5, atol=1e-6), (k, err.max())


def test_taa_rgba16f(gpu):
    import snkpt
    frames, vels = _frames(40, 56, 3, seed=9)
    got = _run(snkpt.Context(0), frames, vels, snkpt.FORMAT_RGBA16F)
    want = _oracle(frames, vels, half=True)
    for k in range(3):
        assert np.allclose(got[k], want[k], rtol=2e-3, atol=1e-4), k  # one half ulp


def test_taa_in_place_equals_out_of_place(gpu):
    import snkpt
    frames, vels = _frames(33, 70, 3, seed=4)
    a = _run(snkpt.Context(0), frames, vels, snkpt.FORMAT_RGBA32F)
    b = _run(snkpt.Context(0), frames, vels, snkpt.FORMAT_RGBA32F, in_place=True)
    for k in range(3):
        assert np.array_equal(a[k], b[k]), k


def test_taa_after_render_converges_toward_the_mean(gpu):
    """A static view: resolving successive 1-spp frames lowers the error
    against the accumulated mean (the resolve is an exponential average)."""
    import torch

    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w = h = 64
    cam = scene.camera(0, w, h)
    vel = torch.zeros(h, w, 2, dtype=torch.float16, device="cuda")
    frame = torch.zeros(h, w, 4, device="cuda")
    res = torch.zeros(h, w, 4, device="cuda")
    acc = torch.zeros(h, w, 4, device="cuda")
    for f in range(64):
        ctx.render(cam, snkpt.default_settings(max_depth=3, frame=f, seed=2), frame, w, h)
        acc += frame
        ctx.taa_resolve(frame, res, vel, w, h, reset=(f == 0))
    torch.cuda.synchronize()
    mean = (acc / 64)[..., :3]
    # away from the emitter: the +-0.55 sigma box of a texel next to radiance 15
    # pulls its history toward the neighbourhood mean (the clip's edge bias)
    hot = torch.nn.functional.max_pool2d(mean.amax(-1)[None, None], 5, stride=1, padding=2)[0, 0] > 1.0
    keep = ~hot
    e_raw = float(((frame[..., :3] - mean) ** 2)[keep].mean().sqrt())
    e_taa = float(((res[..., :3] - mean) ** 2)[keep].mean().sqrt())
    _say(f"RMSE vs 64-frame mean ({int(keep.sum())} texels): raw frame {e_raw:.4f}, resolved {e_taa:.4f}")
    assert e_taa < 0.6 * e_raw


def test_taa_rejects_bad_input(gpu):
    import torch

    import snkpt
    ctx = snkpt.Context(0)
    img = torch.zeros(8, 8, 4, device="cuda")
    vel = torch.zeros(8, 8, 2, dtype=torch.float16, device="cuda")
    with pytest.raises(snkpt.SnkptError):
        ctx.taa_resolve(img, img, None, 8, 8)
    raw = torch.zeros(8 * 8 * 16 + 4, dtype=torch.uint8, device="cuda")
    with pytest.raises(snkpt.SnkptError):
        ctx.taa_resolve(raw[4:], img, vel, 8, 8)  # not 16-byte aligned
    with pytest.raises(snkpt.SnkptError):
        ctx.taa_resolve(img, img, vel, 8, 8, fmt=7)
