"""ReSTIR DI parity: the reused estimator equals the path tracer in
expectation (unbiased 1/Z combination) and has lower error per frame."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(config=3):
    import snkpt
    scene = snkpt.SceneData(config)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    return scene, ctx


def _oracle(ctx, cam, w, h, spp, depth):
    import torch

    import snkpt
    s = snkpt.default_settings(max_depth=depth, seed=1234)
    acc = torch.zeros(h * w * 3, dtype=torch.float64, device="cuda")
    for k in range(spp // 256):
        s.frame = k
        ctx.oracle_accumulate(cam, s, acc, w, h, 256)
    torch.cuda.synchronize()
    return (acc / spp).view(h, w, 3).cpu().numpy()


def _frames(ctx, cam, w, h, nframes, **kw):
    import torch

    import snkpt
    out = torch.zeros(h, w, 4, device="cuda")
    imgs = []
    for f in range(nframes):
        s = snkpt.default_settings(frame=f, **kw)
        ctx.render(cam, s, out, w, h)
        imgs.append(out[..., :3].cpu().numpy().copy())
    return np.stack(imgs)


@pytest.mark.parametrize("temporal,spatial", [(0, 0), (1, 0), (1, 5)])
def test_restir_equals_pt_in_expectation(gpu, temporal, spatial):
    import snkpt
    scene, ctx = _setup(3)
    w = h = 48
    cam = scene.camera(0, w, h)
    ref = _oracle(ctx, cam, w, h, 2048, 1)
    imgs = _frames(ctx, cam, w, h, 96, mode=snkpt.MODE_RESTIR, max_depth=1, num_temporal_resamples=temporal,
                   num_spatial_resamples=spatial, seed=5)
    mean = imgs.mean(0)
    tr = mean.reshape(6, 8, 6, 8, 3).mean((1, 3))
    to = ref.reshape(6, 8, 6, 8, 3).mean((1, 3))
    rel = np.abs(tr - to).sum() / to.sum()
    _say(f"temporal={temporal} spatial={spatial}: image mean {mean.mean():.5f} vs {ref.mean():.5f}, tile L1 rel {rel:.4f}")
    assert abs(mean.mean() - ref.mean()) < 0.02 * ref.mean()
    assert rel < 0.05


def test_restir_lowers_per_frame_error(gpu):
    """Many-light scene: one ReSTIR frame (after temporal warm-up) has lower
    RMSE against the converged image than one path-traced frame."""
    import snkpt
    scene, ctx = _setup(3)
    w = h = 64
    cam = scene.camera(0, w, h)
    ref = _oracle(ctx, cam, w, h, 1024, 1)
    pt = _frames(ctx, cam, w, h, 4, mode=snkpt.MODE_PT, max_depth=1, seed=9)
    rs = _frames(ctx, cam, w, h, 12, mode=snkpt.MODE_RESTIR, max_depth=1, num_spatial_resamples=5, seed=9)
    e_pt = np.sqrt(((pt - ref) ** 2).mean())
    e_rs = np.sqrt(((rs[8:] - ref) ** 2).mean())
    _say(f"RMSE one frame: PT {e_pt:.4f}, ReSTIR {e_rs:.4f}")
    assert e_rs < 0.7 * e_pt


def test_restir_with_indirect_bounces(gpu):
    """ReSTIR at the primary vertex + path-traced bounces equals full PT in expectation."""
    import snkpt
    scene, ctx = _setup(1)
    w = h = 32
    cam = scene.camera(0, w, h)
    ref = _oracle(ctx, cam, w, h, 2048, 4)
    imgs = _frames(ctx, cam, w, h, 128, mode=snkpt.MODE_RESTIR, max_depth=4, num_spatial_resamples=3, seed=3)
    mean = imgs.mean(0)
    _say(f"restir+pt mean {mean.mean():.5f} vs {ref.mean():.5f}")
    assert abs(mean.mean() - ref.mean()) < 0.02 * ref.mean()
    tr = mean.reshape(4, 8, 4, 8, 3).mean((1, 3))
    to = ref.reshape(4, 8, 4, 8, 3).mean((1, 3))
    assert np.abs(tr - to).sum() / to.sum() < 0.05
