"""ReSTIR DI parity: the reused estimator equals the path tracer in
expectation (unbiased 1/Z combination) and has lower error per frame."""
import sys

import numpy as np
import pytest

from oracle import binding as O

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


@pytest.mark.parametrize("temporal,spatial", [(0, 0), (1, 0), (1, 5), (1, 8)])
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


def _predicted_reuse(bvh, arrs, cam_prev, cam, w, h, mirror=False):
    """Pixels of `cam` whose primary hit reprojects (through the previous
    camera's proj_view, as the temporal pass does) onto a previous pixel that
    saw a non-emissive surface with a normal within 0.9 and a distance within
    5 %: the history the temporal pass must accept, predicted on the CPU from
    the oracle's traversal and surfaces.  mirror=True flips the reprojected
    row, the mistake the count must be able to tell apart."""
    import raysets

    def surf(c):
        org, dr = raysets.primary_rays(c, w, h)
        t = bvh.trace(np.concatenate([org, dr], 1))[0]
        _, n3, _, _, fl = O.primary_surfaces(arrs, bvh, c, w, h)
        valid = (t >= 0) & ((fl.reshape(-1) & 1) == 0)
        p = org[:, :3].astype(np.float64) + dr[:, :3].astype(np.float64) * t[:, None]
        return valid, p, n3.reshape(-1, 3), t
    v1, p1, n1, _ = surf(cam)
    v0, _, n0, t0 = surf(cam_prev)
    pv = np.array(cam_prev.proj_view, np.float64).reshape(4, 4).T
    cl = np.concatenate([p1, np.ones((p1.shape[0], 1))], 1) @ pv.T
    ok = v1 & (cl[:, 3] > 1e-6)
    fx = (cl[:, 0] / cl[:, 3] * 0.5 + 0.5) * w
    fy = (cl[:, 1] / cl[:, 3] * 0.5 + 0.5) * h
    if mirror:
        fy = h - fy
    ok &= (fx >= 0) & (fy >= 0) & (fx < w) & (fy < h)
    q = np.where(ok, np.floor(fy).astype(np.int64) * w + np.floor(fx).astype(np.int64), 0)
    pos0 = np.array(cam_prev.cam_pos[:3], np.float64)
    dprev = np.linalg.norm(p1 - pos0, axis=1)
    ok &= v0[q] & ((n0[q] * n1).sum(1) >= 0.9) & (np.abs(dprev - t0[q]) <= 0.05 * t0[q])
    return int(ok.sum()), int(v1.sum())


def test_restir_orbit_reuses_reprojected_history(gpu):
    """Config 3's seeded 64-frame camera orbit (an arc in front of the room's
    open side, every frame looking into the lit room): every frame's temporal
    pass accepts the history of the pixels the previous camera saw (count
    within 0.5 % of the CPU prediction, which a mirrored reprojection misses by
    far), and the frames stay equal to the converged image of their own camera
    in expectation (summed over 47 frames of the orbit) with less error than
    plain PT.  Single frames' sums scatter by ~3.5 % (pixels next to the 10k
    emitters are heavy-tailed, in the oracle too), hence 47 frames and a
    4096-spp oracle."""
    import torch

    import snkpt
    scene, ctx = _setup(3)
    _, ref_ctx = _setup(3)  # references on a context of their own: ctx keeps its ReSTIR history
    arrs = O.scene_arrays(scene)
    bvh = O.Bvh(O.flatten(*arrs[:4]))
    w, h = 64, 48
    nf = 48
    out = torch.zeros(h, w, 4, device="cuda")
    num = den = 0.0
    tiles_d = np.zeros((6, 8, 3))
    tiles_o = np.zeros((6, 8, 3))
    e_rs, e_pt = [], []
    miss_mirror = n_valid = 0
    for f in range(nf):
        cam = scene.camera(f, w, h)
        s = snkpt.default_settings(mode=snkpt.MODE_RESTIR, max_depth=1, num_temporal_resamples=1,
                                   num_spatial_resamples=5, seed=21, frame=f)
        ctx.render(cam, s, out, w, h)
        img = out[..., :3].cpu().numpy().copy()
        reused = ctx.frame_stats().restir_history_reused
        if f == 0:
            assert reused == 0
            continue
        want, valid = _predicted_reuse(bvh, arrs, scene.camera(f - 1, w, h), cam, w, h)
        wrong, _ = _predicted_reuse(bvh, arrs, scene.camera(f - 1, w, h), cam, w, h, mirror=True)
        _say(f"orbit frame {f}: reused {reused}, predicted {want} (mirrored {wrong}) of {valid} surface pixels")
        assert want > 0.5 * valid
        assert abs(reused - want) <= 0.005 * valid + 2, (f, reused, want)
        miss_mirror += want - wrong
        n_valid += valid
        ref = _oracle(ref_ctx, cam, w, h, 4096, 1)
        pt = _frames(ref_ctx, cam, w, h, 1, mode=snkpt.MODE_PT, max_depth=1, seed=77 + f)[0]
        num += img.sum()
        den += ref.sum()
        tiles_d += (img - ref).reshape(6, 8, 8, 8, 3).sum((1, 3))
        tiles_o += ref.reshape(6, 8, 8, 8, 3).sum((1, 3))
        if f >= 4:
            e_rs.append(np.sqrt(((img - ref) ** 2).mean()))
            e_pt.append(np.sqrt(((pt - ref) ** 2).mean()))
    # the count tells a mirrored reprojection apart (some orbit views are nearly symmetric)
    assert miss_mirror > 0.15 * n_valid, (miss_mirror, n_valid)
    rel = np.abs(tiles_d).sum() / tiles_o.sum()
    _say(f"orbit: sum ratio {num / den:.4f}, tile L1 rel {rel:.4f}, RMSE ReSTIR {np.mean(e_rs):.4f} PT {np.mean(e_pt):.4f}")
    assert abs(num / den - 1.0) < 0.03
    assert rel < 0.08
    assert np.mean(e_rs) < 0.8 * np.mean(e_pt)
