"""Analytic lights of the light block (directional, point, spot;
LightBuffers.glsl CalcDirectionalLight / CalcPointlight / CalcSpotlight):
a Lambertian floor lit by one such light has the closed-form radiance
albedo/pi * colour * cos / attenuation (times the spot falloff), pixel by pixel;
occluders cast exact shadows; with an emissive quad as well, NEE picks one of
the two light sets per vertex and the image equals the sum of both closed forms
in expectation, in PT and in ReSTIR mode."""
import sys

import numpy as np
import pytest

from helpers import ArrayScene, floor_light_scene, look_at_camera, polygon_irradiance

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _floor_only(albedo=0.5):
    sc = ArrayScene()
    sc.quad([-10, 0, -10], [0, 0, 20], [20, 0, 0], [albedo] * 3, 2)
    return sc


def _view(w=16, h=16):
    cam = look_at_camera([0.2, 1.5, 0.1], [0.3, 0.0, -0.2], fov=40.0, w=w, h=h)
    import raysets
    org, dr = raysets.primary_rays(cam, w, h)
    t = -org[:, 1] / dr[:, 1]
    xs = (org[:, :3] + dr[:, :3] * t[:, None]).astype(np.float64)
    return cam, xs


def _render(ctx, cam, w, h, **kw):
    import torch

    import snkpt
    out = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, snkpt.default_settings(**kw), out, w, h)
    torch.cuda.synchronize()
    return out[..., :3].cpu().numpy().reshape(-1, 3)


def _point_closed(xs, P, I, att, albedo=0.5):
    d = P[None, :] - xs
    dist = np.linalg.norm(d, axis=1)
    cos = np.maximum(d[:, 1] / dist, 0.0)
    a = att[2] + att[1] * dist + att[0] * dist ** 2  # (exp, linear, constant)
    return albedo / np.pi * cos[:, None] * np.asarray(I)[None, :] / a[:, None]


def test_point_light_closed_form_per_pixel(gpu):
    import snkpt
    P, I, att = np.array([0.4, 1.2, -0.1]), (3.0, 2.0, 1.0), (0.5, 0.3, 1.0)
    ctx = snkpt.Context(0)
    _floor_only().load_into(ctx, snkpt.light_block(points=[(I, tuple(P), att)]))
    cam, xs = _view()
    img = _render(ctx, cam, 16, 16, max_depth=1)
    want = _point_closed(xs, P, I, att)
    _say(f"point light: max rel err {np.max(np.abs(img - want) / want):.2e}")
    assert np.allclose(img, want, rtol=2e-4, atol=1e-7)


def test_directional_light_closed_form_per_pixel(gpu):
    import snkpt
    L = np.array([0.3, 1.0, 0.2])
    ctx = snkpt.Context(0)
    _floor_only().load_into(ctx, snkpt.light_block(dir_colour=(1.5, 1.0, 0.5), dir_to_light=tuple(L)))
    cam, xs = _view()
    img = _render(ctx, cam, 16, 16, max_depth=2)
    want = 0.5 / np.pi * (L[1] / np.linalg.norm(L)) * np.array([1.5, 1.0, 0.5])
    assert np.allclose(img, want[None, :], rtol=2e-4)


def test_spot_light_cone_and_falloff(gpu):
    import snkpt
    P, D, I, att, ap = np.array([0.3, 1.0, -0.2]), np.array([0.0, -1.0, 0.0]), (4.0, 4.0, 4.0), (0.0, 0.0, 1.0), 0.985
    ctx = snkpt.Context(0)
    _floor_only().load_into(ctx, snkpt.light_block(spots=[(I, tuple(P), tuple(D), att, ap)]))
    cam, xs = _view(32, 32)
    img = _render(ctx, cam, 32, 32, max_depth=1)
    d = P[None, :] - xs
    dist = np.linalg.norm(d, axis=1)
    wi = d / dist[:, None]
    sf = -(wi @ D)
    inside = (sf >= 1e-4) & (sf >= ap)
    fall = np.where(inside, 1.0 - (1.0 - sf) / max(1.0 - ap, 1e-5), 0.0)
    want = _point_closed(xs, P, I, att) * fall[:, None]
    edge = np.abs(sf - ap) < 1e-4  # pixels on the cone's edge may round either way
    assert inside.sum() > 50 and (~inside).sum() > 50
    # the falloff divides by 1 - aperture = 0.015: FP32 rounding of the cosine
    # (~1e-7) becomes ~1e-5 of the falloff, absolute
    err = np.abs(img - want)[~edge]
    _say(f"spot light: {inside.sum()} pixels inside, max abs err {err.max():.2e} (max radiance {want.max():.3f})")
    assert np.all(img[~inside & ~edge] == 0.0)
    assert np.allclose(img[~edge], want[~edge], rtol=2e-4, atol=2e-5 * want.max())


def test_point_light_shadow(gpu):
    import snkpt
    P, I, att = np.array([0.3, 2.0, -0.2]), (2.0, 2.0, 2.0), (0.0, 0.0, 1.0)
    sc = _floor_only()
    # a 0.2 x 0.2 occluder halfway up: its shadow is the 0.4 x 0.4 square under it
    sc.quad([0.2, 1.0, -0.3], [0.2, 0, 0], [0, 0, 0.2], [0.5] * 3, 2)
    ctx = snkpt.Context(0)
    sc.load_into(ctx, snkpt.light_block(points=[(I, tuple(P), att)]))
    cam, xs = _view(32, 32)
    img = _render(ctx, cam, 32, 32, max_depth=1)
    want = _point_closed(xs, P, I, att)
    # pixels whose camera ray meets the occluder (y = 1) before the floor see the occluder
    import raysets
    org, dr = raysets.primary_rays(cam, 32, 32)
    t1 = (1.0 - org[:, 1]) / dr[:, 1]
    y1 = org[:, :3] + dr[:, :3] * t1[:, None]
    on_occ = (np.abs(y1[:, 0] - 0.3) <= 0.101) & (np.abs(y1[:, 2] + 0.2) <= 0.101)
    dx, dz = np.abs(xs[:, 0] - 0.3), np.abs(xs[:, 2] + 0.2)
    shadow = (dx < 0.19) & (dz < 0.19) & ~on_occ
    lit = ((dx > 0.21) | (dz > 0.21)) & ~on_occ
    assert shadow.sum() > 20 and lit.sum() > 20
    assert np.all(img[shadow] == 0.0)
    assert np.allclose(img[lit], want[lit], rtol=2e-4)


@pytest.mark.parametrize("mode", ["pt", "restir"])
def test_point_light_with_emissive_quad_in_expectation(gpu, mode):
    """Both light sets: per vertex NEE picks one with probability 1/2 (PT) or
    ReSTIR resamples the quad and a second shadow ray takes the point light;
    the mean image is the sum of the two closed forms."""
    import torch

    import snkpt
    sc, verts = floor_light_scene()
    P, I, att = np.array([-0.6, 1.0, 0.3]), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0)
    ctx = snkpt.Context(0)
    sc.load_into(ctx, snkpt.light_block(points=[(I, tuple(P), att)]))
    w = h = 8
    cam = look_at_camera([0.3, 1.0, 0.2], [0.3, 0.0, -0.2], fov=4.0, w=w, h=h)
    import raysets
    org, dr = raysets.primary_rays(cam, w, h)
    t = -org[:, 1] / dr[:, 1]
    xs = (org[:, :3] + dr[:, :3] * t[:, None]).astype(np.float64)
    want = np.array([0.5 / np.pi * polygon_irradiance(x, [0, 1, 0], verts, 10.0) for x in xs])[:, None] + \
        _point_closed(xs, P, I, att)
    m = snkpt.MODE_PT if mode == "pt" else snkpt.MODE_RESTIR
    acc = np.zeros((w * h, 3))
    frames = 400
    for f in range(frames):
        acc += _render(ctx, cam, w, h, mode=m, max_depth=1, frame=f, seed=3, spp=4 if mode == "pt" else 1)
    mean = acc / frames
    _say(f"{mode}: mean {mean.mean():.5f} vs closed form {want.mean():.5f}")
    assert abs(mean.mean() - want.mean()) < 0.01 * want.mean()
    assert np.allclose(mean, want, rtol=0.08)


def test_light_block_validation(gpu):
    import snkpt
    ctx = snkpt.Context(0)
    blk = bytearray(snkpt.light_block(spots=[((1, 1, 1), (0, 1, 0), (0, -1, 0), (0, 0, 1), 0.5)]))
    with pytest.raises(snkpt.SnkptError):
        ctx.set_lights(bytes(blk[:49300]))  # one spot light ends at 112 + 48 * 1024 + 64 = 49328 B
    ctx.set_lights(bytes(blk))


def test_many_point_lights_restir_reservoirs(gpu):
    """64 point lights of random colour, position and falloff over the floor:
    ReSTIR's analytic-light reservoirs (32 candidates, temporal and spatial
    reuse, the reference's point-light reservoir) give the sum of the 64 closed
    forms in expectation, with a lower per-frame error than plain PT's one
    uniformly chosen light per vertex."""
    import snkpt
    rng = np.random.default_rng(12)
    pts = []
    for _ in range(64):
        P = np.array([rng.uniform(-3, 3), rng.uniform(0.3, 2.0), rng.uniform(-3, 3)])
        I = tuple(rng.uniform(0.05, 1.5, 3))
        att = (rng.uniform(0.2, 1.0), rng.uniform(0.0, 0.5), rng.uniform(0.5, 1.0))
        pts.append((I, tuple(P), att))
    sc = _floor_only()
    ctx = snkpt.Context(0)
    sc.load_into(ctx, snkpt.light_block(points=pts))
    w = h = 16
    cam, xs = _view(w, h)
    want = sum(_point_closed(xs, np.array(P), I, att) for I, P, att in pts)
    tiles = lambda a: a.reshape(4, 4, 4, 4, 3).mean((1, 3))  # noqa: E731
    pt_ctx = snkpt.Context(0)
    sc.load_into(pt_ctx, snkpt.light_block(points=pts))
    e_pt = [np.sqrt(((_render(pt_ctx, cam, w, h, mode=snkpt.MODE_PT, max_depth=1, frame=f, seed=9) - want) ** 2)
                    .mean()) for f in range(20)]
    # temporal reuse correlates the frames (M cap 20), so its mean is checked
    # more loosely than that of the spatial-only estimator
    for temporal, frames, tol in ((0, 400, 0.015), (1, 400, 0.03)):
        c = snkpt.Context(0)
        sc.load_into(c, snkpt.light_block(points=pts))
        acc = np.zeros((w * h, 3))
        e_rs = []
        for f in range(frames):
            img = _render(c, cam, w, h, mode=snkpt.MODE_RESTIR, max_depth=1, frame=f, seed=5,
                          num_temporal_resamples=temporal, num_spatial_resamples=5)
            acc += img
            if f >= 10:
                e_rs.append(np.sqrt(((img - want) ** 2).mean()))
        mean = acc / frames
        rel = np.abs(tiles(mean) - tiles(want)).sum() / tiles(want).sum()
        _say(f"64 point lights, temporal={temporal}: ReSTIR mean {mean.mean():.5f} vs closed form "
             f"{want.mean():.5f}, tile L1 rel {rel:.4f}; per-frame RMSE ReSTIR {np.mean(e_rs):.5f}, "
             f"PT {np.mean(e_pt):.5f}")
        assert abs(mean.mean() - want.mean()) < tol * want.mean()
        assert rel < 4 * tol
        assert np.mean(e_rs) < 0.6 * np.mean(e_pt)
