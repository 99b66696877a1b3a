"""GPU parity of the GGX / Cook-Torrance shading (LightBuffers.glsl:65-126) and
of Russian roulette: per pixel against the CPU oracle path tracer on the same
PCG streams, and against a quadrature of f cos L over a square emitter."""
import sys

import numpy as np
import pytest

from helpers import ArrayScene, ggx_room, look_at_camera
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _render(ctx, cam, w, h, **kw):
    import torch

    import snkpt
    out = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, snkpt.default_settings(**kw), out, w, h)
    torch.cuda.synchronize()
    return out[..., :3].cpu().numpy()


def _agree(g, c):
    return (np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-5).all(-1).mean()


@pytest.mark.parametrize("rr_start", [0, 3])
def test_ggx_room_matches_cpu_oracle_per_pixel(gpu, rr_start):
    """Config 2's materials (GGX, metallic 0/1, roughness 0.05-1) in the room,
    depth 8, with and without Russian roulette from bounce 3."""
    import snkpt
    sc, scene = ggx_room(5)
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w = h = 64
    cam = scene.camera(0, w, h)
    _say(f"ggx room rr_start={rr_start}")
    g = _render(ctx, cam, w, h, spp=4, max_depth=8, rr_start=rr_start, seed=7)
    c = O.pt_render(sc.arrays(), O.Bvh(sc.world()), cam, w, h, 4, 8, rr_start, 7)
    frac = _agree(g, c)
    _say(f"pixels agreeing: {frac:.4f}; means {g.mean():.6f} vs {c.mean():.6f}")
    assert np.isfinite(g).all() and g.mean() > 0.01
    assert frac >= 0.97, frac
    assert abs(g.mean() - c.mean()) <= 0.01 * c.mean()


def test_c2_frame_matches_cpu_oracle_per_pixel(gpu):
    """Config 2 itself (1.02M triangles, its GGX materials, depth 8, RR from
    bounce 3) at 128x72, per pixel against the oracle."""
    import snkpt
    scene = snkpt.SceneData(2)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w, h = 128, 72
    cam = scene.camera(0, w, h)
    _say("c2 128x72 vs oracle")
    g = _render(ctx, cam, w, h, spp=2, max_depth=scene.s.max_depth, rr_start=scene.s.rr_start, seed=3)
    arrs = O.scene_arrays(scene)
    b = O.Bvh(O.flatten(*arrs[:4]))
    c = O.pt_render(arrs, b, cam, w, h, 2, scene.s.max_depth, scene.s.rr_start, 3)
    frac = _agree(g, c)
    _say(f"pixels agreeing: {frac:.4f}; means {g.mean():.6f} vs {c.mean():.6f}")
    assert scene.s.rr_start == 3 and scene.s.max_depth == 8
    assert frac >= 0.97, frac
    assert abs(g.mean() - c.mean()) <= 0.01 * c.mean()


def _ggx_floor(rough, metal, L=10.0, side=1.0, height=2.0):
    sc = ArrayScene()
    sc.quad([-10, 0, -10], [0, 0, 20], [20, 0, 0], [0.7, 0.5, 0.3], 0, roughness=rough, metallic=metal)
    hs = side / 2
    sc.quad([-hs, height, -hs], [side, 0, 0], [0, 0, side], [L] * 3, 1)
    return sc


def _emitter_quadrature(x, wo, rough, metal, L=10.0, side=1.0, height=2.0, n=96):
    """f(wo, wi) L cos_x cos_l / r^2 over the emitter square by Gauss-Legendre
    quadrature, f from the oracle's BRDF (pinned by its own closed forms)."""
    gx, gw = np.polynomial.legendre.leggauss(n)
    xs = gx * side / 2
    ws = gw * side / 2
    tot = np.zeros(3)
    nf = np.array([0, 1.0, 0])
    for i in range(n):
        for k in range(n):
            q = np.array([xs[i], height, xs[k]])
            d = q - x
            r2 = d @ d
            wi = d / np.sqrt(r2)
            f, _ = O.brdf_eval([0.7, 0.5, 0.3], rough, metal, 0, nf, wo, wi)
            tot += ws[i] * ws[k] * f * L * wi[1] * wi[1] / r2
    return tot


@pytest.mark.parametrize("rough,metal", [(0.3, 1.0), (0.5, 0.0), (0.8, 1.0)])
def test_ggx_floor_direct_light_quadrature(gpu, rough, metal):
    """A GGX floor under a square emitter: the radiance towards the eye equals
    the integral of f cos L over the emitter, by NEE alone (depth 1) and by
    NEE + BRDF sampling under the power heuristic (depth 2)."""
    import snkpt
    sc = _ggx_floor(rough, metal)
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    eye = np.array([1.6, 1.2, 0.4])
    for x0 in ([-0.6, 0.0, 0.1], [0.2, 0.0, -0.3]):
        x0 = np.array(x0)
        wo = (eye - x0) / np.linalg.norm(eye - x0)
        want = _emitter_quadrature(x0, wo, rough, metal)
        cam = look_at_camera(eye, x0, fov=0.5, w=4, h=4)
        for depth in (1, 2):
            img = _render(ctx, cam, 4, 4, spp=20000, max_depth=depth, seed=13 + depth).reshape(-1, 3).mean(0)
            _say(f"rough={rough} metal={metal} x={x0} depth={depth}: {img} vs {want}")
            assert np.allclose(img, want, rtol=0.02, atol=1e-4), (rough, metal, depth, img, want)
