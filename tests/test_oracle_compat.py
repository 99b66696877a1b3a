"""The oracle's restatement of the reference's BounceDiffuse estimator
(RayGen.rgen:123-213 with the payload of RayClosestHit.rchit:87-142), the
term SNKPT_MODE_REFERENCE_COMPAT adds to the direct light, pinned by closed
forms.  Every outer ray leaves along rd = UniformSampleHemisphere(r1, r2), so
dot(rd, n) = r1 is uniform on [0, 1) and E[max(cos(dot(rd, n)), 0)] = sin(1):
the reference's cos-of-a-cosine weight makes sin(1) the signature of each
closed form below (the cosine itself would give 1/2)."""
import math

import numpy as np

from helpers import ArrayScene, look_at_camera
from oracle import binding as O

REF_PI = np.float32(3.1415926)
W = H = 48


def _bounce(sc, cam, seed=5, dir_lights=None, frames=(0,)):
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    out = [O.compat_bounce(arrs, b, cam, W, H, seed, frame=f, dir_lights=dir_lights) for f in frames]
    return np.mean(out, axis=0).reshape(-1, 3)


def _down_camera(y=0.5):
    return look_at_camera([0.0, y, 0.0], [0.001, 0.0, 0.0], fov=20.0, w=W, h=H)


def test_all_emitter_hemisphere_is_2pi_albedo_le_sin1():
    """A floor inside a box whose walls and ceiling all emit: every outer ray
    hits an emitter and no inner ray is sent."""
    a, le = np.array([0.8, 0.6, 0.4]), np.array([1.0, 0.5, 0.25])
    S, Hc = 5.0, 5.0
    sc = ArrayScene()
    sc.quad([-S, 0, -S], [0, 0, 2 * S], [2 * S, 0, 0], a, 2)
    sc.quad([-S, Hc, -S], [2 * S, 0, 0], [0, 0, 2 * S], le, 1)
    for o, u, v in (([-S, 0, -S], [2 * S, 0, 0], [0, Hc, 0]), ([-S, 0, S], [0, Hc, 0], [2 * S, 0, 0]),
                    ([-S, 0, -S], [0, Hc, 0], [0, 0, 2 * S]), ([S, 0, -S], [0, 0, 2 * S], [0, Hc, 0])):
        sc.quad(o, u, v, le, 1)
    img = _bounce(sc, _down_camera(1.0))
    want = 2.0 * float(REF_PI) * a * le * math.sin(1.0)
    got = img.mean(axis=0)
    # 48 x 48 pixels x 8 rays: the sample mean of cos(r1) has a relative sd of 1e-3
    assert np.allclose(got, want, rtol=5e-3), (got, want)
    # per pixel the value is 2 pi a le (1/8) sum_k cos(r1_k), inside [cos 1, 1] x 2 pi a le
    lo, hi = 2 * float(REF_PI) * a * le * math.cos(1.0), 2 * float(REF_PI) * a * le
    assert np.all(img >= lo * (1 - 1e-5)) and np.all(img <= hi * (1 + 1e-5))


def test_inner_rays_see_the_emitter_below_a_grey_ceiling():
    """Outer rays hit a non-emissive ceiling (payload 0) and each sends two
    inner rays down into an emissive floor: E = 2 pi a le (16 / 8) 0.25 sin 1."""
    a, le = np.array([0.7, 0.5, 0.3]), np.array([2.0, 1.0, 0.5])
    S = 1000.0
    sc = ArrayScene()
    # the primary patch, just above the emissive floor
    sc.quad([-0.05, 0.001, -0.05], [0, 0, 0.1], [0.1, 0, 0], a, 2)
    sc.quad([-S, 0, -S], [0, 0, 2 * S], [2 * S, 0, 0], le, 1)
    # a grey ceiling whose normal faces the floor
    sc.quad([-S, 1.0, -S], [2 * S, 0, 0], [0, 0, 2 * S], [0.5, 0.5, 0.5], 2)
    img = _bounce(sc, _down_camera(0.5))
    want = 2.0 * float(REF_PI) * a * le * 2.0 * 0.25 * math.sin(1.0)
    got = img.mean(axis=0)
    # 48 x 48 x 16 inner samples; rays that leave the 2000 x 2000 planes within tmax are negligible
    assert np.allclose(got, want, rtol=5e-3), (got, want)


def test_directional_light_payload_of_a_lambert_ceiling():
    """Outer rays hit the top-lit Lambert ceiling (its normal faces up, the
    light is straight above): payload = a_c / PI colour, unshadowed; the inner
    rays leave upwards and miss.  E = 2 PI a (a_c / PI) colour sin 1."""
    a, ac, col = np.array([0.9, 0.6, 0.3]), np.array([0.5, 0.4, 0.8]), np.array([1.5, 1.0, 2.0])
    S = 1000.0
    sc = ArrayScene()
    sc.quad([-S, 0, -S], [0, 0, 2 * S], [2 * S, 0, 0], a, 2)
    sc.quad([-S, 1.0, -S], [0, 0, 2 * S], [2 * S, 0, 0], ac, 2)
    dl = np.array([[*col, 0.0, 1.0, 0.0]], np.float32)
    img = _bounce(sc, _down_camera(0.5), dir_lights=dl)
    want = 2.0 * a * ac * col * math.sin(1.0)
    got = img.mean(axis=0)
    assert np.allclose(got, want, rtol=5e-3), (got, want)
    # the same ceiling lit from below its authored normal is dark: n.l clamps to 0
    dark = _bounce(sc, _down_camera(0.5), dir_lights=np.array([[*col, 0.0, -1.0, 0.0]], np.float32))
    assert np.all(dark == 0.0)


def test_no_emitters_and_no_lights_give_zero_and_misses_are_skipped():
    sc = ArrayScene()
    sc.quad([-5, 0, -5], [0, 0, 10], [10, 0, 0], [0.5, 0.5, 0.5], 2)
    sc.quad([-5, 1, -5], [10, 0, 0], [0, 0, 10], [0.5, 0.5, 0.5], 2)
    img = _bounce(sc, _down_camera(0.5))
    assert np.all(img == 0.0)
    # a camera looking into empty space: every pixel misses, the term is 0
    cam = look_at_camera([0.0, 0.5, 0.0], [0.0, 0.5, -1.0], fov=20.0, w=W, h=H)
    sc2 = ArrayScene()
    sc2.quad([-5, 0, -5], [0, 0, 10], [10, 0, 0], [0.5, 0.5, 0.5], 1)
    assert np.all(_bounce(sc2, cam) == 0.0)


# ---- the reference's own ReSTIR DI (Q10), RayGen.rgen:215-528 ----
#
# With one light the reservoirs' arithmetic has a closed form: every
# UpdateReservoir adds 1 to M, so after the spatial pass's combine of the
# pixel's own reservoir W = T.M T.W / (T.M + 1), and the temporal pass carries
# w_sum along: in frame f >= 1 (before the M cap at 20 x 32 bites, f < 20) the
# running w_sum holds 32 f candidates against M = 33 f, so the estimate is
# 32/33 of the light (exactly for one point light, in expectation over the
# candidates for one emissive triangle).  Frame 0 has no history, and the
# reference discards a fresh reservoir without one: only the directional
# light is left.


def _lamp_scene(albedo):
    sc = ArrayScene()
    sc.quad([-5, 0, -5], [0, 0, 10], [10, 0, 0], albedo, 0, roughness=1.0, metallic=0.0)
    return sc


def test_q10_single_point_light_is_32_33_of_calcpointlight():
    """One point light straight above a rough dielectric floor seen straight
    down (v = l = n): CalcPointlight = (0.96 / PI + 0.04 / (4.0001 PI)) albedo
    colour / (c + l d + e d^2), and the reservoirs deliver 32/33 of it."""
    import snkpt
    alb = np.array([0.7, 0.5, 0.3])
    sc = _lamp_scene(alb)
    col, hl = np.array([2.0, 3.0, 4.0]), 1.5
    att = (0.5, 0.2, 1.0)  # exp, linear, constant
    lb = snkpt.light_block(points=[(tuple(col), (0.2, hl, -0.1), att)])
    cam = look_at_camera([0.2, 3.0, -0.1], [0.2, 0.0, -0.1], fov=0.2)
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    brdf = 0.96 / REF_PI + 0.04 / (4.0001 * REF_PI)
    light = brdf * alb * col / (att[2] + att[1] * hl + att[0] * hl * hl)
    assert np.allclose(O.compat_direct(arrs, b, cam, 1, 1, 1, lb, 1, 0)[0, 0], 0.0)
    for frames in (2, 5, 12):
        got = O.compat_direct(arrs, b, cam, 1, 1, frames, lb, 1, 0)[0, 0]
        assert np.allclose(got, 32.0 / 33.0 * light, rtol=2e-4), (frames, got, light)
    # without temporal reuse every fresh reservoir is discarded (no history)
    assert np.allclose(O.compat_direct(arrs, b, cam, 1, 1, 5, lb, 0, 0)[0, 0], 0.0)


def test_q10_directional_light_is_calcdirectionallight_shadowed():
    """The directional light alone, every frame: (kd albedo / PI + spec) n.l
    colour with v = n, l = n at roughness 1 (h = n, F = 0.04, D = 1 / PI, G = 1)."""
    import snkpt
    alb = np.array([0.7, 0.5, 0.3])
    sc = _lamp_scene(alb)
    col = np.array([1.5, 1.0, 0.5])
    lb = snkpt.light_block(dir_colour=tuple(col), dir_to_light=(0.0, 1.0, 0.0))
    cam = look_at_camera([0.2, 3.0, -0.1], [0.2, 0.0, -0.1], fov=0.2)
    b = O.Bvh(sc.world())
    want = (0.96 * alb / REF_PI + 0.04 / (4.0001 * REF_PI)) * col
    for frames in (1, 3):
        got = O.compat_direct(sc.arrays(), b, cam, 1, 1, frames, lb, 1, 2)[0, 0]
        assert np.allclose(got, want, rtol=2e-4), (got, want)


def test_q10_single_emissive_triangle_expectation():
    """One emissive triangle over a white rough floor: the pdf the candidates
    divide by is 1 / (N_inst N_tri area) and the target has no emitter cosine,
    so frames 1..16 average 32/33 of the integral of Le f n.l / d^2 over the
    triangle (quadrature, f from the oracle's BRDF, which equals the
    reference's CookTorranceBRDF x albedo for a white floor)."""
    sc = _lamp_scene([1.0, 1.0, 1.0])
    le = np.array([6.0, 4.0, 2.0])
    tri = np.array([[-0.6, 1.2, -0.4], [0.7, 1.4, -0.3], [0.1, 1.1, 0.8]])
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    sc.add(tri, [0, 1, 2], [n / np.linalg.norm(n)] * 3, le, 1)
    w = h = 6
    cam = look_at_camera([1.8, 0.8, 0.3], [0.1, 0.0, 0.0], fov=30.0, w=w, h=h)  # below the emitter
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    frames = [O.compat_direct(arrs, b, cam, w, h, f, None, 1, 0, seed=3) for f in range(2, 18)]
    got = np.mean(frames, axis=0).reshape(-1, 3)
    import raysets
    org, dr = raysets.primary_rays(cam, w, h)
    xs = org[:, :3].astype(np.float64) + dr[:, :3] * (-org[:, 1] / dr[:, 1])[:, None]
    # Gauss-Legendre on the unit square mapped onto the triangle (Duffy)
    g, gw = np.polynomial.legendre.leggauss(24)
    g, gw = 0.5 * (g + 1), 0.5 * gw
    area2 = np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    want = np.zeros((w * h, 3))
    for p_i, x in enumerate(xs):
        wo = org[p_i, :3] - x
        wo /= np.linalg.norm(wo)
        acc = np.zeros(3)
        for a, wa in zip(g, gw):
            for c, wc in zip(g, gw):
                s1, s2 = a, a * c  # (1 - s1) v0 + (s1 - s2) v1 + s2 v2
                q = (1 - s1) * tri[0] + (s1 - s2) * tri[1] + s2 * tri[2]
                d = q - x
                r2 = d @ d
                wi = d / np.sqrt(r2)
                f, _ = O.brdf_eval([1.0, 1.0, 1.0], 1.0, 0.0, 0, [0, 1.0, 0], wo, wi)
                acc += wa * wc * a * le * f * max(wi[1], 0.0) / r2
        want[p_i] = acc * area2
    ratio = got.sum(0) / (32.0 / 33.0 * want.sum(0))
    assert np.allclose(ratio, 1.0, atol=0.02), ratio
    assert np.allclose(got, 32.0 / 33.0 * want, rtol=0.1)
