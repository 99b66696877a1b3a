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
