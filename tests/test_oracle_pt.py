"""CPU tests of the shading and path-tracing oracle against closed forms."""
import numpy as np
import pytest

from helpers import floor_light_scene, look_at_camera, polygon_irradiance
from oracle import binding as O

REF_PI = np.float32(3.1415926)


def test_brdf_closed_forms():
    n = np.array([0, 0, 1.0])
    # Lambert: albedo / PI for any pair of directions
    f, pdf = O.brdf_eval([0.3, 0.5, 0.7], 1.0, 0.0, 2, n, [0.6, 0, 0.8], [0, -0.6, 0.8])
    assert np.allclose(f, np.array([0.3, 0.5, 0.7], np.float32) / REF_PI, rtol=1e-6)
    assert np.isclose(pdf, 0.8 / np.pi, rtol=1e-6)  # cosine pdf
    # GGX, normal incidence (h = n, h.v = 1): F = f0, G = 1, D = 1 / (PI a^2)
    for rough, metal in [(1.0, 0.0), (0.5, 0.0), (0.5, 1.0), (0.2, 0.3)]:
        alb = np.array([0.5, 0.25, 0.125])
        f, _ = O.brdf_eval(alb, rough, metal, 0, n, n, n)
        a2 = rough ** 4
        f0 = 0.04 + (alb - 0.04) * metal
        D = 1.0 / (np.pi * a2)
        want = (1 - f0) * (1 - metal) * alb / np.pi + f0 * D / 4.0001
        # (n.h^2 (a^2-1) + 1) cancels in FP32 for small a, as in the reference formula
        assert np.allclose(f, want, rtol=2e-4), (rough, metal, f, want)
    # grazing view: n.v = 0 -> G = 0, no specular; diffuse (1-F)(1-m) albedo/PI with F = f0 + (1-f0) (1-h.v)^5
    wo = np.array([1.0, 0, 0])
    f, _ = O.brdf_eval([0.5] * 3, 0.5, 0.0, 0, n, wo, n)
    h = (wo + n) / np.linalg.norm(wo + n)
    F = 0.04 + 0.96 * (1 - h @ wo) ** 5
    assert np.allclose(f, (1 - F) * 0.5 / np.pi, rtol=1e-5)


@pytest.mark.parametrize("rough,metal,flags", [(1.0, 0.0, 2), (0.8, 0.0, 0), (0.3, 0.0, 0), (0.3, 1.0, 0),
                                               (0.5, 0.3, 0)])
def test_brdf_pdf_normalised_and_consistent(rough, metal, flags):
    rng = np.random.default_rng(1)
    n = np.array([0, 0, 1.0])
    th = 0.7
    wo = np.array([np.sin(th), 0, np.cos(th)])
    # integral of the pdf over the hemisphere = 1 (stratified quadrature in (cos, phi))
    N = 400
    c = (np.arange(N) + 0.5) / N
    ph = (np.arange(2 * N) + 0.5) / (2 * N) * 2 * np.pi
    total = 0.0
    for ci in c:
        si = np.sqrt(1 - ci * ci)
        for p in ph[::4]:
            wi = np.array([si * np.cos(p), si * np.sin(p), ci])
            total += O.brdf_eval([0.5] * 3, rough, metal, flags, n, wo, wi)[1]
    integral = total * (1.0 / N) * (2 * np.pi / (2 * N / 4))
    # E[1 / pdf(wi)] over the sampler's valid samples = solid angle of the hemisphere,
    # and the pdf integrates to the sampler's acceptance rate (GGX reflections that
    # fall below the surface are rejected, so the valid density is sub-normalised)
    acc, m, accepted = 0.0, 20000, 0
    for _ in range(m):
        ok, wi = O.brdf_sample([0.5] * 3, rough, metal, flags, n, wo, *rng.random(3))
        if ok:
            accepted += 1
            acc += 1.0 / O.brdf_eval([0.5] * 3, rough, metal, flags, n, wo, wi)[1]
    # 1/pdf is heavy-tailed for a pure specular lobe: wider tolerance there
    assert abs(acc / m / (2 * np.pi) - 1.0) < (0.05 if metal < 1 else 0.12), acc / m
    assert abs(integral - accepted / m) < 0.02, (integral, accepted / m)
    if flags == 2:
        assert abs(integral - 1.0) < 0.01


def _floor_render(depth, spp, seed=3):
    sc, verts = floor_light_scene()
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    x = np.array([0.3, 0.0, -0.2])
    cam = look_at_camera([0.3, 1.0, 0.2], x, fov=2.0)
    img = O.pt_render(arrs, b, cam, 1, 1, spp, depth, 0, seed)
    E = polygon_irradiance(x, [0, 1, 0], verts, 10.0)
    return img[0, 0], 0.5 / np.pi * E


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_direct_light_matches_lambert_polygon_formula(depth):
    """Floor radiance = albedo/pi * irradiance from the square emitter
    (closed form), whether it arrives by NEE alone (depth 1) or by NEE + BRDF
    sampling combined with MIS (depth >= 2): the MIS weights sum to one."""
    got, want = _floor_render(depth, 40000)
    assert np.allclose(got, want, rtol=0.01), (got, want)
