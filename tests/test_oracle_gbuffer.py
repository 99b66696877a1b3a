"""CPU tests of the oracle's G-buffer-input mode (RayGen.rgen:531-555): the
texel decoders against IEEE binary16 and the Vulkan SNORM rule, the primary
vertex reconstructed from depth against a closed form, and agreement in
expectation with the camera-ray path tracer on the same scene."""
import numpy as np
import pytest

from helpers import floor_light_scene, look_at_camera, polygon_irradiance
from oracle import binding as O


def test_half_decoder_matches_ieee_binary16_exhaustively():
    bits = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    want = bits.view(np.float16).astype(np.float32)
    got = np.array([O.half_to_float(int(b)) for b in bits], np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def _floor_gbuffer(x, eye):
    sc, verts = floor_light_scene()
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    cam = look_at_camera(eye, x, fov=2.0)
    depth, n3, a3, rm, fl = O.primary_surfaces(arrs, b, cam, 1, 1)
    return sc, verts, arrs, b, cam, O.encode_gbuffer(depth, n3, a3, rm, fl)


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_gbuffer_direct_light_matches_lambert_polygon_formula(depth):
    """A G-buffer texel of the Lambertian floor lit by the square emitter:
    radiance = decoded albedo / pi * irradiance at the position rebuilt from
    depth (closed form), by NEE alone or NEE + BRDF sampling under MIS."""
    x = np.array([0.3, 0.0, -0.2])
    sc, verts, arrs, b, cam, gb = _floor_gbuffer(x, [0.3, 1.0, 0.2])
    assert gb["depth"][0, 0] < 1.0 and gb["mat_flags"][0, 0] == 2
    albedo = np.float32(gb["albedo"][0, 0, 0]) / np.float32(127.0)  # SNORM8: c / 127
    img = O.pt_render_gbuffer(arrs, b, cam, gb, 1, 1, 40000, depth, 0, seed=5)
    want = albedo / np.pi * polygon_irradiance(x, [0, 1, 0], verts, 10.0)
    assert np.allclose(img[0, 0], want, rtol=0.01), (img[0, 0], want)


def test_gbuffer_position_from_depth_reproduces_the_hit():
    """A wrong unprojection moves the rebuilt point: its irradiance, a steep
    function of position under the small emitter, then misses the closed form
    at the true hit.  Probe points at different distances from the emitter's
    axis all match."""
    for xz in ([0.0, 0.0], [0.45, 0.1], [-0.8, 0.6]):
        x = np.array([xz[0], 0.0, xz[1]])
        sc, verts, arrs, b, cam, gb = _floor_gbuffer(x, [xz[0] + 0.5, 1.5, xz[1] + 0.7])
        albedo = np.float32(gb["albedo"][0, 0, 0]) / np.float32(127.0)
        img = O.pt_render_gbuffer(arrs, b, cam, gb, 1, 1, 20000, 1, 0, seed=9)
        want = albedo / np.pi * polygon_irradiance(x, [0, 1, 0], verts, 10.0)
        assert np.allclose(img[0, 0], want, rtol=0.015), (xz, img[0, 0], want)


def test_gbuffer_background_and_emissive_texels():
    """Cleared depth (1) renders black; an emissive texel's radiance is its
    decoded albedo (RayGen.rgen:539-542), whatever the light."""
    sc, verts = floor_light_scene()
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    cam = look_at_camera([0.0, 1.0, 0.3], [0.0, 0.0, 0.0], fov=30.0, w=2, h=1)
    depth = np.array([[1.0, 0.5]], np.float32)
    n3 = np.zeros((1, 2, 3), np.float32)
    n3[..., 1] = 1.0
    a3 = np.array([[[0.2, 0.4, 0.6], [0.9, 0.5, 0.1]]], np.float32)
    rm = np.full((1, 2, 2), 0.5, np.float32)
    fl = np.array([[1, 1]], np.uint32)
    gb = O.encode_gbuffer(depth, n3, a3, rm, fl)
    img = O.pt_render_gbuffer(arrs, b, cam, gb, 2, 1, 1, 4, 0, seed=1)
    assert np.array_equal(img[0, 0], np.zeros(3, np.float32))
    dec = gb["albedo"][0, 1, :3].astype(np.float32) / np.float32(127.0)
    assert np.array_equal(img[0, 1], dec)


def test_gbuffer_paths_match_camera_paths_in_expectation():
    """Room scene: a G-buffer rasterized from the camera rays' own hits gives
    the camera-ray image in expectation (quantised albedo / normal and the
    larger ray bias shift it by well under the tolerance)."""
    import snkpt
    scene = snkpt.SceneData(1)
    arrs = O.scene_arrays(scene)
    pos, idx, inst, m4, _ = scene.arrays()
    b = O.Bvh(O.flatten(pos, idx, inst, m4))
    w = h = 24
    cam = scene.camera(0, w, h)
    depth, n3, a3, rm, fl = O.primary_surfaces(arrs, b, cam, w, h)
    gb = O.encode_gbuffer(depth, n3, a3, rm, fl)
    spp = 512
    ref = O.pt_render(arrs, b, cam, w, h, spp, 4, 0, 11)
    got = O.pt_render_gbuffer(arrs, b, cam, gb, w, h, spp, 4, 0, 11)
    # a directly seen emitter is its SNORM8 albedo, clamped to 1 by the format
    em = (fl & 1) != 0
    assert em.any() and np.allclose(got[em], 1.0)
    lit = ~em
    assert abs(got[lit].mean() / ref[lit].mean() - 1.0) < 0.03, (got[lit].mean(), ref[lit].mean())
    # per-pixel: most pixels within the Monte Carlo noise of 512 spp
    rel = np.abs(got - ref).mean(-1)[lit] / (ref.mean(-1)[lit] + 0.05)
    assert np.mean(rel < 0.25) > 0.9, np.mean(rel < 0.25)
