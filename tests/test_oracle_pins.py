"""CPU pins of the oracle functions the path tracer's parity rests on:
GGX direct light against quadrature, unequal-power emitter selection (closed
forms and a chi-square test of the alias sampler), Russian roulette in
expectation, and the cache MLP's forward pass on hand-set weights."""
import numpy as np
import pytest

from helpers import ArrayScene, ggx_room, look_at_camera, polygon_irradiance
from oracle import binding as O
from test_gpu_ggx import _emitter_quadrature, _ggx_floor


@pytest.mark.parametrize("rough,metal", [(0.3, 1.0), (0.5, 0.0)])
def test_oracle_ggx_floor_matches_quadrature(rough, metal):
    """The oracle's path tracer on a GGX floor equals the quadrature of f cos L
    over the emitter (NEE alone at depth 1; NEE + BRDF sample under MIS at 2)."""
    sc = _ggx_floor(rough, metal)
    b = O.Bvh(sc.world())
    eye = np.array([1.6, 1.2, 0.4])
    x0 = np.array([-0.6, 0.0, 0.1])
    wo = (eye - x0) / np.linalg.norm(eye - x0)
    want = _emitter_quadrature(x0, wo, rough, metal)
    cam = look_at_camera(eye, x0, fov=0.5, w=2, h=2)
    for depth in (1, 2):
        got = O.pt_render(sc.arrays(), b, cam, 2, 2, 20000, depth, 0, 5 + depth).reshape(-1, 3).mean(0)
        assert np.allclose(got, want, rtol=0.02, atol=1e-4), (depth, got, want)


def _two_emitter_scene():
    """Lambert floor under two emitters of unequal size and radiance (and so
    unequal power); B is coloured so the luminance weighting matters."""
    sc = ArrayScene()
    sc.quad([-10, 0, -10], [0, 0, 20], [20, 0, 0], [0.5] * 3, 2)
    qa = ([-1.3, 2.0, -0.5], [1.0, 0, 0], [0, 0, 1.0])
    qb = ([0.6, 1.5, 0.0], [0.6, 0, 0], [0, 0, 0.6])
    La, Lb = np.array([10.0, 10.0, 10.0]), np.array([40.0, 20.0, 5.0])
    sc.quad(*qa, La, 1)
    sc.quad(*qb, Lb, 1)

    def verts(q):
        o, a, b = (np.asarray(v) for v in q)
        return [o, o + a, o + a + b, o + b]
    return sc, verts(qa), verts(qb), La, Lb


@pytest.mark.parametrize("depth", [1, 2])
def test_unequal_emitters_sum_of_polygon_forms(depth):
    """Floor radiance = albedo/pi * (E_a La + E_b Lb) per channel (two Lambert
    polygon closed forms): the emitter selection pdf the estimator divides by
    must be the probability its alias table draws with."""
    sc, va, vb, La, Lb = _two_emitter_scene()
    b = O.Bvh(sc.world())
    for x in ([0.2, 0.0, 0.3], [-0.9, 0.0, -0.2], [1.2, 0.0, 0.6]):
        x = np.array(x)
        want = 0.5 / np.pi * (polygon_irradiance(x, [0, 1, 0], va, 1.0) * La +
                              polygon_irradiance(x, [0, 1, 0], vb, 1.0) * Lb)
        cam = look_at_camera([x[0] + 0.1, 1.0, x[2] + 0.3], x, fov=1.0)
        got = O.pt_render(sc.arrays(), b, cam, 1, 1, 40000, depth, 0, 11 + depth)[0, 0]
        assert np.allclose(got, want, rtol=0.012), (x, depth, got, want)


def _independent_power(arrs):
    """Per emissive triangle (world id -> power): luminance of the emission
    (Rec. 709 weights) times the triangle's world-space area, in float64."""
    pos, idx, inst, m4, mats, nrm, em = arrs
    world = O.flatten(pos, idx, inst, m4).astype(np.float64)
    off = np.concatenate([[0], np.cumsum(inst[:, 4].astype(np.int64) // 3)])
    power = {}
    for i in em:
        al = mats[inst[i, 3], 8:11].view(np.float32).astype(np.float64)
        lum = 0.2126 * al[0] + 0.7152 * al[1] + 0.0722 * al[2]
        for g in range(int(off[i]), int(off[i + 1])):
            v = world[g, :9].reshape(3, 3)
            area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
            if lum * area > 0:
                power[g] = lum * area
    return power


def test_alias_sampler_frequencies_on_many_lights():
    """Config 3's 10,000 emitters of log-uniform radiance: the table's pdf is
    power / total power, and 4M draws of its alias sampler follow that pdf
    (chi-square over all emitters)."""
    import snkpt
    scene = snkpt.SceneData(3)  # the arrays are views of its memory
    arrs = O.scene_arrays(scene)
    power = _independent_power(arrs)
    n = 4_000_000
    gid, pdf, cnt = O.emitter_sample(arrs, n, seed=3)
    assert gid.size == len(power) >= 10000
    total = sum(power.values())
    want = np.array([power[int(g)] / total for g in gid])
    assert np.allclose(pdf, want, rtol=1e-5)
    exp = n * want
    chi2 = ((cnt - exp) ** 2 / exp).sum()
    dof = gid.size - 1
    assert abs(chi2 - dof) < 6 * np.sqrt(2 * dof), (chi2, dof)
    assert cnt.sum() == n
    # a uniform draw (the reference's per-instance choice) would fail the same test
    chi2_u = ((cnt - n / gid.size) ** 2 / (n / gid.size)).sum()
    assert chi2_u > dof + 60 * np.sqrt(2 * dof)


def _tile_stats(arrs, b, cam, w, h, rr_start, runs, spp, seed0):
    means = np.stack([O.pt_render(arrs, b, cam, w, h, spp, 8, rr_start, seed0 + k) for k in range(runs)])
    t = means.reshape(runs, h // 4, 4, w // 4, 4, 3).mean((2, 4))
    return t.mean(0), t.std(0, ddof=1) / np.sqrt(runs)


def test_russian_roulette_is_unbiased():
    """Depth-8 paths on GGX materials: RR from bounce 1 and from bounce 3 give
    the image without RR in expectation (z-test per 4x4 tile and channel), and
    RR does change the individual samples."""
    sc, scene = ggx_room(5)
    arrs = sc.arrays()
    b = O.Bvh(sc.world())
    w = h = 16
    cam = scene.camera(0, w, h)
    m0, s0 = _tile_stats(arrs, b, cam, w, h, 0, 8, 64, 100)
    for rr in (1, 3):
        m1, s1 = _tile_stats(arrs, b, cam, w, h, rr, 8, 64, 200 + 10 * rr)
        z = (m1 - m0) / np.sqrt(s0 ** 2 + s1 ** 2 + 1e-12)
        assert np.abs(z).max() < 4.5, (rr, np.abs(z).max())
        assert abs(m1.mean() - m0.mean()) < 0.02 * m0.mean(), (rr, m1.mean(), m0.mean())
    a = O.pt_render(arrs, b, cam, w, h, 4, 8, 0, 9)
    r = O.pt_render(arrs, b, cam, w, h, 4, 8, 1, 9)
    assert (np.abs(a - r) > 1e-6).any(-1).mean() > 0.2


def routing_net(n=200, seed=4):
    """Hand-set weights that route chosen inputs through the 4 ReLU layers,
    an affine level-0 grid, n samples, and the closed-form outputs:
    y0 = relu(s) - relu(-s) = s (s = sin(pi d_x): only correct with the ReLU),
    y1 = 1.5 (roughness + 2 albedo_r) (a scale applied in the third layer),
    y2 = the level-0 grid feature of an affine grid = its affine closed form.
    Units are permuted between layers, so a transposed or misordered weight
    layout routes the wrong values."""
    W = np.zeros(O.MLP_PARAMS, np.float32)

    def w(layer, o, i, v):
        W[layer * 4096 + o * 64 + i] = v
    # layer 0: inputs -> units 5, 9, 17, 30, 41
    w(0, 5, 42, 1.0)           # sin(pi d_x)
    w(0, 9, 42, -1.0)
    w(0, 17, 38, 1.0)          # roughness
    w(0, 17, 39, 2.0)          # + 2 albedo_r
    w(0, 30, 0, 1.0)           # level-0 feature 0
    w(0, 41, 0, -1.0)
    # layers 1-3: a different permutation each time, a scale of 3 on the roughness path
    perm = [{5: 60, 9: 2, 17: 33, 30: 7, 41: 50}, {60: 11, 2: 12, 33: 13, 7: 14, 50: 15},
            {11: 0, 12: 1, 13: 2, 14: 3, 15: 4}]
    for layer, p in enumerate(perm, start=1):
        for i, o in p.items():
            w(layer, o, i, 3.0 if (layer == 2 and i == 33) else 1.0)
    out = 4 * 4096
    W[out + 0 * 64 + 0] = 1.0
    W[out + 0 * 64 + 1] = -1.0
    W[out + 1 * 64 + 2] = 0.5
    W[out + 2 * 64 + 3] = 1.0
    W[out + 2 * 64 + 4] = -1.0
    g = O.CacheGeom([0, 0, 0], [1, 1, 1])
    grid = np.zeros(O.GRID_PARAMS, np.float32)
    r = int(g.res[0])
    for z in range(r + 1):
        for y in range(r + 1):
            for x in range(r + 1):
                grid[2 * O.grid_index(x, y, z, r, int(g.dense[0]))] = (x - 2 * y + 3 * z) / r
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    s[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 6:9] = [0, 1, 0]
    s[:, 9] = rng.uniform(0.05, 1, n)
    s[:, 10:13] = rng.uniform(0.1, 0.9, (n, 3))
    p = s[:, 0:3].astype(np.float64)
    want = np.stack([np.sin(np.pi * s[:, 3].astype(np.float64)),
                     1.5 * (s[:, 9].astype(np.float64) + 2 * s[:, 10]),
                     p[:, 0] - 2 * p[:, 1] + 3 * p[:, 2]], 1)
    return g, W, grid, s, want


def test_mlp_forward_closed_form():
    """The oracle's encoding + forward pass on routing_net's weights gives the
    closed-form outputs."""
    g, W, grid, s, want = routing_net()
    y = O.cache_infer(g, grid, W, s)
    assert np.allclose(y, want, rtol=1e-6, atol=1e-6), np.abs(y - want).max()
