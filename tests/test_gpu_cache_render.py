"""Cache render mode: paths terminated into the radiance cache at vertex Q,
training records built from full paths.

Pins: a zero network turns the mode into the path tracer truncated before the
cache vertex (expectation against the FP64 oracle); the applied cache term is
exactly throughput * max(network(query), 0) of the exported queries (checked
through snkpt_cache_infer); query records at the primary hit match the traced
geometry; training targets at a last-bounce vertex average to the closed-form
Lambertian radiance there (the subtraction of the path prefix and the division
by the throughput both matter); online training lowers the error."""
import sys

import numpy as np
import pytest

from helpers import floor_light_scene, look_at_camera, polygon_irradiance

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _room(w, h):
    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    return scene, ctx, scene.camera(0, w, h)


def _render(ctx, cam, w, h, **kw):
    import torch

    import snkpt
    out = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, snkpt.default_settings(**kw), out, w, h)
    torch.cuda.synchronize()
    return out[..., :3].cpu().numpy().copy()


def _oracle(ctx, cam, w, h, spp, depth):
    import torch

    import snkpt
    s = snkpt.default_settings(max_depth=depth, seed=77)
    acc = torch.zeros(h * w * 3, dtype=torch.float64, device="cuda")
    for k in range(spp // 256):
        s.frame = k
        ctx.oracle_accumulate(cam, s, acc, w, h, 256)
    torch.cuda.synchronize()
    return (acc / spp).view(h, w, 3).cpu().numpy()


def _zero_net(ctx):
    import snkpt
    ctx.cache_reset(1)
    ctx.cache_set_params(0, np.zeros(ctx.cache_params(0).size, np.float32))


def test_zero_cache_equals_pt_truncated_before_query(gpu):
    """Q = 2, zero network: E[image] = PT with max_depth 1 (direct light at the
    primary hit under full MIS, emission seen directly)."""
    import snkpt
    w = h = 32
    scene, ctx, cam = _room(w, h)
    _zero_net(ctx)
    ref = _oracle(ctx, cam, w, h, 2048, 1)
    acc = np.zeros((h, w, 3))
    frames = 128
    for f in range(frames):
        acc += _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=4, cache_query_depth=2, frame=f, seed=3,
                       flags=snkpt.FLAG_NO_TRAIN)
    mean = acc / frames
    st = ctx.frame_stats()
    tr = mean.reshape(4, 8, 4, 8, 3).mean((1, 3))
    to = ref.reshape(4, 8, 4, 8, 3).mean((1, 3))
    rel = np.abs(tr - to).sum() / to.sum()
    _say(f"zero cache: mean {mean.mean():.5f} vs PT depth 1 {ref.mean():.5f}, tile L1 rel {rel:.4f}, "
         f"queries {st.cache_queries}")
    assert st.cache_queries > 0 and st.cache_train_samples == 0
    assert abs(mean.mean() - ref.mean()) < 0.02 * ref.mean()
    assert rel < 0.05


def test_zero_indirect_cache_equals_pt_to_query_depth(gpu):
    """Indirect split, Q = 2, zero network: the query vertex keeps its NEE
    (weight 1) and nothing else, so E[image] = PT with max_depth 2."""
    import snkpt
    w = h = 32
    scene, ctx, cam = _room(w, h)
    _zero_net(ctx)
    ref = _oracle(ctx, cam, w, h, 2048, 2)
    acc = np.zeros((h, w, 3))
    frames = 128
    for f in range(frames):
        acc += _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=6, cache_query_depth=2, frame=f, seed=4,
                       flags=snkpt.FLAG_NO_TRAIN | snkpt.FLAG_CACHE_INDIRECT)
    mean = acc / frames
    tr = mean.reshape(4, 8, 4, 8, 3).mean((1, 3))
    to = ref.reshape(4, 8, 4, 8, 3).mean((1, 3))
    rel = np.abs(tr - to).sum() / to.sum()
    _say(f"zero indirect cache: mean {mean.mean():.5f} vs PT depth 2 {ref.mean():.5f}, tile L1 rel {rel:.4f}")
    assert abs(mean.mean() - ref.mean()) < 0.02 * ref.mean()
    assert rel < 0.05


@pytest.mark.parametrize("q", [1, 2])
def test_zero_indirect_cache_with_mis_probe_equals_pt_to_query_depth(gpu, q):
    """Indirect split with the MIS probe, zero network: the query vertex's
    direct light is NEE plus a traced BRDF sample under the power heuristic, so
    E[image] = PT with max_depth Q, as without the probe."""
    import snkpt
    w = h = 32
    scene, ctx, cam = _room(w, h)
    _zero_net(ctx)
    ref = _oracle(ctx, cam, w, h, 2048, q)
    acc = np.zeros((h, w, 3))
    frames = 128
    for f in range(frames):
        acc += _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=6, cache_query_depth=q, frame=f, seed=6,
                       flags=snkpt.FLAG_NO_TRAIN | snkpt.FLAG_CACHE_INDIRECT | snkpt.FLAG_CACHE_MIS)
    mean = acc / frames
    st = ctx.frame_stats()
    tr = mean.reshape(4, 8, 4, 8, 3).mean((1, 3))
    to = ref.reshape(4, 8, 4, 8, 3).mean((1, 3))
    rel = np.abs(tr - to).sum() / to.sum()
    _say(f"zero indirect cache + MIS probe, Q={q}: mean {mean.mean():.5f} vs PT depth {q} {ref.mean():.5f}, "
         f"tile L1 rel {rel:.4f}, extension rays {st.extension_rays}")
    assert st.extension_rays > 0
    assert abs(mean.mean() - ref.mean()) < 0.02 * ref.mean()
    assert rel < 0.05


@pytest.mark.parametrize("form", ["full", "indirect_mis"])
def test_cache_term_is_throughput_times_network_output(gpu, form):
    """render(W) - render(0) with the same seed = scatter of thr * max(y, 0)
    over the exported queries, y from snkpt_cache_infer on those queries (also
    with the indirect split and its MIS probe, whose continuation must not
    change the throughput recorded for the query)."""
    import torch

    import snkpt
    w = h = 48
    scene, ctx, cam = _room(w, h)
    ctx.cache_reset(5)
    W = ctx.cache_params(0).copy()
    W[4 * 4096:] *= 8.0  # a larger output layer so the term is well above rounding
    fl = snkpt.FLAG_NO_TRAIN
    if form == "indirect_mis":
        fl |= snkpt.FLAG_CACHE_INDIRECT | snkpt.FLAG_CACHE_MIS
    kw = dict(mode=snkpt.MODE_CACHE, max_depth=5, cache_query_depth=2, seed=21, frame=4, flags=fl)
    ctx.cache_set_params(0, W)
    img_w = _render(ctx, cam, w, h, **kw)
    q = ctx.cache_params(5).reshape(-1, 16)
    meta = ctx.cache_params(6).reshape(-1, 4)
    y = ctx.cache_infer(torch.from_numpy(q).cuda()).cpu().numpy()[:, :3]
    ctx.cache_set_params(0, np.zeros_like(W))
    img_0 = _render(ctx, cam, w, h, **kw)
    path = meta[:, 3].view(np.uint32)
    assert len(q) > w * h // 4 and path.max() < w * h and len(np.unique(path)) == len(path)
    want = np.zeros((w * h, 3))
    want[path] = meta[:, :3].astype(np.float64) * np.maximum(y, 0.0)
    got = (img_w - img_0).reshape(-1, 3).astype(np.float64)
    tol = 2e-6 * np.maximum(np.abs(img_w).reshape(-1, 3), 1.0)
    _say(f"{len(q)} queries, term mean {want.mean():.4f}, max err {np.abs(got - want).max():.2e}")
    assert want.max() > 1e-3
    assert np.all(np.abs(got - want) <= tol)


def test_primary_queries_record_the_hit(gpu):
    """Q = 1 on a Lambertian floor: each query is the traced primary hit."""
    import raysets

    import snkpt
    sc, verts = floor_light_scene(albedo=0.4)
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w = h = 16
    eye = np.array([0.2, 1.0, 0.1])
    cam = look_at_camera(eye, [0.3, 0.0, -0.2], fov=30.0, w=w, h=h)
    ctx.cache_reset(2)
    _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=3, cache_query_depth=1, seed=8,
            flags=snkpt.FLAG_NO_TRAIN)
    q = ctx.cache_params(5).reshape(-1, 16)
    meta = ctx.cache_params(6).reshape(-1, 4)
    assert len(q) == w * h
    path = meta[:, 3].view(np.uint32)
    org, dr = raysets.primary_rays(cam, w, h)
    d = dr[path, :3].astype(np.float64)
    t = -org[path, 1] / d[:, 1]
    x = org[path, :3] + d * t[:, None]
    assert np.allclose(q[:, 0:3], x, atol=1e-4)
    assert np.allclose(q[:, 3:6], -d, atol=1e-5)
    assert np.allclose(q[:, 6:9], [0, 1, 0], atol=1e-6)
    assert np.allclose(q[:, 9], 1.0) and np.allclose(q[:, 10:13], 0.4)
    assert np.allclose(meta[:, :3], 1.0)


def test_training_at_the_query_vertex_only(gpu):
    """SNKPT_FLAG_CACHE_TRAIN_AT_QUERY with Q = 1 on a Lambertian floor: the
    training samples are exactly the training paths' primary hits, one per path
    (every camera ray of this view hits the floor): points of the floor seen
    along camera rays, direction back to the eye, the floor's normal and albedo."""
    import snkpt
    sc, verts = floor_light_scene(albedo=0.4)
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w = h = 32
    eye = np.array([0.2, 1.0, 0.1])
    cam = look_at_camera(eye, [0.3, 0.0, -0.2], fov=30.0, w=w, h=h)
    ctx.cache_reset(2)
    _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=4, cache_query_depth=1, seed=8,
            cache_train_paths=w * h, flags=snkpt.FLAG_CACHE_INDIRECT | snkpt.FLAG_CACHE_TRAIN_AT_QUERY)
    n = int(ctx.frame_stats().cache_train_samples)
    assert n == w * h, n
    s = ctx.cache_params(7).reshape(-1, 16)[:n]
    assert np.allclose(s[:, 1], 0.0, atol=1e-4)
    to_eye = eye[None, :] - s[:, 0:3]
    to_eye /= np.linalg.norm(to_eye, axis=1, keepdims=True)
    assert np.allclose(s[:, 3:6], to_eye, atol=1e-3)
    assert np.allclose(s[:, 6:9], [0, 1, 0], atol=1e-6)
    assert np.allclose(s[:, 10:13], 0.4)


def _wall_scene(albedo_floor=0.5, albedo_wall=0.7):
    sc, verts = floor_light_scene(albedo=albedo_floor)
    # a Lambertian wall at x = 1.2 facing -x, below the emitter's height
    sc.quad([1.2, 0.0, -3.0], [0, 0, 6.0], [0, 1.8, 0], [albedo_wall] * 3, 2)
    return sc, verts


def test_training_targets_are_unbiased_radiance(gpu):
    """max_depth 2, camera on the floor: targets recorded at wall vertices
    (the last bounce: NEE only) average to albedo/pi * E(x) from Lambert's
    polygon formula.  Without the prefix subtraction they would carry the
    floor's direct light; without the throughput division they would be
    scaled by the floor albedo."""
    import snkpt
    sc, verts = _wall_scene()
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w = h = 64
    cam = look_at_camera([0.0, 1.5, 0.0], [0.0, 0.0, -0.3], fov=60.0, w=w, h=h)
    ctx.cache_reset(3)
    ratios_wall, ratios_floor_d1 = [], []
    for f in range(24):
        _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=2, cache_query_depth=2, frame=f, seed=13,
                cache_lr=1e-9)
        t = ctx.cache_params(7).reshape(-1, 16)
        on_wall = np.abs(t[:, 0] - 1.2) < 1e-4
        for s in t[on_wall]:
            e = polygon_irradiance(s[0:3], [-1, 0, 0], verts, 10.0)
            if e > 1e-3:
                ratios_wall.append(s[13] / (0.7 / np.pi * e))
    for f in range(6):
        _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=1, cache_query_depth=2, frame=f, seed=14,
                cache_lr=1e-9)
        t = ctx.cache_params(7).reshape(-1, 16)
        assert np.all(np.abs(t[:, 1]) < 1e-4)  # only primary (floor) vertices with depth 1
        for s in t[:2000]:
            ratios_floor_d1.append(s[13] / (0.5 / np.pi * polygon_irradiance(s[0:3], [0, 1, 0], verts, 10.0)))
    rw, rf = np.array(ratios_wall), np.array(ratios_floor_d1)
    _say(f"wall targets: n={len(rw)} mean ratio {rw.mean():.4f} +- {rw.std() / np.sqrt(len(rw)):.4f}; "
         f"floor depth-1: n={len(rf)} mean ratio {rf.mean():.4f} +- {rf.std() / np.sqrt(len(rf)):.4f}")
    assert len(rw) > 2000
    assert abs(rw.mean() - 1.0) < max(0.03, 4 * rw.std() / np.sqrt(len(rw)))
    assert abs(rf.mean() - 1.0) < max(0.02, 4 * rf.std() / np.sqrt(len(rf)))


def test_online_training_learns_the_room(gpu):
    """Room, depth 4, Q = 2: after online training the cache image is close
    to the converged path-traced image and a single cache frame has lower
    error than a single path-traced frame."""
    import snkpt
    w = h = 64
    scene, ctx, cam = _room(w, h)
    ref = _oracle(ctx, cam, w, h, 2048, 4)
    ctx.cache_reset(4)
    kw = dict(mode=snkpt.MODE_CACHE, max_depth=4, cache_query_depth=2, seed=31)
    for f in range(300):
        _render(ctx, cam, w, h, frame=f, **kw)
    st = ctx.frame_stats()
    assert st.cache_train_samples > 0
    frames = [_render(ctx, cam, w, h, frame=1000 + f, flags=snkpt.FLAG_NO_TRAIN, **kw) for f in range(64)]
    pts = [_render(ctx, cam, w, h, frame=f, mode=snkpt.MODE_PT, max_depth=4, seed=32) for f in range(64)]
    mean = np.mean(frames, 0)
    bias = abs(mean.mean() - ref.mean()) / ref.mean()
    e_cache = np.sqrt(np.mean([(f_ - ref) ** 2 for f_ in frames]))
    e_pt = np.sqrt(np.mean([(p - ref) ** 2 for p in pts]))
    _say(f"cache mean {mean.mean():.5f} vs PT {ref.mean():.5f} (rel bias {bias:.4f}); "
         f"one-frame RMSE cache {e_cache:.4f} vs PT {e_pt:.4f}")
    assert bias < 0.1
    assert e_cache < e_pt


def test_training_is_deterministic(gpu):
    import snkpt
    w = h = 32
    scene, ctx, cam = _room(w, h)
    outs = []
    for _ in range(2):
        ctx.cache_reset(6)
        for f in range(5):
            img = _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=3, frame=f, seed=2)
        outs.append((img, ctx.cache_params(0).copy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def _wall_grid(verts, nz=120, ny=36):
    """Midpoint quadrature of the wall x = 1.2, z in [-3, 3], y in [0, 1.8]:
    points, cell area and the direct irradiance at each point."""
    zs = -3.0 + (np.arange(nz) + 0.5) * 6.0 / nz
    ys = (np.arange(ny) + 0.5) * 1.8 / ny
    Z, Y = np.meshgrid(zs, ys)
    P = np.stack([np.full_like(Z, 1.2), Y, Z], -1).reshape(-1, 3)
    E = np.array([polygon_irradiance(p, [-1, 0, 0], verts, 10.0) for p in P])
    return P, (6.0 / nz) * (1.8 / ny), E


def test_indirect_training_targets_match_one_bounce_quadrature(gpu):
    """Indirect split, max_depth 2: a floor vertex's target is the radiance
    its continuation brings back from the wall (the wall's NEE, the last
    vertex), excluding the floor's own direct light.  Its expectation is
    rho_f/pi * integral over the wall of (rho_w/pi) E_d(y) cos cos / r^2 dA,
    with E_d from Lambert's polygon formula.  Wall vertices (last bounce)
    have no indirect light: their targets are exactly 0."""
    import snkpt
    sc, verts = _wall_scene()
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w = h = 64
    cam = look_at_camera([0.0, 1.5, 0.0], [0.0, 0.0, -0.3], fov=60.0, w=w, h=h)
    ctx.cache_reset(3)
    P, dA, Ew = _wall_grid(verts)
    Lw = 0.7 / np.pi * Ew
    got, want, n_wall = 0.0, 0.0, 0
    for f in range(24):
        _render(ctx, cam, w, h, mode=snkpt.MODE_CACHE, max_depth=2, cache_query_depth=2, frame=f, seed=17,
                cache_lr=1e-9, flags=snkpt.FLAG_CACHE_INDIRECT)
        t = ctx.cache_params(7).reshape(-1, 16).astype(np.float64)
        on_wall = np.abs(t[:, 0] - 1.2) < 1e-4
        assert np.all(t[on_wall, 13:16] == 0.0)
        n_wall += int(on_wall.sum())
        fl = t[np.abs(t[:, 1]) < 1e-4]
        d = P[None, :, :] - fl[:, None, 0:3]
        r2 = (d ** 2).sum(-1)
        cx = d[..., 1] / np.sqrt(r2)
        cy = d[..., 0] / np.sqrt(r2)
        e = 0.5 / np.pi * (Lw[None, :] * cx * cy / r2).sum(1) * dA
        got += fl[:, 13].sum()
        want += e.sum()
    _say(f"indirect targets at the floor: sum {got:.3f} vs quadrature {want:.3f} (ratio {got / want:.4f}); "
         f"{n_wall} wall samples, all 0")
    assert n_wall > 1000
    assert abs(got / want - 1.0) < 0.03
