"""G-buffer-input mode (snkpt_render_gbuffer; RayGen.rgen:531-555) on the GPU:
per pixel against the CPU oracle's G-buffer mode, directly seen emission and
background, ReSTIR and cache modes over a G-buffer, motion vectors, the RT
mirror and argument checks.  The G-buffers are rasterized stand-ins built on
the CPU from the oracle's primary hits, in the reference's image formats."""
import sys

import numpy as np
import pytest

from helpers import floor_light_scene, look_at_camera
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene_gbuffer(cfg, w, h):
    import snkpt
    scene = snkpt.SceneData(cfg)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    arrs = O.scene_arrays(scene)
    pos, idx, inst, m4, _ = scene.arrays()
    b = O.Bvh(O.flatten(pos, idx, inst, m4))
    cam = scene.camera(0, w, h)
    depth, n3, a3, rm, fl = O.primary_surfaces(arrs, b, cam, w, h)
    return scene, ctx, arrs, b, cam, O.encode_gbuffer(depth, n3, a3, rm, fl)


def _dev(gb):
    import torch
    out = {}
    for k, v in gb.items():
        a = np.ascontiguousarray(v)
        if a.dtype == np.uint16:
            a = a.view(np.int16)
        out[k] = torch.from_numpy(a).cuda()
    return out


def _render_gb(ctx, cam, gbd, w, h, **kw):
    import torch

    import snkpt
    out = torch.zeros(h, w, 4, device="cuda")
    ctx.render_gbuffer(cam, snkpt.default_settings(**kw), gbd, out, w, h)
    torch.cuda.synchronize()
    return out[..., :3].cpu().numpy().copy()


def test_gbuffer_pt_matches_cpu_oracle_per_pixel(gpu):
    _say("room, G-buffer PT vs oracle")
    w = h = 64
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    g = _render_gb(ctx, cam, _dev(gb), w, h, spp=4, max_depth=4, rr_start=0, seed=7)
    c = O.pt_render_gbuffer(arrs, b, cam, gb, w, h, 4, 4, 0, 7)
    assert np.isfinite(g).all() and g.mean() > 0.01
    close = np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-5
    frac = close.all(-1).mean()
    _say(f"pixels agreeing: {frac:.4f}; means {g.mean():.6f} vs {c.mean():.6f}")
    assert frac >= 0.97, frac
    assert abs(g.mean() - c.mean()) <= 0.01 * c.mean()
    st = ctx.frame_stats()
    assert st.camera_rays == 0 and st.extension_rays > 0 and st.shadow_rays > 0


def test_gbuffer_background_and_emissive_texels(gpu):
    """Cleared depth renders black; an emissive texel is its decoded albedo."""
    import snkpt
    sc, _ = floor_light_scene()
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w, h = 4, 2
    cam = look_at_camera([0.0, 1.0, 0.3], [0.0, 0.0, 0.0], fov=30.0, w=w, h=h)
    depth = np.full((h, w), 0.5, np.float32)
    depth[0, :2] = 1.0
    n3 = np.zeros((h, w, 3), np.float32)
    n3[..., 1] = 1.0
    a3 = np.random.default_rng(2).random((h, w, 3)).astype(np.float32)
    rm = np.full((h, w, 2), 0.5, np.float32)
    fl = np.ones((h, w), np.uint32)
    gb = O.encode_gbuffer(depth, n3, a3, rm, fl)
    g = _render_gb(ctx, cam, _dev(gb), w, h, max_depth=4, seed=1)
    assert np.array_equal(g[0, :2], np.zeros((2, 3), np.float32))
    dec = gb["albedo"][..., :3].astype(np.float32) / np.float32(127.0)
    assert np.array_equal(g[0, 2:], dec[0, 2:]) and np.array_equal(g[1], dec[1])


def test_gbuffer_whole_paths_flag_keeps_the_wavefront_image(gpu):
    import snkpt
    w = h = 40
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    a = _render_gb(ctx, cam, gbd, w, h, max_depth=5, rr_start=3, seed=2)
    c = _render_gb(ctx, cam, gbd, w, h, max_depth=5, rr_start=3, seed=2, flags=snkpt.FLAG_WHOLE_PATHS)
    assert np.array_equal(a, c)


def test_gbuffer_tail_bounces_keep_the_wavefront_image(gpu):
    """The bounces after the G-buffer vertex, run one thread per path from
    any bounce on: the wavefront's image bit for bit."""
    w = h = 40
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    a = _render_gb(ctx, cam, gbd, w, h, max_depth=5, rr_start=3, seed=2, tail_bounce=0xFFFFFFFF)
    for t in (1, 2, 3, 4):
        c = _render_gb(ctx, cam, gbd, w, h, max_depth=5, rr_start=3, seed=2, tail_bounce=t)
        assert np.array_equal(a, c), t


@pytest.mark.parametrize("temporal,frames,tol", [(0, 96, 0.05), (1, 384, 0.08)])
def test_gbuffer_restir_equals_pt_in_expectation(gpu, temporal, frames, tol):
    """Many-light scene over a G-buffer: ReSTIR frames average to the
    G-buffer path tracer's converged image (CPU oracle, 2048 spp).  Temporal
    reuse correlates the frames, so it gets more of them and a wider tile
    tolerance (measured: tile L1 0.044 at 384 frames, 0.022 at 1536)."""
    import snkpt
    w = h = 48
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(3, w, h)
    gbd = _dev(gb)
    ref = O.pt_render_gbuffer(arrs, b, cam, gb, w, h, 2048, 1, 0, 77, threads=16)
    acc = np.zeros((h, w, 3))
    for f in range(frames):
        acc += _render_gb(ctx, cam, gbd, w, h, mode=snkpt.MODE_RESTIR, max_depth=1, num_temporal_resamples=temporal,
                          num_spatial_resamples=3, frame=f, seed=5)
    mean = acc / frames
    tr = mean.reshape(6, 8, 6, 8, 3).mean((1, 3))
    to = ref.reshape(6, 8, 6, 8, 3).mean((1, 3))
    rel = np.abs(tr - to).sum() / to.sum()
    _say(f"G-buffer ReSTIR (temporal {temporal}, {frames} frames) mean {mean.mean():.5f} vs PT {ref.mean():.5f}, "
         f"tile L1 rel {rel:.4f}")
    assert abs(mean.mean() - ref.mean()) < 0.02 * ref.mean()
    assert rel < tol


def test_gbuffer_restir_zero_motion_equals_reprojection(gpu):
    """Static camera: zero motion vectors pick the same history texel as
    reprojection, so the frames are bit-identical."""
    import torch

    import snkpt
    w = h = 40
    scene, ctx_a, arrs, b, cam, gb = _scene_gbuffer(3, w, h)
    ctx_b = snkpt.Context(0)
    ctx_b.load_scene(scene)
    ctx_b.build_accel()
    gbd = _dev(gb)
    gbm = dict(gbd)
    gbm["motion"] = torch.zeros(h, w, 2, dtype=torch.float16, device="cuda")
    for f in range(3):
        kw = dict(mode=snkpt.MODE_RESTIR, max_depth=2, num_temporal_resamples=1, num_spatial_resamples=2, frame=f,
                  seed=8)
        a = _render_gb(ctx_a, cam, gbd, w, h, **kw)
        m = _render_gb(ctx_b, cam, gbm, w, h, **kw)
        assert np.array_equal(a, m), f


def test_gbuffer_zero_indirect_cache_at_primary_equals_direct_light(gpu):
    """Cache mode over a G-buffer, indirect split at the primary vertex, zero
    network: the texel keeps its NEE (weight 1) and the cache adds exactly 0,
    so the image is bit-identical to G-buffer PT with max_depth 1."""
    import snkpt
    w = h = 48
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    ctx.cache_reset(1)
    ctx.cache_set_params(0, np.zeros(ctx.cache_params(0).size, np.float32))
    for f in range(2):
        c = _render_gb(ctx, cam, gbd, w, h, mode=snkpt.MODE_CACHE, max_depth=4, cache_query_depth=1, frame=f, seed=6,
                       flags=snkpt.FLAG_NO_TRAIN | snkpt.FLAG_CACHE_INDIRECT, cache_ema=0.0)
        st = ctx.frame_stats()
        p = _render_gb(ctx, cam, gbd, w, h, max_depth=1, frame=f, seed=6)
        assert st.cache_queries > 0 and st.camera_rays == 0
        assert np.array_equal(c, p), (f, np.abs(c - p).max())


def test_gbuffer_cache_mode_trains_and_renders(gpu):
    """Cache mode over a G-buffer with training: finite images, queries at the
    G-buffer vertices (Q = 1 indirect and Q = 2 full) and training samples
    from camera paths."""
    import snkpt
    w = h = 64
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    for q, fl in ((1, snkpt.FLAG_CACHE_INDIRECT), (2, 0)):
        ctx.cache_reset(3)
        for f in range(4):
            g = _render_gb(ctx, cam, gbd, w, h, mode=snkpt.MODE_CACHE, max_depth=5, cache_query_depth=q, frame=f,
                           seed=4, flags=fl)
        st = ctx.frame_stats()
        _say(f"Q={q}: mean {g.mean():.5f}, queries {st.cache_queries}, training samples {st.cache_train_samples}")
        assert np.isfinite(g).all() and g.mean() > 0.0
        assert st.cache_queries > 0 and st.cache_train_samples > 0


def test_rt_mirror_takes_the_gbuffer_resources(gpu):
    import torch

    import snkpt
    w = h = 32
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    rt = snkpt.RT()
    out = torch.zeros(h, w, 4, device="cuda")
    rt.Init(scene, out)
    s = snkpt.default_settings(max_depth=3, seed=9)
    rt.RecordRenderCmdBuf(None, out, scene, s, cam=cam, resources=gbd)
    torch.cuda.synchronize()
    want = _render_gb(ctx, cam, gbd, w, h, max_depth=3, seed=9)
    assert np.array_equal(out[..., :3].cpu().numpy(), want)


def test_render_gbuffer_rejects_bad_input(gpu):
    import torch

    import snkpt
    w = h = 16
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    out = torch.zeros(h, w, 4, device="cuda")
    s = snkpt.default_settings()
    missing = dict(gbd)
    missing["depth"] = None
    with pytest.raises(snkpt.SnkptError):
        ctx.render_gbuffer(cam, s, missing, out, w, h)
    raw = torch.zeros(w * h * 8 + 2, dtype=torch.uint8, device="cuda")
    bad = dict(gbd)
    bad["normal"] = raw[2:]  # 2-byte offset: not 8-byte aligned
    with pytest.raises(snkpt.SnkptError):
        ctx.render_gbuffer(cam, s, bad, out, w, h)
    empty = snkpt.Context(0)
    with pytest.raises(snkpt.SnkptError):
        empty.render_gbuffer(cam, s, gbd, out, w, h)  # no scene built


def test_gbuffer_reference_compat_matches_camera_rays(gpu):
    """SNKPT_MODE_REFERENCE_COMPAT over a G-buffer: its primary vertices are the
    camera rays' hits rebuilt from depth, so with the same seeds the image
    matches the camera-ray compat image up to the depth round trip."""
    import torch

    import snkpt
    w = h = 64
    scene, ctx, arrs, b, cam, gb = _scene_gbuffer(1, w, h)
    gbd = _dev(gb)
    ctx2 = snkpt.Context(0)
    ctx2.load_scene(scene)
    ctx2.build_accel()
    g_img, c_img = [], []
    for f in range(4):
        kw = dict(mode=snkpt.MODE_REFERENCE_COMPAT, seed=9, frame=f, num_temporal_resamples=0)
        g_img.append(_render_gb(ctx, cam, gbd, w, h, **kw))
        out = torch.zeros(h, w, 4, device="cuda")
        ctx2.render(cam, snkpt.default_settings(**kw), out, w, h)
        torch.cuda.synchronize()
        c_img.append(out[..., :3].cpu().numpy().copy())
    g, c = np.mean(g_img, 0), np.mean(c_img, 0)
    assert np.all(np.isfinite(g)) and g.mean() > 0.0
    # emissive texels hold the SNORM8-clamped albedo, not the emission: left out
    keep = (gb["mat_flags"].astype(np.int64) & 1) == 0
    g, c = g[keep], c[keep]
    close = np.abs(g - c).max(axis=-1) <= 1e-2 * (1.0 + np.abs(c).max(axis=-1))
    _say(f"compat G-buffer vs camera: {close.mean():.3f} of pixels agree, means {g.mean():.5f} {c.mean():.5f}")
    # the G-buffer vertex carries the reference's ray bias (0.01 + 0.001 t), camera
    # hits their own: shadow rays near contact edges may differ
    assert close.mean() >= 0.8
    assert abs(g.mean() - c.mean()) <= 0.02 * c.mean()
