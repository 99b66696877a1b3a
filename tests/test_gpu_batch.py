"""Batched rendering (snkpt_batch_*): many views of several scenes in one set
of queues must give, view by view, exactly the image snkpt_render gives for
that view alone (same per-view seed, same frame); grouping, launch shapes,
graph capture/replay and the autotuner must not change a single bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (config, variant) of the scenes; (scene, w, h, seed) of the views
SCENES = [(5, 1), (5, 2), (1, 0), (5, 11)]
VIEWS = [(0, 96, 64, 11), (1, 128, 128, 12), (2, 200, 150, 13), (0, 64, 64, 14), (3, 150, 100, 15),
         (2, 33, 17, 16)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup():
    import torch

    import snkpt
    scenes, ctxs = [], []
    for cfg, var in SCENES:
        sd = snkpt.SceneData(cfg, var)
        c = snkpt.Context(0)
        c.load_scene(sd)
        c.build_accel()
        scenes.append(sd)
        ctxs.append(c)
    b = snkpt.Batch(0)
    for c in ctxs:
        b.add_scene(c)
    outs, cams = [], []
    for k, (si, w, h, seed) in enumerate(VIEWS):
        cam = scenes[si].camera(k, w, h)
        out = torch.zeros(h, w, 4, device="cuda")
        b.add_view(si, cam, w, h, seed, out)
        outs.append(out)
        cams.append(cam)
    return scenes, ctxs, b, outs, cams


def _settings(**kw):
    import snkpt
    base = dict(mode=snkpt.MODE_PT, max_depth=6, rr_start=3)
    base.update(kw)
    return snkpt.default_settings(**base)


def _single(ctxs, cams, frame):
    import torch
    res = []
    for k, (si, w, h, seed) in enumerate(VIEWS):
        o = torch.zeros(h, w, 4, device="cuda")
        ctxs[si].render(cams[k], _settings(seed=seed, frame=frame), o, w, h)
        res.append(o.cpu().numpy())
    return res


def _grab(outs):
    import torch
    torch.cuda.synchronize()
    return [o.cpu().numpy().copy() for o in outs]


def test_batch_equals_single_view_renders(gpu):
    scenes, ctxs, b, outs, cams = _setup()
    for frame in (0, 3):
        b.render(_settings(frame=frame))
        got = _grab(outs)
        want = _single(ctxs, cams, frame)
        for k in range(len(VIEWS)):
            assert np.array_equal(got[k], want[k]), (frame, k, np.abs(got[k] - want[k]).max())
        _say(f"frame {frame}: {len(VIEWS)} views bit-identical, mean {np.mean([g.mean() for g in got]):.5f}")
    st = b.frame_stats()
    assert st.camera_rays == sum(w * h for _, w, h, _ in VIEWS)
    assert st.extension_rays > 0 and st.shadow_rays > 0


def test_grouping_and_launch_shapes_do_not_change_images(gpu):
    import snkpt
    scenes, ctxs, b, outs, cams = _setup()
    b.render(_settings(frame=5))
    ref = _grab(outs)
    for tb, tpc, spc, cap in ((64, 4, 2, 20000), (128, 32, 16, 1), (64, 8, 4, 0)):
        t = snkpt.BatchTuning()
        t.trace_block, t.trace_blocks_per_cu, t.shade_blocks_per_cu, t.queue_capacity = tb, tpc, spc, cap
        b.set_tuning(t)
        for o in outs:
            o.zero_()
        b.render(_settings(frame=5))
        got = _grab(outs)
        for k in range(len(VIEWS)):
            assert np.array_equal(got[k], ref[k]), (tb, tpc, spc, cap, k)


def test_tail_bounce_does_not_change_images_or_ray_counts(gpu):
    """The paths still alive from any bounce on run one thread per path: the
    images and ray counts of the launch-per-bounce frame, bit for bit."""
    import snkpt
    scenes, ctxs, b, outs, cams = _setup()
    s = _settings(frame=4)
    t = snkpt.BatchTuning()
    t.trace_block, t.trace_blocks_per_cu, t.shade_blocks_per_cu, t.tail_bounce = 128, 16, 8, 0xFFFFFFFF
    b.set_tuning(t)
    b.render(s)
    ref = _grab(outs)
    rs = b.frame_stats()
    assert rs.extension_rays > 0
    for tail in range(1, s.max_depth):
        t.tail_bounce = tail
        b.set_tuning(t)
        for o in outs:
            o.zero_()
        b.render(s)
        got = _grab(outs)
        st = b.frame_stats()
        for k in range(len(VIEWS)):
            assert np.array_equal(got[k], ref[k]), (tail, k)
        assert (st.extension_rays, st.shadow_rays) == (rs.extension_rays, rs.shadow_rays), tail


def test_graph_replays_match_eager_frames(gpu):
    import torch
    scenes, ctxs, b, outs, cams = _setup()
    st = torch.cuda.Stream()
    eager = []
    for f in range(4):
        b.render(_settings(frame=f), use_graph=False, stream=st)
        st.synchronize()
        eager.append(_grab(outs))
    for f in range(4):
        for o in outs:
            o.zero_()  # the program's own eager work between replays
        torch.cuda.synchronize()
        b.render(_settings(frame=f), use_graph=True, stream=st)
        st.synchronize()
        got = _grab(outs)
        for k in range(len(VIEWS)):
            assert np.array_equal(got[k], eager[f][k]), (f, k)
    _say("4 graph replays bit-identical to eager frames")


def test_accumulation_is_the_running_mean(gpu):
    import snkpt
    scenes, ctxs, b, outs, cams = _setup()
    frames = []
    for f in range(3):
        b.render(_settings(frame=f))
        frames.append(_grab(outs))
    for f in range(3):
        fl = snkpt.FLAG_ACCUMULATE | (snkpt.FLAG_RESET if f == 0 else 0)
        b.render(_settings(frame=f, flags=fl))
    got = _grab(outs)
    for k in range(len(VIEWS)):
        want = (frames[0][k].astype(np.float64) + frames[1][k] + frames[2][k]) / 3.0
        assert np.allclose(got[k][..., :3], want[..., :3], rtol=1e-5, atol=1e-6)


def test_autotune_keeps_images(gpu):
    import torch
    scenes, ctxs, b, outs, cams = _setup()
    st = torch.cuda.Stream()
    b.render(_settings(frame=2), stream=st)
    st.synchronize()
    ref = _grab(outs)
    t = b.autotune(_settings(frame=0), stream=st)
    _say(f"autotuned: {t.as_dict()}")
    assert t.ms_per_frame > 0 and t.trace_block in (64, 128)
    b.render(_settings(frame=2), stream=st)
    st.synchronize()
    got = _grab(outs)
    for k in range(len(VIEWS)):
        assert np.array_equal(got[k], ref[k])


def test_batch_rejects_bad_input(gpu):
    import snkpt
    b = snkpt.Batch(0)
    with pytest.raises(snkpt.SnkptError):
        b.render(_settings())  # no views
    c = snkpt.Context(0)
    with pytest.raises(snkpt.SnkptError):
        b.add_view(0, snkpt.Camera(), 8, 8, 1, None)
    b.add_scene(c)  # not built: rejected at render time
    import torch
    out = torch.zeros(8, 8, 4, device="cuda")
    b.add_view(0, snkpt.Camera(), 8, 8, 1, out)
    with pytest.raises(snkpt.SnkptError):
        b.render(_settings())
    with pytest.raises(snkpt.SnkptError):
        b.render(_settings(mode=snkpt.MODE_RESTIR))


def _contexts(scene_ids):
    import snkpt
    out = []
    for cfg, var in scene_ids:
        sd = snkpt.SceneData(cfg, var)
        c = snkpt.Context(0)
        c.load_scene(sd)
        c.build_accel()
        out.append((sd, c))
    return out


@pytest.mark.parametrize("form", ["full", "indirect", "indirect_mis"])
def test_batch_cache_form_equals_single_view_cache_renders(gpu, form):
    """Cache form, one view per scene, training on: each batched view equals
    the single-context cache render of that view (own networks, same seeds),
    frame after frame, bit for bit."""
    import torch

    import snkpt
    ids = [(5, 1), (1, 0), (5, 2)]
    views = [(0, 96, 64, 21), (1, 64, 64, 22), (2, 80, 48, 23)]
    A = _contexts(ids)
    Bs = _contexts(ids)
    for (_, c) in A + Bs:
        c.cache_reset(7)
    b = snkpt.Batch(0)
    for _, c in A:
        b.add_scene(c)
    outs, cams = [], []
    for k, (si, w, h, seed) in enumerate(views):
        cam = A[si][0].camera(k, w, h)
        o = torch.zeros(h, w, 4, device="cuda")
        b.add_view(si, cam, w, h, seed, o)
        outs.append(o)
        cams.append(cam)
    indirect = form != "full"
    fl = (snkpt.FLAG_CACHE_INDIRECT if indirect else 0) | (snkpt.FLAG_CACHE_MIS if form == "indirect_mis" else 0)
    kw = dict(mode=snkpt.MODE_CACHE, max_depth=4, rr_start=3, cache_train_paths=1 << 20,
              cache_query_depth=1 if indirect else 2, flags=fl)
    for frame in range(3):
        b.render(_settings(frame=frame, **kw))
        got = _grab(outs)
        for k, (si, w, h, seed) in enumerate(views):
            o = torch.zeros(h, w, 4, device="cuda")
            Bs[si][1].render(cams[k], _settings(frame=frame, seed=seed, **kw), o, w, h)
            torch.cuda.synchronize()
            assert np.array_equal(got[k], o.cpu().numpy()), (frame, k)
    st = b.frame_stats()
    _say(f"cache form: {st.cache_queries} queries, {st.cache_train_samples} training samples in the last frame")
    assert st.cache_queries > 0 and st.cache_train_samples > 0
    for (_, ca), (_, cb) in zip(A, Bs):
        assert np.array_equal(ca.cache_params(0), cb.cache_params(0))


@pytest.mark.parametrize("form", ["full", "indirect_mis"])
def test_batch_cache_form_graph_replays_match_eager_frames(gpu, form):
    """The cache form recorded once as a hipGraph (queries grouped by scene on
    the device, training batch sizes and Adam steps on the device, training
    streams from the frame index in device memory) and replayed: every view and
    every scene's network equal an eager batch over twin contexts after each
    frame, bit for bit, with several views per scene and accumulation on."""
    import torch

    import snkpt
    _say(f"batch cache form: graph replays against eager frames, form {form}")
    _, ctx_e, b_e, outs_e, _ = _setup()
    _, ctx_g, b_g, outs_g, _ = _setup()
    for c in ctx_e + ctx_g:
        c.cache_reset(9)
    indirect = form != "full"
    fl = snkpt.FLAG_ACCUMULATE | ((snkpt.FLAG_CACHE_INDIRECT | snkpt.FLAG_CACHE_MIS) if indirect else 0)
    kw = dict(mode=snkpt.MODE_CACHE, max_depth=5, rr_start=3, cache_train_paths=8192, cache_lr=2.5e-3,
              cache_ema=0.0, cache_query_depth=1 if indirect else 2)
    stream = torch.cuda.Stream()
    for frame in range(4):
        f = fl | (snkpt.FLAG_RESET if frame == 0 else 0)
        b_e.render(_settings(frame=frame, flags=f, **kw), use_graph=False, stream=stream)
        b_g.render(_settings(frame=frame, flags=f, **kw), use_graph=True, stream=stream)
        stream.synchronize()
        ge, gg = _grab(outs_e), _grab(outs_g)
        for k in range(len(VIEWS)):
            assert np.array_equal(ge[k], gg[k]), (frame, k)
        _say(f"batch cache form: frame {frame} equal")
    assert not np.array_equal(ge[0], np.zeros_like(ge[0]))
    for ce, cg in zip(ctx_e, ctx_g):
        assert np.array_equal(ce.cache_params(0), cg.cache_params(0))
    st = b_g.frame_stats(stream)
    assert st.cache_queries > 0 and st.cache_train_samples > 0
    b_e.close()
    b_g.close()


def test_batch_cache_form_frozen_many_views(gpu):
    """Several views per scene, networks frozen: every batched view equals the
    single-context cache render of that view."""
    import torch

    import snkpt
    scenes, ctxs, b, outs, cams = _setup()
    single = _contexts(SCENES)
    for (_, c), c2 in zip(single, ctxs):
        c.cache_reset(3)
        c2.cache_reset(3)
    kw = dict(mode=snkpt.MODE_CACHE, max_depth=5, rr_start=3, cache_query_depth=2, flags=snkpt.FLAG_NO_TRAIN)
    b.render(_settings(frame=1, **kw))
    got = _grab(outs)
    for k, (si, w, h, seed) in enumerate(VIEWS):
        o = torch.zeros(h, w, 4, device="cuda")
        single[si][1].render(cams[k], _settings(frame=1, seed=seed, **kw), o, w, h)
        torch.cuda.synchronize()
        assert np.array_equal(got[k], o.cpu().numpy()), k


def test_full_config5_batch_matches_single_views_and_oracle(gpu):
    """The whole config-5 workload (16 scenes of 10k-2M triangles, 64 views of
    128^2-1024^2, depth 6, RR from bounce 3, 1 spp): one eager frame and one
    graph-replayed frame, every view bit-identical to its own snkpt_render,
    and the smallest scene's smallest view per pixel against the CPU oracle."""
    import torch

    import snkpt
    from oracle import binding as O
    scenes, ctxs = [], []
    for v in range(16):
        sd = snkpt.SceneData(5, v)
        c = snkpt.Context(0)
        c.load_scene(sd)
        c.build_accel()
        scenes.append(sd)
        ctxs.append(c)
    tris = [int(sd.s.n_idx // 3) for sd in scenes]
    _say(f"config 5: 16 scenes, {min(tris)}..{max(tris)} triangles")
    b = snkpt.Batch(0)
    for c in ctxs:
        b.add_scene(c)
    views, outs = [], []
    for view in range(64):
        var, w, h = snkpt.scene_view(5, view)
        cam = scenes[var].camera(view, w, h)
        out = torch.zeros(h, w, 4, device="cuda")
        b.add_view(var, cam, w, h, 0x5000 + view, out)
        views.append((var, w, h, cam))
        outs.append(out)
    assert sum(w * h for _, w, h, _ in views) > 10_000_000
    st = torch.cuda.Stream()
    for frame, graph in ((0, False), (1, True)):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        b.render(_settings(frame=frame), use_graph=graph, stream=st)
        st.synchronize()
        got = _grab(outs)
        for k, (var, w, h, cam) in enumerate(views):
            o = torch.zeros(h, w, 4, device="cuda")
            ctxs[var].render(cam, _settings(seed=0x5000 + k, frame=frame), o, w, h)
            want = o.cpu().numpy()
            assert np.array_equal(got[k], want), (frame, graph, k, np.abs(got[k] - want).max())
        _say(f"frame {frame} (graph={graph}): 64 views bit-identical to single-view renders")
        assert all(np.isfinite(g).all() for g in got)
    k = min(range(64), key=lambda k: (tris[views[k][0]], views[k][1] * views[k][2]))
    var, w, h, cam = views[k]
    arrs = O.scene_arrays(scenes[var])
    c = O.pt_render(arrs, O.Bvh(O.flatten(*arrs[:4])), cam, w, h, 1, 6, 3, 0x5000 + k, frame=1)
    g = got[k][..., :3]
    frac = (np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-5).all(-1).mean()
    _say(f"view {k} ({w}x{h}, scene {var}, {tris[var]} tris) vs oracle: {frac:.4f} of the pixels agree")
    assert frac >= 0.97, frac
    b.close()


def test_batch_follows_scene_rebuilds_between_frames(gpu):
    """A scene context rebuilt with moved instances and given lights between
    two graph-replayed batch frames: the batch sees the new generation,
    rebuilds its tables and graph, and every view still equals its own
    snkpt_render of the changed scene."""
    import torch

    import snkpt
    scenes, ctxs, b, outs, cams = _setup()
    st = torch.cuda.Stream()
    b.render(_settings(frame=0), use_graph=True, stream=st)
    st.synchronize()
    want0 = _single(ctxs, cams, 0)
    got0 = _grab(outs)
    for k in range(len(VIEWS)):
        assert np.array_equal(got0[k], want0[k]), k
    g0 = ctxs[0].generation()
    pos, idx, inst, m4, mats = scenes[0].arrays()
    moved = np.asarray(m4, np.float32).reshape(-1, 16).copy()
    moved[:, 12] += 0.15  # column-major x translation of every transform
    ctxs[0].set_instances(inst, moved)
    ctxs[0].set_lights(snkpt.light_block(points=[((4.0, 3.0, 2.0), (0.2, 1.6, 0.1), (0.0, 0.0, 1.0))]), None)
    ctxs[0].build_accel()
    ctxs[2].set_lights(snkpt.light_block(dir_colour=(0.5, 0.5, 0.5), dir_to_light=(0.3, 1.0, 0.2)), None)
    ctxs[2].build_accel()
    assert ctxs[0].generation() != g0
    for frame in (1, 2):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        b.render(_settings(frame=frame), use_graph=True, stream=st)
        st.synchronize()
        got = _grab(outs)
        want = _single(ctxs, cams, frame)
        for k in range(len(VIEWS)):
            assert np.array_equal(got[k], want[k]), (frame, k, np.abs(got[k] - want[k]).max())
    moved_views = [k for k, v in enumerate(VIEWS) if v[0] in (0, 2)]
    assert any(not np.array_equal(want0[k], _single(ctxs, cams, 0)[k]) for k in moved_views)
    _say("batch after scene rebuilds: views bit-identical to single-view renders")
    b.close()
