"""SNKPT_FLAG_NO_STATS: snkpt_render without its end-of-frame readback.  The
frame is the same image; a plain PT frame recorded into a HIP graph by stream
capture renders that image again on every replay, with uncaptured frames of
the same context run between the replays; what cannot be captured is refused
before anything is launched."""
import sys

import numpy as np
import pytest

from helpers import make_ctx

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _settings(**kw):
    import snkpt
    kw.setdefault("max_depth", 4)
    return snkpt.default_settings(**kw)


def test_no_stats_frame_is_the_same_image_and_keeps_the_stats(gpu):
    import torch

    import snkpt
    _say("capture: no-stats frame, room 64x48")
    scene = snkpt.SceneData(1)
    ctx = make_ctx(scene)
    w, h = 64, 48
    cam = scene.camera(0, w, h)
    a = torch.zeros(h, w, 4, device="cuda")
    b = torch.zeros_like(a)
    ctx.render(cam, _settings(seed=7), a, w, h)
    torch.cuda.synchronize()
    st = ctx.frame_stats()
    assert st.camera_rays == w * h and st.extension_rays > 0
    ctx.render(cam, _settings(seed=9, max_depth=2), b, w, h)
    torch.cuda.synchronize()
    st2 = ctx.frame_stats()
    ctx.render(cam, _settings(seed=7, flags=snkpt.FLAG_NO_STATS), a.new_zeros(a.shape), w, h)
    c = a.new_zeros(a.shape)
    ctx.render(cam, _settings(seed=7, flags=snkpt.FLAG_NO_STATS), c, w, h)
    torch.cuda.synchronize()
    # the no-stats frames left the last readback (seed 9, depth 2) in place
    st3 = ctx.frame_stats()
    assert (st3.camera_rays, st3.extension_rays, st3.shadow_rays) == \
        (st2.camera_rays, st2.extension_rays, st2.shadow_rays)
    assert torch.equal(a, c)


@pytest.mark.parametrize("whole_paths", [False, True])
def test_captured_pt_frame_replays_match_the_uncaptured_frame(gpu, whole_paths):
    import torch

    import snkpt
    _say(f"capture: PT graph replays, whole_paths={whole_paths}")
    scene = snkpt.SceneData(1)
    ctx = make_ctx(scene)
    w, h = 96, 64
    cam = scene.camera(0, w, h)
    fl = snkpt.FLAG_NO_STATS | (snkpt.FLAG_WHOLE_PATHS if whole_paths else 0)
    s = _settings(seed=21, frame=3, flags=fl)
    want = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, s, want, w, h)  # uncaptured: sizes the context's buffers
    torch.cuda.synchronize()
    assert float(want[..., :3].sum()) > 0.0
    got = torch.zeros_like(want)
    other = torch.zeros_like(want)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.render(cam, s, got, w, h)
    torch.cuda.synchronize()
    for k in range(4):
        got.zero_()
        g.replay()
        torch.cuda.synchronize()
        ok = torch.equal(got, want)
        _say(f"capture: replay {k} equal={ok}")
        assert ok
        # uncaptured work on the same context between replays: another seed
        ctx.render(cam, _settings(seed=100 + k), other, w, h)
        torch.cuda.synchronize()
        assert not torch.equal(other, want)


def test_capture_refuses_other_modes_and_unsized_frames(gpu):
    import torch

    import snkpt
    _say("capture: refusals")
    scene = snkpt.SceneData(1)
    ctx = make_ctx(scene)
    w, h = 32, 32
    cam = scene.camera(0, w, h)
    out = torch.zeros(h, w, 4, device="cuda")
    ctx.render(cam, _settings(seed=1, flags=snkpt.FLAG_NO_STATS), out, w, h)
    torch.cuda.synchronize()
    big = torch.zeros(2 * h, 2 * w, 4, device="cuda")
    cases = [
        (_settings(seed=1), out, w, h, "without the no-stats flag"),
        (_settings(seed=1, mode=snkpt.MODE_RESTIR, max_depth=1, flags=snkpt.FLAG_NO_STATS), out, w, h, "ReSTIR"),
        (_settings(seed=1, flags=snkpt.FLAG_NO_STATS), big, 2 * w, 2 * h, "larger than any uncaptured frame"),
        (_settings(seed=1, flags=snkpt.FLAG_NO_STATS, frame=0), big, w // 2, h, "smaller than the last frame"),
        (_settings(seed=1, flags=snkpt.FLAG_NO_STATS | snkpt.FLAG_ACCUMULATE), big, w, h // 2,
         "another shape of the same size"),
    ]
    for s, o, ww, hh, what in cases:
        _say(f"capture: refuse {what}")
        g = torch.cuda.CUDAGraph()
        with pytest.raises(snkpt.SnkptError):
            with torch.cuda.graph(g):
                ctx.render(cam, s, o, ww, hh)
        torch.cuda.synchronize()
    # the context still renders uncaptured afterwards
    again = torch.zeros_like(out)
    ctx.render(cam, _settings(seed=1, flags=snkpt.FLAG_NO_STATS), again, w, h)
    torch.cuda.synchronize()
    assert torch.equal(again, out)
    assert np.isfinite(again.cpu().numpy()).all()


def test_generation_advances_on_every_state_change(gpu):
    """snkpt_get_generation changes after each call that may move or change what
    a captured frame or a batch reads (scene input, build, render-buffer
    growth, cache reset), and stays put across frames of one size."""
    import torch

    import snkpt
    from helpers import floor_light_scene
    scene = snkpt.SceneData(1)
    ctx = make_ctx(scene)
    w, h = 32, 32
    cam = scene.camera(0, w, h)
    out = torch.zeros(2 * h, 2 * w, 4, device="cuda")
    g = [ctx.generation()]

    def step(what, fn):
        fn()
        torch.cuda.synchronize()
        g.append(ctx.generation())
        _say(f"generation after {what}: {g[-1]}")
        assert g[-1] != g[-2], what
    step("first frame", lambda: ctx.render(cam, _settings(seed=1), out, w, h))
    before = ctx.generation()
    ctx.render(cam, _settings(seed=2), out, w, h)
    ctx.render(cam, _settings(seed=3, flags=snkpt.FLAG_NO_STATS), out, w, h)
    torch.cuda.synchronize()
    assert ctx.generation() == before, "frames of one size keep the generation"
    step("a larger frame", lambda: ctx.render(cam, _settings(seed=1), out, 2 * w, 2 * h))
    pos, idx, inst, m4, mats = scene.arrays()
    step("set_instances", lambda: ctx.set_instances(inst, m4))
    step("build_accel", lambda: ctx.build_accel())
    ms = [snkpt.Material.from_buffer_copy(m.tobytes()) for m in mats]
    step("set_materials", lambda: ctx.set_materials(ms))
    step("set_lights", lambda: ctx.set_lights(snkpt.light_block(points=[((1, 1, 1), (0, 1.5, 0), (0, 0, 1))]), None))
    sc, _ = floor_light_scene()
    step("upload_mesh", lambda: ctx.upload_mesh(*sc.arrays()[:2]))
    step("cache_reset", lambda: ctx.cache_reset(1, [-1, 0, -1], [1, 2, 1]))


@pytest.mark.parametrize("form", ["lo_q2", "indirect_mis_q1"])
def test_captured_cache_frame_replays_match_eager_frames(gpu, form):
    """A cache-mode frame (render, queries, training paths on the side stream,
    optimiser step) recorded by stream capture: the training batch size and the
    Adam step counter live on the device, so each replay trains one more step
    exactly as an uncaptured frame with the same settings does.  Two contexts
    with the same cache seed run the same frames, one eagerly and one as
    replays, and give the same bits after every frame."""
    import torch

    import snkpt
    _say(f"capture: cache-mode graph replays, form {form}")
    scene = snkpt.SceneData(1)
    w, h = 96, 64
    cam = scene.camera(0, w, h)
    kw = dict(mode=snkpt.MODE_CACHE, seed=5, cache_train_paths=2048, cache_lr=2.5e-3, cache_ema=0.0,
              flags=snkpt.FLAG_NO_STATS)
    if form == "lo_q2":
        kw["cache_query_depth"] = 2
    else:
        kw["cache_query_depth"] = 1
        kw["flags"] |= snkpt.FLAG_CACHE_INDIRECT | snkpt.FLAG_CACHE_MIS
    eager, capt = make_ctx(scene), make_ctx(scene)
    a = torch.zeros(h, w, 4, device="cuda")
    b = torch.zeros_like(a)
    for ctx, o in ((eager, a), (capt, b)):
        ctx.render(cam, _settings(frame=0, **kw), o, w, h)  # uncaptured: sizes the buffers, first step
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    s1 = _settings(frame=1, **kw)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        capt.render(cam, s1, b, w, h)
    torch.cuda.synchronize()
    outs = []
    for k in range(4):
        eager.render(cam, s1, a, w, h)
        b.zero_()
        g.replay()
        torch.cuda.synchronize()
        ok = torch.equal(a, b)
        _say(f"capture: cache replay {k} equal={ok}")
        assert ok
        assert torch.isfinite(b).all()
        outs.append(b.clone())
    # every replay trained the network: the same frame's image keeps changing
    assert not torch.equal(outs[0], outs[-1])
    # both contexts' networks are the same after the same number of steps
    wa = eager.cache_params(0)
    wb = capt.cache_params(0)
    assert np.array_equal(wa, wb)


def test_capture_refuses_an_unsized_cache_frame(gpu):
    import torch

    import snkpt
    _say("capture: cache-mode refusals")
    scene = snkpt.SceneData(1)
    ctx = make_ctx(scene)
    w, h = 64, 48
    cam = scene.camera(0, w, h)
    kw = dict(mode=snkpt.MODE_CACHE, seed=5, cache_train_paths=1024, flags=snkpt.FLAG_NO_STATS)
    out = torch.zeros(h, w, 4, device="cuda")
    g = torch.cuda.CUDAGraph()
    with pytest.raises(snkpt.SnkptError):  # no uncaptured cache frame yet
        with torch.cuda.graph(g):
            ctx.render(cam, _settings(max_depth=4, **kw), out, w, h)
    torch.cuda.synchronize()
    ctx.render(cam, _settings(max_depth=4, **kw), out, w, h)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(snkpt.SnkptError):  # deeper paths need larger training buffers
        with torch.cuda.graph(g):
            ctx.render(cam, _settings(max_depth=8, **kw), out, w, h)
    torch.cuda.synchronize()
    again = torch.zeros_like(out)
    ctx.render(cam, _settings(max_depth=4, **kw), again, w, h)
    torch.cuda.synchronize()
    assert torch.isfinite(again).all()
