"""GPU parity of the path-trace pass (plain PT mode) against the CPU oracle
path tracer and against closed forms."""
import sys

import numpy as np
import pytest

from helpers import floor_light_scene, look_at_camera, polygon_irradiance
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _render(ctx, cam, w, h, **kw):
    import torch

    import snkpt
    fmt = kw.pop("fmt", snkpt.FORMAT_RGBA32F)
    s = snkpt.default_settings(**kw)
    dt = torch.float16 if fmt == snkpt.FORMAT_RGBA16F else torch.float32
    out = torch.zeros(h, w, 4, device="cuda", dtype=dt)
    ctx.render(cam, s, out, w, h, fmt)
    torch.cuda.synchronize()
    return out.float().cpu().numpy()


def test_room_pt_matches_cpu_oracle_per_pixel(gpu):
    import snkpt
    _say("room pt vs oracle")
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w = h = 64
    cam = scene.camera(0, w, h)
    g = _render(ctx, cam, w, h, spp=4, max_depth=4, rr_start=0, seed=7)[..., :3]
    pos, idx, inst, m4, mats = scene.arrays()
    b = O.Bvh(O.flatten(pos, idx, inst, m4))
    c = O.pt_render(O.scene_arrays(scene), b, cam, w, h, 4, 4, 0, 7)
    assert np.isfinite(g).all() and g.mean() > 0.01
    close = np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-5
    frac = close.all(-1).mean()
    _say(f"pixels agreeing: {frac:.4f}; means {g.mean():.6f} vs {c.mean():.6f}")
    assert frac >= 0.97, frac
    assert abs(g.mean() - c.mean()) <= 0.01 * c.mean()
    st = ctx.frame_stats()
    assert st.camera_rays == w * h * 4 and st.extension_rays > 0 and st.shadow_rays > 0


def test_direct_light_closed_form_gpu(gpu):
    """Floor radiance under a square emitter equals albedo/pi * E (Lambert's
    polygon formula) with NEE alone and with NEE + BRDF sampling under MIS."""
    import snkpt
    sc, verts = floor_light_scene()
    ctx = snkpt.Context(0)
    sc.load_into(ctx)
    w = h = 8
    eye = np.array([0.3, 1.0, 0.2])
    x0 = np.array([0.3, 0.0, -0.2])
    cam = look_at_camera(eye, x0, fov=4.0, w=w, h=h)
    import raysets
    org, dr = raysets.primary_rays(cam, w, h)
    t = -org[:, 1] / dr[:, 1]
    xs = org[:, :3] + dr[:, :3] * t[:, None]
    want = np.array([0.5 / np.pi * polygon_irradiance(x, [0, 1, 0], verts, 10.0) for x in xs]).reshape(h, w)
    for depth in (1, 2):
        img = _render(ctx, cam, w, h, spp=4000, max_depth=depth, seed=11)[..., 0]
        assert np.allclose(img.mean(), want.mean(), rtol=0.005), (depth, img.mean(), want.mean())
        assert np.allclose(img, want, rtol=0.05), depth


def test_oracle_accumulate_f64_matches_render_in_expectation(gpu):
    import torch

    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w = h = 32
    cam = scene.camera(0, w, h)
    s = snkpt.default_settings(max_depth=4, seed=3)
    acc = torch.zeros(h * w * 3, dtype=torch.float64, device="cuda")
    ctx.oracle_accumulate(cam, s, acc, w, h, 256)
    torch.cuda.synchronize()
    o = (acc / 256).view(h, w, 3).cpu().numpy()
    r = _render(ctx, cam, w, h, spp=256, max_depth=4, seed=99)[..., :3]
    # tile means agree within a few standard errors
    to = o.reshape(4, 8, 4, 8, 3).mean((1, 3))
    tr = r.reshape(4, 8, 4, 8, 3).mean((1, 3))
    assert np.allclose(to, tr, rtol=0.08, atol=2e-3)
    assert abs(o.mean() - r.mean()) < 0.02 * o.mean()


def test_render_deterministic_and_half_output(gpu):
    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w = h = 48
    cam = scene.camera(0, w, h)
    a = _render(ctx, cam, w, h, spp=2, max_depth=4, seed=5)
    b = _render(ctx, cam, w, h, spp=2, max_depth=4, seed=5)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    c = _render(ctx, cam, w, h, spp=2, max_depth=4, seed=5, fmt=snkpt.FORMAT_RGBA16F)
    assert np.allclose(c, a, rtol=2e-3, atol=1e-3)
    # accumulation: the running mean of two frames equals the mean of the two images
    f0 = _render(ctx, cam, w, h, spp=1, max_depth=4, seed=5, frame=0, flags=snkpt.FLAG_ACCUMULATE | snkpt.FLAG_RESET)
    f1 = _render(ctx, cam, w, h, spp=1, max_depth=4, seed=5, frame=1, flags=snkpt.FLAG_ACCUMULATE)
    s0 = _render(ctx, cam, w, h, spp=1, max_depth=4, seed=5, frame=0)
    s1 = _render(ctx, cam, w, h, spp=1, max_depth=4, seed=5, frame=1)
    assert np.allclose(f0, s0) and np.allclose(f1[..., :3], 0.5 * (s0 + s1)[..., :3], rtol=1e-5, atol=1e-6)


def test_spheres_1m_frame_renders(gpu):
    """Config 2 at a reduced resolution: finite, non-black, RR active."""
    import snkpt
    scene = snkpt.SceneData(2)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w, h = 480, 270
    img = _render(ctx, scene.camera(0, w, h), w, h, spp=1, max_depth=8, rr_start=3, seed=1)
    assert np.isfinite(img).all() and img[..., :3].mean() > 0


@pytest.mark.parametrize("config,w,h", [(1, 96, 64), (4, 160, 90), (2, 128, 72)])
def test_whole_path_kernel_matches_wavefront(gpu, config, w, h):
    """One thread per whole path (no queues) gives the wavefront's image bit
    for bit: same vertex logic, random stream and order of additions."""
    import torch

    import snkpt
    scene = snkpt.SceneData(config)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    cam = scene.camera(0, w, h)
    outs = []
    for flags in (0, snkpt.FLAG_WHOLE_PATHS):
        o = torch.zeros(h, w, 4, device="cuda")
        s = snkpt.default_settings(max_depth=scene.s.max_depth, rr_start=scene.s.rr_start, seed=5, frame=2,
                                   flags=flags, spp=2)
        ctx.render(cam, s, o, w, h)
        st = ctx.frame_stats()
        outs.append((o.cpu().numpy(), st.extension_rays, st.shadow_rays))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1:] == outs[1][1:]


@pytest.mark.parametrize("config,w,h", [(1, 96, 64), (4, 160, 90), (2, 128, 72), (4, 640, 360)])
def test_per_thread_tail_matches_wavefront(gpu, config, w, h):
    """From any bounce on, the paths still alive can run one thread per path
    instead of a launch per bounce: the image and the ray counts stay those of
    the pure wavefront bit for bit, for every tail bounce and for the bounce the
    context picks itself from the previous frame's queue lengths."""
    import torch

    import snkpt
    never = 0xFFFFFFFF
    scene = snkpt.SceneData(config)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    cam = scene.camera(0, w, h)
    md = scene.s.max_depth

    def run(tail, frame):
        o = torch.zeros(h, w, 4, device="cuda")
        s = snkpt.default_settings(max_depth=md, rr_start=scene.s.rr_start, seed=5, frame=frame, spp=2,
                                   tail_bounce=tail)
        ctx.render(cam, s, o, w, h)
        st = ctx.frame_stats()
        return o.cpu().numpy(), st.extension_rays, st.shadow_rays

    ref = run(never, 2)
    assert ref[1] > 0 and ref[2] > 0
    for t in range(1, md):
        got = run(t, 2)
        assert np.array_equal(ref[0], got[0]), t
        assert ref[1:] == got[1:], t
    for f in range(4):
        want = run(never, f)
        got = run(0, f)
        assert np.array_equal(want[0], got[0]) and want[1:] == got[1:], f


def test_automatic_tail_follows_resolution_and_depth_changes(gpu):
    """The context's tail bounce comes from its previous frame; frames of other
    sizes and depths in between (which reset or move it) still give the pure
    wavefront's image and ray counts."""
    import torch

    import snkpt
    scene = snkpt.SceneData(4)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()

    def run(tail, w, h, depth, frame):
        o = torch.zeros(h, w, 4, device="cuda")
        s = snkpt.default_settings(max_depth=depth, rr_start=2, seed=9, frame=frame, tail_bounce=tail)
        ctx.render(scene.camera(0, w, h), s, o, w, h)
        st = ctx.frame_stats()
        return o.cpu().numpy(), st.extension_rays, st.shadow_rays

    for f, (w, h, depth) in enumerate([(160, 90, 8), (96, 64, 8), (96, 64, 3), (320, 180, 6), (160, 90, 8)]):
        got = run(0, w, h, depth, f)
        want = run(0xFFFFFFFF, w, h, depth, f)
        assert np.array_equal(got[0], want[0]) and got[1:] == want[1:], (w, h, depth)


def test_weighted_accumulation_is_the_weighted_mean(gpu):
    """SNKPT_FLAG_ACCUMULATE with accum_weight w_f: the output is
    sum w_f x_f / sum w_f over the accumulated frames."""
    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    w = h = 32
    cam = scene.camera(0, w, h)
    wts = [1.0, 4.0, 9.0]
    singles = [_render(ctx, cam, w, h, spp=1, max_depth=4, seed=5, frame=f)[..., :3].astype(np.float64)
               for f in range(3)]
    for f, wt in enumerate(wts):
        fl = snkpt.FLAG_ACCUMULATE | (snkpt.FLAG_RESET if f == 0 else 0)
        got = _render(ctx, cam, w, h, spp=1, max_depth=4, seed=5, frame=f, flags=fl, accum_weight=wt)[..., :3]
    want = sum(wt * x for wt, x in zip(wts, singles)) / sum(wts)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
