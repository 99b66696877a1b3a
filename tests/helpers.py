"""Shared helpers for the GPU parity tests."""
import numpy as np

from oracle import binding as O


def scene_world(scene):
    pos, idx, inst, m4, mats = scene.arrays()
    return O.flatten(pos, idx, inst, m4)


def make_ctx(scene=None):
    import snkpt
    ctx = snkpt.Context(0)
    if scene is not None:
        ctx.load_scene(scene)
        ctx.build_accel()
    return ctx


def tri_soup_ctx(pos, idx):
    """Context holding one identity-transformed instance of (pos, idx)."""
    import snkpt
    ctx = snkpt.Context(0)
    ctx.upload_mesh(pos, idx)
    ctx.set_instances(np.array([[0, 0, 0, 0, idx.size]], np.uint32), np.eye(4, dtype=np.float32).reshape(1, 16))
    m = snkpt.Material()
    m.albedo_tex_idx = m.normal_tex_idx = m.roughness_tex_idx = m.metallic_tex_idx = m.ao_tex_idx = 0xFFFFFFFF
    m.albedo[:] = [0.5, 0.5, 0.5, 1.0]
    m.roughness = 1.0
    ctx.set_materials([m])
    ctx.set_lights(None, None)
    ctx.build_accel()
    return ctx


def gpu_trace(ctx, rays_np, any_hit=False, counted=True):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays_np, np.float32).reshape(-1, 8)).cuda()
    org = r[:, :4].contiguous()
    dr = r[:, 4:].contiguous()
    out = ctx.trace(org, dr, any_hit=any_hit, counted=counted)
    torch.cuda.synchronize()
    return [x.cpu().numpy() if x is not None else None for x in out]


def as_bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bits_equal(a, b, what):
    a, b = as_bits(a), as_bits(b)
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{what}: {bad.size} of {a.size} differ, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"
