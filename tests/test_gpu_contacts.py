"""Short collision rays with contact records (snkpt_trace_contacts;
ParticleUpdate.rgen / .rchit) on the GPU: the first hit equals the oracle's
any-hit traversal bit for bit, the contact record equals its numpy
restatement, a moving instance gets the closed-form velocity."""
import sys

import numpy as np
import pytest

from oracle import binding as O
from oracle import contacts as K

pytestmark = pytest.mark.gpu
F = np.float32


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _room():
    import snkpt
    scene = snkpt.SceneData(1)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    arrs = O.scene_arrays(scene)
    b = O.Bvh(O.flatten(arrs[0], arrs[1], arrs[2], arrs[3]))
    return scene, ctx, arrs, b


def _particle_rays(n, seed):
    """Particle-style rays inside the room: from the back of a particle of
    radius r along a random direction, tmax = speed * dt + 2 r (ParticleUpdate.rgen:182)."""
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(-0.95, 0.95, n), rng.uniform(0.05, 1.95, n), rng.uniform(-0.95, 0.95, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.01, 0.08, n)
    speed = rng.uniform(0.0, 30.0, n)
    rays = np.zeros((n, 8), F)
    rays[:, :3] = p - d * r[:, None]
    rays[:, 4:7] = d
    rays[:, 7] = speed / 120.0 + 2 * r
    return rays


def _trace(ctx, rays, dt):
    import torch
    o = torch.from_numpy(np.ascontiguousarray(rays[:, :4])).cuda()
    d = torch.from_numpy(np.ascontiguousarray(rays[:, 4:])).cuda()
    out = ctx.trace_contacts(o, d, dt)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in out]


def test_contacts_match_oracle_any_hit_and_restatement(gpu):
    scene, ctx, arrs, b = _room()
    rays = _particle_rays(20000, 3)
    dt = 1.0 / 120.0
    t, nw, cp, vel, inst, prim = _trace(ctx, rays, dt)
    ct, ci, cpr, cu, cv, _, _ = b.trace(rays, any_hit=True)
    hit = ct >= 0
    _say(f"{int(hit.sum())} of {len(rays)} collision rays hit")
    assert hit.sum() > 1000
    assert np.array_equal(t.view(np.uint32), ct.view(np.uint32))
    assert np.array_equal(inst.view(np.uint32), ci) and np.array_equal(prim.view(np.uint32), cpr)
    kn, kp, kv = K.contacts(arrs, arrs[3], rays, ct, ci, cpr, cu, cv, dt)
    assert np.allclose(nw[:, :3], kn, atol=2e-6)
    assert np.allclose(cp[:, :3], kp, atol=1e-6)
    assert np.allclose(vel[:, :3], kv, atol=1e-3)  # (cp - M p) / dt: 120x the rounding of either point
    assert not nw[~hit].any() and not vel[~hit].any() and (inst.view(np.uint32)[~hit] == 0xFFFFFFFF).all()


def test_contacts_on_the_deep_bvh_match_oracle(gpu):
    """The 1.02M-triangle scene (BVH8 depth 12: the 12-entry stack variant, held
    to 80 VGPRs): particle rays over its volume, first hit bit for bit against the
    oracle's any-hit traversal, contact record against the restatement."""
    import snkpt
    scene = snkpt.SceneData(2)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    ctx.build_accel()
    assert 11 <= ctx.accel_info().depth <= 13
    arrs = O.scene_arrays(scene)
    b = O.Bvh(O.flatten(arrs[0], arrs[1], arrs[2], arrs[3]))
    n = 200000
    rng = np.random.default_rng(11)
    p = np.stack([rng.uniform(-50, 50, n), rng.uniform(0.0, 12.0, n), rng.uniform(-50, 50, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.01, 0.08, n)
    rays = np.zeros((n, 8), F)
    rays[:, :3] = p - d * r[:, None]
    rays[:, 4:7] = d
    rays[:, 7] = rng.uniform(0.0, 30.0, n) / 120.0 + 2 * r
    dt = 1.0 / 120.0
    t, nw, cp, vel, inst, prim = _trace(ctx, rays, dt)
    ct, ci, cpr, cu, cv, _, _ = b.trace(rays, any_hit=True)
    hit = ct >= 0
    _say(f"deep BVH: {int(hit.sum())} of {n} collision rays hit")
    assert hit.sum() > 1000
    assert np.array_equal(t.view(np.uint32), ct.view(np.uint32))
    assert np.array_equal(inst.view(np.uint32), ci) and np.array_equal(prim.view(np.uint32), cpr)
    kn, kp, kv = K.contacts(arrs, arrs[3], rays, ct, ci, cpr, cu, cv, dt)
    assert np.allclose(nw[:, :3], kn, atol=2e-6)
    assert np.allclose(cp[:, :3], kp, atol=1e-4)  # world coordinates up to 50 m
    assert not nw[~hit].any() and not vel[~hit].any()


def test_moving_instance_velocity(gpu):
    """The previous frame had instance k's transform 0.02 lower along y: its
    contacts move at (0, 0.02 / dt, 0), everything else is at rest."""
    scene, ctx, arrs, b = _room()
    inst5, m4 = arrs[2], np.asarray(arrs[3], F).reshape(-1, 16)
    k = int(np.argmax(inst5[:, 4]))  # the instance with the most triangles (the sphere)
    ti = int(inst5[k, 0])
    assert (inst5[:, 0] == ti).sum() == 1, "transform shared with another instance"
    prev = m4.copy()
    prev[ti, 13] -= 0.02  # column-major: translation y
    ctx.set_prev_transforms(prev)
    dt = 1.0 / 120.0
    rays = _particle_rays(20000, 8)
    t, nw, cp, vel, inst, prim = _trace(ctx, rays, dt)
    on = (t >= 0) & (inst == k)
    off = (t >= 0) & (inst != k)
    _say(f"hits on the moving instance: {int(on.sum())}, elsewhere {int(off.sum())}")
    assert on.sum() > 20 and off.sum() > 100
    assert np.allclose(vel[on, :3], [0, 0.02 / dt, 0], atol=2e-3)
    assert np.abs(vel[off, :3]).max() < 2e-3
    kn, kp, kv = K.contacts(arrs, prev, rays, *b.trace(rays, any_hit=True)[:5], dt)
    assert np.allclose(vel[:, :3], kv, atol=1e-3)


def test_contacts_reject_bad_input(gpu):
    import torch

    import snkpt
    scene, ctx, arrs, b = _room()
    o = torch.zeros(8, 4, device="cuda")
    with pytest.raises(snkpt.SnkptError):
        ctx.trace_contacts(o, o, 0.0)  # timestep must be > 0
    with pytest.raises(snkpt.SnkptError):
        ctx.set_prev_transforms(np.eye(4, dtype=F).reshape(1, 16).repeat(3, 0))  # wrong count
    empty = snkpt.Context(0)
    with pytest.raises(snkpt.SnkptError):
        empty.trace_contacts(o, o)  # nothing built
