"""CPU tests pinning the contact-record restatement (ParticleUpdate.rchit)
with closed forms on a rotated, scaled, translated quad."""
import numpy as np

from oracle import binding as O
from oracle import contacts as K

F = np.float32


def _quad_scene(m4):
    # unit quad in the object's x-z plane, normal +y, one instance with transform m4
    pos = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], F)
    nrm = np.tile(np.array([[0, 1, 0]], F), (4, 1))
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    inst = np.array([[0, 0, 0, 0, 6]], np.uint32)
    mats = np.zeros((1, 16), np.uint32)
    return pos, idx, inst, m4.reshape(1, 16).astype(F), mats, nrm, np.zeros(0, np.uint32)


def _transform(angle, scale, shift):
    c, s = np.cos(angle), np.sin(angle)
    rot = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])  # about z: tilts the quad's normal
    m = np.eye(4)
    m[:3, :3] = rot @ np.diag([scale, 1.0, scale])
    m[:3, 3] = shift
    return m.T.reshape(-1).astype(F), rot  # column-major


def test_contact_record_closed_forms():
    angle, shift = 0.3, np.array([0.2, 1.0, -0.1])
    m4, rot = _transform(angle, 2.0, shift)
    arrs = _quad_scene(m4)
    b = O.Bvh(O.flatten(arrs[0], arrs[1], arrs[2], arrs[3]))
    n_world = rot @ np.array([0, 1.0, 0])
    # rays toward the plane's points x0 from 0.5 along the normal
    rng = np.random.default_rng(1)
    pts = [rot @ np.array([rng.uniform(-1.5, 1.5), 0, rng.uniform(-1.5, 1.5)]) + shift for _ in range(32)]
    rays = np.zeros((32, 8), F)
    for k, x in enumerate(pts):
        rays[k, :3] = x + 0.5 * n_world
        rays[k, 4:7] = -n_world
        rays[k, 7] = 1.0
    t, inst, prim, u, v, _, _ = b.trace(rays, any_hit=True)
    assert (t >= 0).all() and np.allclose(t, 0.5, atol=2e-6)
    # the previous frame had the quad 0.05 lower along world y: velocity = (0, 0.05, 0) / dt
    prev_m = m4.reshape(4, 4).copy()
    prev_m[3, 1] -= 0.05
    dt = 1.0 / 120.0
    nw, cp, vel = K.contacts(arrs, prev_m.reshape(1, 16), rays, t, inst, prim, u, v, dt)
    assert np.allclose(nw, n_world, atol=1e-6)
    assert np.allclose(cp, np.array(pts), atol=1e-5)
    assert np.allclose(vel, [0, 0.05 / dt, 0], atol=2e-3)
    # static: previous == current -> velocity ~ 0 (float rounding of the two paths only)
    _, _, vel0 = K.contacts(arrs, m4.reshape(1, 16), rays, t, inst, prim, u, v, dt)
    assert np.abs(vel0).max() < 1e-3


def test_misses_are_zero():
    m4, _ = _transform(0.0, 1.0, np.zeros(3))
    arrs = _quad_scene(m4)
    b = O.Bvh(O.flatten(arrs[0], arrs[1], arrs[2], arrs[3]))
    rays = np.array([[0, 1, 0, 0, 0, 1, 0, 5.0], [5, 1, 5, 0, 0, -1, 0, 5.0], [0, 1, 0, 0, 0, -1, 0, 0.5]], F)
    t, inst, prim, u, v, _, _ = b.trace(rays, any_hit=True)
    assert (t < 0).all()  # away from the quad, outside it, and too short
    nw, cp, vel = K.contacts(arrs, m4.reshape(1, 16), rays, t, inst, prim, u, v, 0.01)
    assert not nw.any() and not cp.any() and not vel.any()
