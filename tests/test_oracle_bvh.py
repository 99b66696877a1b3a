"""CPU tests: the oracle is pinned against known answers, closed forms and
invariants before it is trusted as the GPU checker.  No GPU needed."""
import numpy as np
import pytest

from oracle import binding as O


def test_pcg32_known_answer():
    # pcg32 reference demo (O'Neill, pcg-c-basic): pcg32_srandom(42, 54)
    got = O.pcg32(42, 54, 6)
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert [int(x) for x in got] == want


def test_intersect_closed_form():
    v0, v1, v2 = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    hit, t, u, v = O.intersect((0.25, 0.25, 1.0), (0, 0, -1), v0, v1, v2)
    assert hit and t == 1.0 and u == 0.25 and v == 0.25
    # back face is hit too (no cull flags in the reference)
    hit, t, u, v = O.intersect((0.25, 0.5, -2.0), (0, 0, 1), v0, v1, v2)
    assert hit and t == 2.0 and u == 0.25 and v == 0.5
    # outside
    assert not O.intersect((0.8, 0.8, 1.0), (0, 0, -1), v0, v1, v2)[0]
    # parallel ray in the triangle's plane: no hit
    assert not O.intersect((-1.0, 0.25, 0.0), (1, 0, 0), v0, v1, v2)[0]
    # degenerate triangle
    assert not O.intersect((0.5, 0.0, 1.0), (0, 0, -1), v0, v1, (2, 0, 0))[0]


def test_intersect_t_matches_double_precision():
    rng = np.random.default_rng(7)
    worst = 0.0
    n_hit = 0
    for _ in range(3000):
        tri = rng.uniform(-1, 1, (3, 3)).astype(np.float32).astype(np.float64)
        o = rng.uniform(-3, 3, 3).astype(np.float32).astype(np.float64)
        target = tri.T @ rng.dirichlet([1, 1, 1])  # a point inside the triangle
        d = target - o
        d = (d / np.linalg.norm(d)).astype(np.float32).astype(np.float64)
        hit, t, u, v = O.intersect(o, d, tri[0], tri[1], tri[2])
        # exact (double) ray/plane distance
        nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        denom = nrm @ d
        if abs(denom) < 1e-2 * np.linalg.norm(nrm):
            continue
        assert hit
        n_hit += 1
        t_exact = nrm @ (tri[0] - o) / denom
        worst = max(worst, abs(t - t_exact) / abs(t_exact))
        p = o + t * d
        q = (1 - u - v) * tri[0] + u * tri[1] + v * tri[2]
        assert np.allclose(p, q, atol=1e-4)
    assert n_hit > 2500
    assert worst < 1e-5


def _grid_mesh(nx, ny, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1), indexing="ij")
    z = rng.uniform(-0.05, 0.05, xs.shape)
    jitter = rng.uniform(-0.2, 0.2, (2,) + xs.shape) / max(nx, ny)
    jitter[:, 0, :] = jitter[:, -1, :] = 0
    jitter[:, :, 0] = jitter[:, :, -1] = 0
    pos = np.stack([xs + jitter[0], ys + jitter[1], z], -1).reshape(-1, 3).astype(np.float32)
    idx = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = i * (ny + 1) + j, (i + 1) * (ny + 1) + j, (i + 1) * (ny + 1) + j + 1, i * (ny + 1) + j + 1
            idx += [a, b, c, a, c, d]
    return pos, np.array(idx, np.uint32)


def _records(pos, idx):
    inst = np.array([[0, 0, 0, 0, idx.size]], np.uint32)
    return O.flatten(pos, idx, inst, np.eye(4, dtype=np.float32).reshape(1, 16))


def test_watertight_shared_edges_and_vertices():
    """Rays aimed exactly at shared edges and vertices of a closed mesh never leak."""
    pos, idx = _grid_mesh(12, 12, 3)
    world = _records(pos, idx)
    rng = np.random.default_rng(11)
    tris = idx.reshape(-1, 3)
    targets = []
    for a, b, c in tris:
        for (p, q) in ((a, b), (b, c), (c, a)):
            s = rng.uniform(0.05, 0.95)
            targets.append((1 - s) * pos[p].astype(np.float64) + s * pos[q])
        targets.append(pos[a].astype(np.float64))
    targets = np.array(targets)
    # keep interior targets (boundary edges have only one triangle)
    inner = (targets[:, 0] > 1e-3) & (targets[:, 0] < 1 - 1e-3) & (targets[:, 1] > 1e-3) & (targets[:, 1] < 1 - 1e-3)
    targets = targets[inner]
    o = targets + np.array([0.013, -0.021, 1.5]) * rng.uniform(0.5, 1.5, (len(targets), 1))
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(targets), 8), np.float32)
    rays[:, :3] = o
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    t, inst, prim, u, v = O.trace_brute(world, rays)
    assert (t >= 0).all(), f"{(t < 0).sum()} rays leaked through shared edges/vertices"


def _room():
    import snkpt
    scene = snkpt.SceneData(1)
    pos, idx, inst, m4, mats = scene.arrays()
    return scene, O.flatten(pos, idx, inst, m4)


def test_flatten_closed_form():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([0, 1, 2], np.uint32)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1], m[2, 2] = 2, 3, 4
    m[3, :3] = (10, 20, 30)  # translation in column-major element 12..14
    inst = np.array([[0, 0, 0, 0, 3]], np.uint32)
    rec = O.flatten(pos, idx, inst, m.reshape(1, 16))
    assert np.array_equal(rec[0, :9], [10, 20, 30, 12, 20, 30, 10, 23, 30])
    assert rec[0, 9:].view(np.uint32).tolist() == [0, 0, 0]


def test_scene_generators_counts_and_determinism():
    import snkpt
    s1 = snkpt.SceneData(1)
    _, idx, inst, _, mats = s1.arrays()
    ntri = int(inst[:, 4].sum()) // 3
    # 5 walls x 2 + 2 boxes x 12 + UV sphere 64x32 with pole fans (3968) + light quad 2
    assert ntri == 10 + 24 + 3968 + 2
    assert s1.s.n_emissive == 1
    s1b = snkpt.SceneData(1)
    assert np.array_equal(s1.arrays()[0], s1b.arrays()[0])
    s2 = snkpt.SceneData(2)
    ntri2 = int(s2.arrays()[2][:, 4].sum()) // 3
    assert ntri2 == 2 + 200 * 5120
    assert s2.s.n_emissive == 16
    s3 = snkpt.SceneData(3)
    inst3 = s3.arrays()[2]
    em = np.ctypeslib.as_array(s3.s.emissive_instances, (s3.s.n_emissive,))
    assert int(inst3[em, 4].sum()) // 3 == 10000 + 2
    assert s3.s.n_emissive <= 17


def test_morton_corners():
    # two triangles whose box centres are the centroid-bounds corners get keys 0 and 0x3FFFFFFF
    pos = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 2, 3], [1, 2, 3], [1, 2, 3]], np.float32)
    idx = np.array([0, 1, 2, 3, 4, 5], np.uint32)
    world = _records(pos, idx)
    b = O.Bvh(world)
    assert b.keys.tolist() == [0, 0x3FFFFFFF]


def _check_bvh_invariants(b: O.Bvh, world: np.ndarray):
    n = b.n
    n_int = n - 1
    assert np.all(np.diff(b.keys.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(b.ids), np.arange(n))
    # radix tree: children partition the parent's range; split at the highest differing bit
    for i in range(n_int):
        f, c = int(b.first[i]), int(b.count[i])
        ranges = []
        for ch in b.child[2 * i: 2 * i + 2]:
            ch = int(ch)
            if ch & 0x80000000:
                ranges.append((ch & 0x7FFFFFFF, 1))
            else:
                ranges.append((int(b.first[ch]), int(b.count[ch])))
        assert ranges[0][0] == f and ranges[0][0] + ranges[0][1] == ranges[1][0]
        assert ranges[1][0] + ranges[1][1] == f + c
        if i >= 1 and i >= n_int:  # pragma: no cover
            pass
    aabb = b.aabb.reshape(-1, 6)

    def box(tag):
        return aabb[n_int + (tag & 0x7FFFFFFF)] if tag & 0x80000000 else aabb[tag]

    for i in range(n_int):
        bb = aabb[i]
        for ch in b.child[2 * i: 2 * i + 2]:
            cb = box(int(ch))
            assert np.all(bb[:3] <= cb[:3]) and np.all(bb[3:] >= cb[3:])
    # BVH8: every triangle exactly once; quantised boxes contain the children
    gids = b.leaf.reshape(-1, 12)[:, 11].view(np.uint32)
    assert np.array_equal(np.sort(gids), np.arange(n))
    nodes = b.nodes.reshape(-1, 20)
    seen_children = set()
    for k, nd in enumerate(nodes):
        p = nd[0:3].view(np.float32).astype(np.float64)
        e = [int((nd[3] >> (8 * a)) & 0xFF) - 127 for a in range(3)]
        imask = int(nd[3] >> 24)
        meta = [(int(nd[6 + (s >> 2)]) >> (8 * (s & 3))) & 0xFF for s in range(8)]
        q = lambda w, s: (int(nd[w + (s >> 2)]) >> (8 * (s & 3))) & 0xFF
        ci = 0
        for s in range(8):
            if meta[s] == 0:
                continue
            lo = [p[a] + q(8 + 2 * a, s) * 2.0 ** e[a] for a in range(3)]
            hi = [p[a] + q(14 + 2 * a, s) * 2.0 ** e[a] for a in range(3)]
            if imask >> s & 1:
                child = int(nd[4]) + ci
                ci += 1
                seen_children.add(child)
                cn = nodes[child]
                # child node box origin lies inside the quantised box
                cp = cn[0:3].view(np.float32)
                assert all(lo[a] <= cp[a] + 1e-12 for a in range(3))
            else:
                cnt = bin(meta[s] >> 5).count("1")
                off = meta[s] & 31
                for tt in range(cnt):
                    r = b.leaf.reshape(-1, 12)[int(nd[5]) + off + tt]
                    v = r[:9].reshape(3, 3).astype(np.float64)
                    assert np.all(v.min(0) >= np.array(lo) - 1e-12) and np.all(v.max(0) <= np.array(hi) + 1e-12)
    assert seen_children == set(range(1, len(nodes)))


def test_build_invariants_room():
    _, world = _room()
    b = O.Bvh(world)
    assert b.n == world.shape[0]
    _check_bvh_invariants(b, world)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9, 33])
def test_build_small(n):
    rng = np.random.default_rng(n)
    pos = rng.uniform(-1, 1, (3 * n, 3)).astype(np.float32)
    world = _records(pos, np.arange(3 * n, dtype=np.uint32))
    b = O.Bvh(world)
    _check_bvh_invariants(b, world) if n >= 2 else None
    rays = np.zeros((512, 8), np.float32)
    org = rng.uniform(-2, 2, (512, 3))
    d = rng.normal(size=(512, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = org, d, 1e30
    got = b.trace(rays)
    want = O.trace_brute(world, rays)
    for a, c in zip(got[:5], want):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_bvh8_traversal_equals_brute_force_room():
    import raysets
    scene, world = _room()
    b = O.Bvh(world)
    lo, hi = world[:, :9].reshape(-1, 3).min(0), world[:, :9].reshape(-1, 3).max(0)
    org, dr = raysets.random_rays(20000, lo, hi, seed=5)
    porg, pdr = raysets.primary_rays(scene.camera(0, 64, 64), 64, 64)
    rays = np.concatenate([np.concatenate([org, dr], 1), np.concatenate([porg, pdr], 1)])
    got = b.trace(rays)
    want = O.trace_brute(world, rays)
    assert (want[0] >= 0).mean() > 0.5
    for a, c in zip(got[:5], want):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    # any-hit agrees on occlusion
    ga = b.trace(rays, any_hit=True)
    wa = O.trace_brute(world, rays, any_hit=True)
    assert np.array_equal(ga[0] >= 0, wa[0] >= 0)


def test_axis_parallel_rays_in_wall_planes_cpu():
    """Rays with zero direction components, many lying exactly in wall planes:
    BVH8 traversal still equals brute force (parallel axes never cull)."""
    _, world = _room()
    b = O.Bvh(world)
    rng = np.random.default_rng(4)
    n = 6000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-1.2, 1.2, (n, 3)) + np.array([0, 1, 0])
    axis = rng.integers(0, 3, n)
    rays[np.arange(n), 4 + axis] = rng.choice([-1.0, 1.0], n)
    rays[: n // 3, 1] = 0.0
    rays[n // 3: 2 * n // 3, 0] = 1.0
    rays[:, 7] = 1e30
    got = b.trace(rays)
    want = O.trace_brute(world, rays)
    for a, c in zip(got[:5], want):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_invalid_rays_miss_cpu():
    """NaN / inf components, zero directions and tmin > tmax are misses."""
    _, world = _room()
    b = O.Bvh(world)
    rays = np.zeros((6, 8), np.float32)
    rays[:, 1] = 1.0
    rays[:, 4] = 1.0
    rays[:, 7] = 1e30
    rays[0, 4] = np.nan
    rays[1, 0] = np.inf
    rays[2, 4:7] = 0.0
    rays[3, 3], rays[3, 7] = 2.0, 1.0
    rays[4, 7] = np.nan
    t, inst, prim, u, v, nv, tv = b.trace(rays)
    assert (t[:5] == -1).all() and (nv[:5] == 0).all()
    assert t[5] > 0  # the control ray hits the right wall
    bt = O.trace_brute(world, rays)[0]
    assert (bt[:5] == -1).all() and bt[5] == t[5]


def test_vertical_ray_from_ground_visits_few_nodes_cpu():
    """Regression: a ray along +y from the ground of the 1.02M-triangle scene
    (two parallel axes) culls normally instead of walking the tree."""
    import snkpt
    scene = snkpt.SceneData(2)
    pos, idx, inst, m4, mats = scene.arrays()
    world = O.flatten(pos, idx, inst, m4)
    b = O.Bvh(world)
    rays = np.array([[-13.745107650756836, 0.010000953450798988, 41.78258514404297, 0.0, 0.0, 1.0, 0.0, 1e30],
                     [5.0, 0.01, 5.0, 0.0, 0.0, 1.0, 0.0, 1e30],
                     [5.0, 30.0, 5.0, 0.0, 0.0, -1.0, 0.0, 1e30],
                     [-100.0, 1.0, 3.0, 0.0, 1.0, 0.0, 0.0, 1e30]], np.float32)
    t, ii, pp, u, v, nv, tv = b.trace(rays)
    assert nv.max() < 200 and tv.max() < 200, (nv, tv)
    w = O.trace_brute(world, rays)
    for a, c in zip((t, ii, pp, u, v), w):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
