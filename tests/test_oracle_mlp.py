"""CPU tests of the radiance-cache oracle: finite-difference gradients, Adam's
closed first step, hash-grid geometry and index closed forms."""
import numpy as np
import pytest

from oracle import binding as O


def _rand_params(seed, grid_scale=0.5):
    rng = np.random.default_rng(seed)
    W = np.zeros(O.MLP_PARAMS, np.float32)
    for l in range(4):
        lim = np.sqrt(6 / 128)
        W[l * 4096:(l + 1) * 4096] = rng.uniform(-lim, lim, 4096)
    W[4 * 4096: 4 * 4096 + 3 * 64] = rng.uniform(-np.sqrt(6 / 67), np.sqrt(6 / 67), 3 * 64)
    grid = (rng.uniform(-1, 1, O.GRID_PARAMS) * grid_scale).astype(np.float32)
    return W, grid


def _samples(n, seed, lo=(-1, 0, -1), hi=(1, 2, 1)):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    s[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    s[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    s[:, 9] = rng.uniform(0.05, 1, n)
    s[:, 10:13] = rng.uniform(0.1, 0.9, (n, 3))
    s[:, 13:16] = rng.uniform(0, 2, (n, 3))
    return s


def test_grid_geometry_and_hash_closed_forms():
    g = O.CacheGeom([-1, 0, -1], [1, 2, 1])
    assert g.res[0] == 16 and g.res[15] == 2048
    assert np.all(np.diff(g.res.astype(np.int64)) > 0)
    # dense levels are exactly those whose (N+1)^3 vertices fit in 2^19 entries
    assert np.array_equal(g.dense, ((g.res.astype(np.int64) + 1) ** 3 <= (1 << 19)).astype(np.uint32))
    assert g.lo_inv[3] == np.float32(0.5)
    # Mueller et al. 2022 primes (1, 2654435761, 805459861), xor, modulo 2^19
    assert O.grid_index(1, 0, 0, 2048, 0) == 1
    assert O.grid_index(0, 1, 0, 2048, 0) == 2654435761 % (1 << 19)
    assert O.grid_index(0, 0, 1, 2048, 0) == 805459861 % (1 << 19)
    assert O.grid_index(3, 5, 7, 2048, 0) == (3 ^ (5 * 2654435761 % 2**32) ^ (7 * 805459861 % 2**32)) % (1 << 19)
    # dense: x + y (N+1) + z (N+1)^2, a bijection onto [0, (N+1)^3)
    assert O.grid_index(2, 3, 4, 16, 1) == 2 + 3 * 17 + 4 * 289
    idx = {O.grid_index(x, y, z, 16, 1) for x in range(17) for y in range(17) for z in range(17)}
    assert len(idx) == 17 ** 3 and max(idx) == 17 ** 3 - 1


def test_encode_trilinear_reproduces_affine_grid():
    """A grid whose level-0 features are an affine function of the vertex
    coordinates is reproduced exactly by trilinear interpolation."""
    g = O.CacheGeom([0, 0, 0], [1, 1, 1])
    grid = np.zeros(O.GRID_PARAMS, np.float32)
    r = int(g.res[0])
    for z in range(r + 1):
        for y in range(r + 1):
            for x in range(r + 1):
                i = O.grid_index(x, y, z, r, 1)
                grid[2 * i] = (x + 2 * y + 3 * z) / r
                grid[2 * i + 1] = 1.0
    s = _samples(50, 1, lo=(0, 0, 0), hi=(1, 1, 1))
    enc = O.cache_encode(g, grid, s)
    want = s[:, 0] + 2 * s[:, 1] + 3 * s[:, 2]
    assert np.allclose(enc[:, 0], want, atol=1e-5)
    assert np.allclose(enc[:, 1], 1.0, atol=1e-12)
    assert np.allclose(enc[:, 32:35], 0.5 * (s[:, 3:6] + 1), atol=1e-6)


def test_gradients_match_finite_differences():
    g = O.CacheGeom([-1, 0, -1], [1, 2, 1])
    W, grid = _rand_params(3)
    s = _samples(6, 4)
    loss, dW, dg, den = O.cache_grads(g, grid, W, s, want_den=True)
    assert loss > 0
    rng = np.random.default_rng(5)
    h = 1e-3

    def loss_at(Wx, gx):  # denominators frozen: the loss's stop-gradient
        return O.cache_grads(g, gx, Wx, s, want_grid=False, den=den)[0]

    nz = np.nonzero(dW)[0]
    for p in rng.choice(nz, 24, replace=False):
        Wp, Wm = W.copy(), W.copy()
        Wp[p] += h
        Wm[p] -= h
        fd = (loss_at(Wp, grid) - loss_at(Wm, grid)) / (float(Wp[p]) - float(Wm[p]))
        assert abs(fd - dW[p]) <= 2e-3 * abs(dW[p]) + 1e-6, (p, fd, dW[p])
    gz = np.nonzero(dg)[0]
    assert gz.size > 0
    hg = 1e-4  # grid features feed 8 ReLU layers' worth of kinks: a smaller step
    for p in rng.choice(gz, 24, replace=False):
        gp, gm = grid.copy(), grid.copy()
        gp[p] += hg
        gm[p] -= hg
        fd = (loss_at(W, gp) - loss_at(W, gm)) / (float(gp[p]) - float(gm[p]))
        assert abs(fd - dg[p]) <= 2e-3 * abs(dg[p]) + 1e-6, (p, fd, dg[p])


def test_adam_closed_first_step():
    rng = np.random.default_rng(2)
    p0 = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100)
    lr = 1e-2
    p1, m1, v1 = O.adam(p0, np.zeros(100), np.zeros(100), g, 1, lr)
    # t = 1 with zero moments: the bias-corrected step is lr * g / (|g| + eps / sqrt(1 - b2))
    want = p0 - lr * g / (np.abs(g) + 1e-15 / np.sqrt(0.01))
    assert np.allclose(p1, want, rtol=1e-6, atol=1e-7)
    assert np.allclose(m1, 0.1 * g, rtol=1e-6) and np.allclose(v1, 0.01 * g * g, rtol=1e-6)
