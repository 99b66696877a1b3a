"""GPU parity of the radiance-cache kernels (hash-grid encoding, MFMA MLP
forward/backward, Adam) against the FP64 CPU oracle, plus learning and
determinism checks."""
import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu

LO, HI = [-1.0, 0.0, -1.0], [1.0, 2.0, 1.0]


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(LO, HI, (n, 3))
    d = rng.normal(size=(n, 3))
    s[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    s[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    s[:, 9] = rng.uniform(0.05, 1, n)
    s[:, 10:13] = rng.uniform(0.1, 0.9, (n, 3))
    x = s[:, 0:3]
    base = 0.5 + 0.4 * np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 2]) + 0.2 * x[:, 1]
    s[:, 13] = base
    s[:, 14] = base * 0.8 + 0.1 * s[:, 3]
    s[:, 15] = base * 0.6 + 0.1 * s[:, 11]
    return s


def _ctx_with_params(seed=1, grid_scale=0.3):
    import snkpt
    ctx = snkpt.Context(0)
    ctx.cache_reset(seed, LO, HI)
    rng = np.random.default_rng(seed)
    W = ctx.cache_params(0).copy()
    grid = (rng.uniform(-1, 1, O.GRID_PARAMS) * grid_scale).astype(np.float32)
    ctx.cache_set_params(1, grid)
    return ctx, W, grid


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def test_encode_and_forward_match_oracle(gpu):
    ctx, W, grid = _ctx_with_params(1)
    g = O.CacheGeom(LO, HI)
    s = _samples(1000, 2)
    y = ctx.cache_infer(_dev(s)).cpu().numpy()[:, :3]
    enc = ctx.cache_params(4).reshape(-1, 64).astype(np.float64)
    want_enc = O.cache_encode(g, grid, s)
    assert np.allclose(enc, want_enc, rtol=4e-3, atol=2e-3)
    # the oracle runs on the GPU's fp16-rounded weights and grid, in FP64
    W16 = W.astype(np.float16).astype(np.float32)
    grid16 = grid.astype(np.float16).astype(np.float32)
    want = O.cache_infer(g, grid16, W16, s)
    err = np.abs(y - want)
    assert np.all(err <= 3e-2 * np.abs(want) + 1e-2), err.max()
    assert np.sqrt((err ** 2).mean()) < 5e-3 * np.sqrt((want ** 2).mean()) + 1e-3


def test_forward_closed_form_on_routing_weights(gpu):
    """The MFMA forward on hand-set weights that route inputs through the
    ReLU layers gives the closed-form outputs (FP16 inputs and activations)."""
    from test_oracle_pins import routing_net
    import snkpt
    g, W, grid, s, want = routing_net(1000)
    ctx = snkpt.Context(0)
    ctx.cache_reset(1, [0, 0, 0], [1, 1, 1])
    ctx.cache_set_params(0, W)
    ctx.cache_set_params(1, grid)
    y = ctx.cache_infer(_dev(s)).cpu().numpy()[:, :3].astype(np.float64)
    err = np.abs(y - want)
    assert np.all(err <= 4e-3 * np.abs(want) + 4e-3), err.max()


def test_gradients_match_oracle(gpu):
    ctx, W, grid = _ctx_with_params(3)
    g = O.CacheGeom(LO, HI)
    s = _samples(4096, 4)
    loss_gpu = ctx.cache_train(_dev(s), lr=1e-2, apply=False)
    gW = ctx.cache_params(2).astype(np.float64)
    gG = ctx.cache_params(3).astype(np.float64)
    W16 = W.astype(np.float16).astype(np.float32)
    grid16 = grid.astype(np.float16).astype(np.float32)
    loss, dW, dG = O.cache_grads(g, grid16, W16, s)
    assert abs(loss_gpu - loss) <= 2e-2 * loss
    cos = gW @ dW / (np.linalg.norm(gW) * np.linalg.norm(dW))
    rel = np.linalg.norm(gW - dW) / np.linalg.norm(dW)
    assert cos > 0.999 and rel < 3e-2, (cos, rel)
    tz = np.nonzero(dG)[0]
    relg = np.linalg.norm(gG[tz] - dG[tz]) / np.linalg.norm(dG[tz])
    assert relg < 5e-2, relg
    assert np.count_nonzero(gG) <= tz.size  # gradients only where the oracle touches the grid


def test_adam_step_matches_oracle(gpu):
    ctx, W, grid = _ctx_with_params(5)
    s = _dev(_samples(2048, 6))
    ctx.cache_train(s, apply=False)
    gW = ctx.cache_params(2).astype(np.float64)
    gG = ctx.cache_params(3).astype(np.float64)
    ctx.cache_train(s, lr=1e-2, apply=True)
    W1 = ctx.cache_params(0)
    G1 = ctx.cache_params(1)
    wW, _, _ = O.adam(W, np.zeros_like(W), np.zeros_like(W), gW, 1, 1e-2)
    wG, _, _ = O.adam(grid, np.zeros_like(grid), np.zeros_like(grid), gG, 1, 1e-2)
    assert np.allclose(W1, wW, rtol=1e-5, atol=1e-7)
    assert np.allclose(G1, wG, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss_and_is_deterministic(gpu):
    import snkpt
    losses = []
    finals = []
    for rep in range(2):
        ctx = snkpt.Context(0)
        ctx.cache_reset(7, LO, HI)
        ls = []
        for it in range(300):
            ls.append(ctx.cache_train(_dev(_samples(8192, 1000 + it)), lr=1e-2))
        losses.append(ls)
        finals.append((ctx.cache_params(0), ctx.cache_params(1)))
    l0, l1 = np.mean(losses[0][:5]), np.mean(losses[0][-10:])
    assert l1 < 0.1 * l0, (l0, l1)
    assert np.array_equal(finals[0][0].view(np.uint32), finals[1][0].view(np.uint32))
    assert np.array_equal(finals[0][1].view(np.uint32), finals[1][1].view(np.uint32))
    # held-out error of the trained cache
    test = _samples(4096, 99)
    y = ctx.cache_infer(_dev(test)).cpu().numpy()[:, :3]
    rel = np.abs(y - test[:, 13:16]).mean() / np.abs(test[:, 13:16]).mean()
    assert rel < 0.1, rel
