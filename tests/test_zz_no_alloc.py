"""snkpt_cache_infer_noalloc (static device workspace, no allocation) against
snkpt_cache_infer, through the C ABI: bit-identical outputs on two calls in a
row, the second smaller than the first, and a clear error above its limit."""
import numpy as np
import pytest

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
    return s


def test_noalloc_queries_equal_allocating_queries(gpu):
    import torch

    import snkpt
    ctx = snkpt.Context(0)
    ctx.cache_reset(3, LO, HI)
    rng = np.random.default_rng(3)
    grid = (rng.uniform(-1, 1, snkpt.CACHE_GRID_PARAMS) * 0.3).astype(np.float32)
    ctx.cache_set_params(1, grid)
    for n, seed in ((65536, 1), (1237, 2)):
        print(f"noalloc query n={n}", flush=True)
        s = torch.from_numpy(_samples(n, seed)).cuda()
        want = ctx.cache_infer(s)
        got = torch.full((n, 4), float("nan"), device="cuda")
        ctx.cache_infer_noalloc(s, got)
        torch.cuda.synchronize()
        w, g = want[:, :3].cpu().numpy(), got[:, :3].cpu().numpy()
        assert np.isfinite(g).all()
        assert np.array_equal(w.view(np.uint32), g.view(np.uint32)), n
    s = torch.zeros(65537, 16, device="cuda")
    out = torch.zeros(65537, 4, device="cuda")
    with pytest.raises(snkpt.SnkptError) as e:
        ctx.cache_infer_noalloc(s, out)
    assert e.value.code == snkpt.ERR_CAPACITY and "65,536" in str(e.value)
    ctx.cache_infer_noalloc(s[:0], out[:0])  # n = 0 is a no-op
