"""CPU tests pinning the numpy restatement of the TAA resolve
(TAA_Resolve.comp:37-86) with closed forms."""
import numpy as np

from oracle import taa

F = np.float32


def _vel(h, w, v=(0.0, 0.0)):
    out = np.zeros((h, w, 2), F)
    out[...] = v
    return out


def test_constant_image_is_a_fixed_point():
    c = np.full((6, 7, 4), 0.37, F)
    out = taa.resolve(c, _vel(6, 7), c, _vel(6, 7))
    assert np.allclose(out[..., :3], 0.37, rtol=1e-6) and np.all(out[..., 3] == 1.0)


def test_clip_aabb_closed_forms():
    lo, hi = np.full((1, 3), -1.0, F), np.full((1, 3), 1.0, F)
    p = np.zeros((1, 3), F)
    for q, want in (([4, 0, 0], [1, 0, 0]), ([2, 2, 0], [1, 1, 0]), ([0.5, -0.25, 0.1], [0.5, -0.25, 0.1]),
                    ([-3, 1.5, 0], [-1, 0.5, 0])):
        got = taa._clip_aabb(lo, hi, p, np.array([q], F))
        assert np.allclose(got, [want], atol=1e-7), (q, got)


def test_bilinear_repeat_sampler():
    img = np.arange(12, dtype=F).reshape(3, 4, 1)
    centre = np.array([[(1 + 0.5) / 4, (2 + 0.5) / 3]], F)
    assert np.allclose(taa._bilinear_repeat(img, centre), img[2, 1])
    half = np.array([[2.0 / 4, 1.5 / 3]], F)  # halfway between columns 1 and 2 of row 1
    assert np.allclose(taa._bilinear_repeat(img, half), 0.5 * (img[1, 1] + img[1, 2]))
    edge = np.array([[0.0, 1.5 / 3]], F)  # u = 0: half column 3 (wrapped), half column 0
    assert np.allclose(taa._bilinear_repeat(img, edge), 0.5 * (img[1, 3] + img[1, 0]))


def test_static_blend_inside_the_box_is_point_nine_history():
    rng = np.random.default_rng(3)
    cur = (0.5 + 0.1 * rng.random((5, 5, 4))).astype(F)
    mean = cur[1:4, 1:4, :3].mean((0, 1))
    hist = np.zeros((5, 5, 4), F)
    hist[..., :3] = mean  # the centre's neighbourhood mean: inside its box
    out = taa.resolve(cur, _vel(5, 5), hist, _vel(5, 5))
    want = cur[2, 2, :3] * F(1.0 - 0.9) + mean * F(0.9)
    assert np.allclose(out[2, 2, :3], want, rtol=1e-5)


def test_history_outside_the_box_is_clipped():
    cur = np.full((5, 5, 4), 0.5, F)
    cur[2, 2, :3] = 0.6  # some spread in the centre neighbourhood
    hist = np.full((5, 5, 4), 5.0, F)
    out = taa.resolve(cur, _vel(5, 5), hist, _vel(5, 5))
    col = cur[1:4, 1:4, 0]
    mu, sigma = col.mean(), np.sqrt(abs((col * col).mean() - col.mean() ** 2))
    clipped = mu + 0.55 * sigma
    assert np.allclose(out[2, 2, 0], 0.1 * 0.6 + 0.9 * clipped, rtol=1e-4)


def test_fast_motion_gives_the_neighbourhood_mean():
    rng = np.random.default_rng(5)
    cur = rng.random((6, 6, 4)).astype(F)
    out = taa.resolve(cur, _vel(6, 6, (0.2, 0.0)), rng.random((6, 6, 4)).astype(F), _vel(6, 6))
    mu = cur[2:5, 2:5, :3].mean((0, 1))  # pixel (3, 3): |v| = 0.2 -> disocclusion weight 1
    assert np.allclose(out[3, 3, :3], mu, rtol=1e-5)
