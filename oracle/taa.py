"""Temporal anti-aliasing resolve restated in numpy (float32).  TEST
INFRASTRUCTURE ONLY (the checker of the HIP kernel, never the product).

Follows TAA_Resolve.comp:37-86 with the pass around it
(TAA_ResolvePass.cpp:39-56):
  - history texcoord = (pixel + 0.5) / size + velocity, clamped to [0, 1];
  - 3x3 neighbourhood of the current frame (edge-clamped texel coordinates),
    each colour max(c, 0): min, max, mean mu and sigma = sqrt(|E[c^2] - mu^2|);
    box mu -/+ 0.55 sigma (TAA_Resolve.comp:55-74);
  - history colour max(h, 0) from a bilinear, repeat-addressed sampler
    (Images.cpp:14-33), clipped toward clamp(mu, min, max) against the box
    (clip_aabb, TAA_Resolve.comp:11-35);
  - previous velocity sampled the same way; disocclusion weight
    clamp((|v_prev + v| - 0.001) * 10, 0, 1);
  - out = mix(mix(current, history, 0.9), mu, disocclusion), alpha 1;
  - the output becomes the next history and the velocity the next previous
    velocity (the pass's two blits).
Every pixel reads the frame as it was before the resolve (the shader reads its
neighbours from the image it writes).  Pinned by tests/test_oracle_taa.py.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _bilinear_repeat(img, uv):
    """Bilinear sample of img (h, w, c) at texcoords uv (..., 2), repeat wrap."""
    h, w = img.shape[:2]
    x = uv[..., 0] * F(w) - F(0.5)
    y = uv[..., 1] * F(h) - F(0.5)
    fx, fy = np.floor(x), np.floor(y)
    ax, ay = (x - fx)[..., None], (y - fy)[..., None]
    x0 = np.mod(fx.astype(np.int64), w)
    y0 = np.mod(fy.astype(np.int64), h)
    x1, y1 = np.mod(x0 + 1, w), np.mod(y0 + 1, h)
    c00, c10, c01, c11 = img[y0, x0], img[y0, x1], img[y1, x0], img[y1, x1]
    one = F(1.0)
    top = (one - ax) * c00 + ax * c10
    bot = (one - ax) * c01 + ax * c11
    return ((one - ay) * top + ay * bot).astype(F)


def _clip_aabb(lo, hi, p, q):
    """clip_aabb (TAA_Resolve.comp:11-35) on (..., 3) arrays; p, q are the
    xyz of (p, 1) and (q, 1), so the w component of r stays 0."""
    r = (q - p).astype(F)
    rmax = (hi - p).astype(F)
    rmin = (lo - p).astype(F)
    eps = F(0.00000001)
    for a in range(3):
        m = r[..., a] > rmax[..., a] + eps
        s = np.where(m, rmax[..., a] / np.where(m, r[..., a], F(1.0)), F(1.0)).astype(F)
        r = (r * s[..., None]).astype(F)
    for a in range(3):
        m = r[..., a] < rmin[..., a] - eps
        s = np.where(m, rmin[..., a] / np.where(m, r[..., a], F(1.0)), F(1.0)).astype(F)
        r = (r * s[..., None]).astype(F)
    return (p + r).astype(F)


def resolve(current, velocity, history, prev_velocity):
    """One resolve.  current / history (h, w, 4) float32, velocity /
    prev_velocity (h, w, 2) float32.  Returns the output (h, w, 4)."""
    current = np.asarray(current, F)
    h, w = current.shape[:2]
    ys, xs = np.mgrid[0:h, 0:w]
    uv = np.stack([(xs.astype(F) + F(0.5)) / F(w), (ys.astype(F) + F(0.5)) / F(h)], -1).astype(F)
    huv = np.clip(uv + velocity.astype(F), F(0.0), F(1.0)).astype(F)
    cur3 = current[..., :3]
    mn = np.full((h, w, 3), F(9999.0), F)
    mx = np.full((h, w, 3), F(-9999.0), F)
    m = np.zeros((h, w, 3), F)
    m2 = np.zeros((h, w, 3), F)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            cx = np.clip(xs + dx, 0, w - 1)
            cy = np.clip(ys + dy, 0, h - 1)
            col = np.maximum(cur3[cy, cx], F(0.0))
            mn = np.minimum(mn, col)
            mx = np.maximum(mx, col)
            m = (m + col).astype(F)
            m2 = (m2 + col * col).astype(F)
    inv9 = F(1.0 / 9.0)
    gamma = F(0.55)
    mu = (m * inv9).astype(F)
    sigma = np.sqrt(np.abs(m2 * inv9 - mu * mu)).astype(F)
    minc = (mu - gamma * sigma).astype(F)
    maxc = (mu + gamma * sigma).astype(F)
    hist = np.maximum(_bilinear_repeat(np.asarray(history, F)[..., :3], huv), F(0.0))
    hist = _clip_aabb(minc, maxc, np.clip(mu, mn, mx), hist)
    pv = _bilinear_repeat(np.asarray(prev_velocity, F), huv)
    v = (pv + velocity.astype(F)).astype(F)
    vlen = np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]).astype(F)
    dis = np.clip((vlen - F(0.001)) * F(10.0), F(0.0), F(1.0))[..., None]
    a = F(0.9)
    fin = (cur3 * (F(1.0) - a) + hist * a).astype(F)
    fin = (fin * (F(1.0) - dis) + mu * dis).astype(F)
    out = np.ones((h, w, 4), F)
    out[..., :3] = fin
    return out
