"""Synthetic ray sets for measurement and parity (no reference counterpart).

primary    one camera ray per pixel through the pixel centre, as the
           reference unprojects pixels (Util.glsl:1-13: inverse(view) *
           inverse(proj) at depth 1, then normalize(p - cam_pos)).
secondary  one cosine-weighted bounce direction per primary hit, origin
           offset along the geometric normal (facing the incoming ray) by
           1e-4 * scene scale: the incoherent ray set.
shadow     primary hit -> uniformly sampled point on a power-sampled emitter.
Everything is seeded; the same call gives the same rays.
"""
from __future__ import annotations

import numpy as np


def camera_matrix(cam) -> np.ndarray:
    view = np.array(cam.view, np.float64).reshape(4, 4).T  # column-major -> row-major
    proj = np.array(cam.proj, np.float64).reshape(4, 4).T
    return np.linalg.inv(view) @ np.linalg.inv(proj)


def primary_rays(cam, w: int, h: int, tmax: float = 1e30):
    """(N, 4) org_tmin and (N, 4) dir_tmax float32, row-major pixels, row 0 at the top."""
    m = camera_matrix(cam)
    ys, xs = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64), indexing="ij")
    u = (xs + 0.5) / w * 2.0 - 1.0
    v = (ys + 0.5) / h * 2.0 - 1.0
    clip = np.stack([u, v, np.ones_like(u), np.ones_like(u)], -1).reshape(-1, 4)
    wp = clip @ m.T
    p = wp[:, :3] / np.maximum(wp[:, 3:4], 1e-6)
    o = np.array(cam.cam_pos[:3], np.float64)
    d = p - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = d.shape[0]
    org = np.zeros((n, 4), np.float32)
    org[:, :3] = o
    org[:, 3] = 0.0
    dr = np.zeros((n, 4), np.float32)
    dr[:, :3] = d
    dr[:, 3] = tmax
    return org, dr


def random_rays(n: int, lo, hi, seed: int, tmax: float = 1e30):
    """Uniform origins in the box [lo, hi] (slightly enlarged) and uniform directions."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    pad = 0.05 * (hi - lo) + 1e-3
    o = rng.uniform(lo - pad, hi + pad, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org = np.zeros((n, 4), np.float32)
    org[:, :3] = o
    dr = np.zeros((n, 4), np.float32)
    dr[:, :3] = d
    dr[:, 3] = tmax
    return org, dr


def secondary_rays_torch(org, dirs, t, inst, prim, world_tris, inst_tri_offset, scale: float, seed: int):
    """Cosine-weighted bounce rays from the primary hits (torch tensors on one device).

    world_tris: (T, 12) float32 records in global-id order (v0, v1, v2, ids).
    inst_tri_offset: (I,) int64 first global id per instance.
    Returns org_tmin (M, 4), dir_tmax (M, 4) for the M primary rays that hit.
    """
    import torch
    hit = t >= 0
    o = org[hit, :3]
    d = dirs[hit, :3]
    th = t[hit].unsqueeze(1)
    gid = inst_tri_offset[inst[hit].long()] + prim[hit].long()
    tri = world_tris[gid]
    v0, v1, v2 = tri[:, 0:3], tri[:, 3:6], tri[:, 6:9]
    n = torch.linalg.cross(v1 - v0, v2 - v0)
    n = n / n.norm(dim=1, keepdim=True).clamp_min(1e-30)
    n = torch.where(((n * d).sum(1, keepdim=True) > 0), -n, n)  # face the incoming ray
    p = o + d * th + n * (1e-4 * scale)
    g = torch.Generator(device=org.device)
    g.manual_seed(seed)
    r1 = torch.rand(p.shape[0], device=org.device, generator=g)
    r2 = torch.rand(p.shape[0], device=org.device, generator=g)
    # cosine-weighted hemisphere around n (Malley's method)
    rr = r1.sqrt()
    phi = 2.0 * np.pi * r2
    lx, ly, lz = rr * torch.cos(phi), rr * torch.sin(phi), (1.0 - r1).clamp_min(0).sqrt()
    a = torch.where(n[:, 0:1].abs() > 0.9, torch.tensor([0.0, 1.0, 0.0], device=org.device),
                    torch.tensor([1.0, 0.0, 0.0], device=org.device))
    tng = torch.linalg.cross(a, n)
    tng = tng / tng.norm(dim=1, keepdim=True)
    btg = torch.linalg.cross(n, tng)
    nd = tng * lx[:, None] + btg * ly[:, None] + n * lz[:, None]
    nd = nd / nd.norm(dim=1, keepdim=True)
    ok = torch.isfinite(p).all(1) & torch.isfinite(nd).all(1)
    p, nd = p[ok], nd[ok]
    m = p.shape[0]
    out_o = torch.zeros(m, 4, device=org.device)
    out_o[:, :3] = p
    out_d = torch.zeros(m, 4, device=org.device)
    out_d[:, :3] = nd
    out_d[:, 3] = 1e30
    return out_o.contiguous(), out_d.contiguous()


def shadow_rays_torch(org, dirs, t, emit_tris, emit_pdf, seed: int, scale: float, normals=None):
    """Shadow rays from primary hits to uniformly sampled points on power-sampled emitters.

    emit_tris: (E, 9) world-space emitter vertices; emit_pdf: (E,) selection probabilities.
    """
    import torch
    hit = t >= 0
    o = org[hit, :3]
    d = dirs[hit, :3]
    p = o + d * t[hit].unsqueeze(1) - d * (1e-4 * scale)
    g = torch.Generator(device=org.device)
    g.manual_seed(seed)
    e = torch.multinomial(emit_pdf, p.shape[0], replacement=True, generator=g)
    r1 = torch.rand(p.shape[0], device=org.device, generator=g)
    r2 = torch.rand(p.shape[0], device=org.device, generator=g)
    su = r1.sqrt()
    b0, b1, b2 = 1 - su, su * (1 - r2), su * r2
    tri = emit_tris[e]
    q = tri[:, 0:3] * b0[:, None] + tri[:, 3:6] * b1[:, None] + tri[:, 6:9] * b2[:, None]
    dv = q - p
    dist = dv.norm(dim=1)
    m = p.shape[0]
    out_o = torch.zeros(m, 4, device=org.device)
    out_o[:, :3] = p
    out_o[:, 3] = 1e-4 * scale
    out_d = torch.zeros(m, 4, device=org.device)
    out_d[:, :3] = dv / dist[:, None]
    out_d[:, 3] = dist * (1 - 1e-4)
    return out_o.contiguous(), out_d.contiguous()
