"""Contact record of a terminate-on-first-hit collision ray, restated in numpy
float32 on top of the oracle's any-hit BVH8 traversal.  TEST INFRASTRUCTURE
ONLY (the checker of snkpt_trace_contacts, never the product).

Follows ParticleUpdate.rchit:38-62: barycentrics (1-u-v, u, v); object-space
normal and position interpolated n0 b0 + n1 b1 + n2 b2; world normal
normalize(transpose(inverse(mat3(M))) n); contact position origin + dir * t;
contact velocity (position - (M_prev * (p, 1)).xyz) / timestep.  Pinned by
tests/test_oracle_contacts.py (closed forms on a transformed quad).
"""
from __future__ import annotations

import numpy as np

F = np.float32


def normal_matrix(m4_colmajor):
    """transpose(inverse(mat3(M))) of a column-major mat4, formed in float64,
    rounded to float32 once (3 x 3, rows)."""
    m = np.asarray(m4_colmajor, np.float64).reshape(4, 4).T  # row-major
    return np.linalg.inv(m[:3, :3]).T.astype(F)


def contacts(arrs, prev_m4, rays, t, inst, prim, u, v, timestep):
    """arrs: (pos, idx, inst5, mat4s, mats, nrm, em); prev_m4: previous-frame
    mat4s (column-major, same count); rays (n, 8); any-hit results.  Returns
    (normal, position, velocity), each (n, 3) float32, zero on misses."""
    pos, idx, inst5, m4, _mats, nrm, _em = arrs
    pos = np.asarray(pos, F).reshape(-1, 3)
    nrm = np.asarray(nrm, F).reshape(-1, 3)
    m4 = np.asarray(m4, F).reshape(-1, 16)
    prev = np.asarray(prev_m4, F).reshape(-1, 16)
    rays = np.asarray(rays, F).reshape(-1, 8)
    n = rays.shape[0]
    out_n = np.zeros((n, 3), F)
    out_p = np.zeros((n, 3), F)
    out_v = np.zeros((n, 3), F)
    ts = F(timestep)
    for k in np.nonzero(np.asarray(t) >= 0)[0]:
        rec = inst5[inst[k]]
        ti, io, vo = int(rec[0]), int(rec[1]), int(rec[2])
        ix = idx[io + 3 * int(prim[k]): io + 3 * int(prim[k]) + 3].astype(np.int64) + vo
        uu, vv = F(u[k]), F(v[k])
        b0 = F(1.0) - uu - vv
        nl = (nrm[ix[0]] * b0 + nrm[ix[1]] * uu) + nrm[ix[2]] * vv
        pl = (pos[ix[0]] * b0 + pos[ix[1]] * uu) + pos[ix[2]] * vv
        nm = normal_matrix(m4[ti])
        nw = np.array([(nm[r, 0] * nl[0] + nm[r, 1] * nl[1]) + nm[r, 2] * nl[2] for r in range(3)], F)
        l2 = (nw[0] * nw[0] + nw[1] * nw[1]) + nw[2] * nw[2]
        if l2 > 0:
            nw = nw * (F(1.0) / np.sqrt(l2))
        o, d, tk = rays[k, :3], rays[k, 4:7], F(t[k])
        cp = (o + d * tk).astype(F)
        q = prev[ti].reshape(4, 4).T  # row-major
        pp = np.array([((q[r, 0] * pl[0] + q[r, 1] * pl[1]) + q[r, 2] * pl[2]) + q[r, 3] for r in range(3)], F)
        out_n[k], out_p[k], out_v[k] = nw, cp, (cp - pp) / ts
    return out_n, out_p, out_v
