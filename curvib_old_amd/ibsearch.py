"""Immersed-boundary node list of the analytic sphere of a ChannelCase: the
records cv_set_ib_nodes takes (IBMInfo, variables.h:36-72).

The reference writes them with its immersed-body search over a triangulated
surface (ibm_search_advanced, then InterceptionPoint, ibm.c:1952-2420).
Here the body is the exact sphere that cases.nvert_field classifies:
  - foot point pmin: the sphere point nearest the IB node's cell centre, so
    d_s = |centre - c| - R and the normal is (centre - pmin) / d_s;
  - interception: the ray from the cell centre along the normal, tried
    against the reference's six 3x3 patches of neighbouring cell centres in
    its order (i-1, i+1, j-1, j+1, k-1, k+1) and, in each, its eight
    triangles (ICP, ibm.c:1861-1950).  The first triangle the ray meets at
    d > 0 gives the three cells, their barycentric weights cr1..cr3 and
    d_i, the distance to the interception point.
The surface is at rest: ua = 0, dun = 0.  nf is the sphere normal at pmin.
"""
from __future__ import annotations

import numpy as np

from . import cases

# ICP's triangles over the 3x3 patch points (ibm.c:1877-1884)
_TRIANGLES = ((0, 1, 4), (1, 2, 4), (2, 4, 5), (4, 5, 8), (4, 7, 8), (4, 6, 7), (3, 4, 6), (3, 4, 0))


def _patch(axis, side):
    """(di, dj, dk) of the nine patch points m = 0..8 (InterceptionPoint's
    ip/jp/kp tables, ibm.c:1974-2359)."""
    m = np.arange(9)
    a, b = m % 3 - 1, m // 3 - 1
    s = np.full(9, side)
    if axis == 0:
        return np.stack([s, a, b], axis=1)
    if axis == 1:
        return np.stack([a, s, b], axis=1)
    return np.stack([a, b, s], axis=1)


def sphere_ib_nodes(case: cases.ChannelCase, xyz: np.ndarray, nvert: np.ndarray):
    """(idx, val): idx (n, 12) int32 = i j k, i1 j1 k1, i2 j2 k2, i3 j3 k3 and
    val (n, 12) float64 = cr1 cr2 cr3, d_s, d_i, ua[3], dun, nf[3], one row
    per IB node (nvert 1.4) in the reference's k, j, i loop order."""
    if case.sphere is None:
        return np.zeros((0, 12), np.int32), np.zeros((0, 12))
    mx, my, mz = case.storage
    cent = np.zeros((mz, my, mx, 3))
    cent[1:-1, 1:-1, 1:-1] = cases.cell_centres(case, xyz)
    ibm = np.zeros_like(nvert, dtype=bool)
    ibm[1:-1, 1:-1, 1:-1] = np.abs(nvert[1:-1, 1:-1, 1:-1] - 1.4) < 1e-12
    kk, jj, ii = np.nonzero(ibm)
    if len(kk) and (ii.min() < 2 or jj.min() < 2 or kk.min() < 2 or ii.max() > mx - 3 or jj.max() > my - 3 or
                    kk.max() > mz - 3):
        raise ValueError("IB nodes must keep a cell of fluid between them and the domain boundary")
    cx, cy, cz, d = case.sphere
    c0, R = np.array([cx, cy, cz]), 0.5 * d
    p = cent[kk, jj, ii]
    v = p - c0
    r = np.linalg.norm(v, axis=1)
    nrm = v / r[:, None]
    pmin = c0 + R * nrm
    d_s = r - R
    n = len(kk)
    idx = np.zeros((n, 12), np.int32)
    val = np.zeros((n, 12))
    idx[:, 0], idx[:, 1], idx[:, 2] = ii, jj, kk
    val[:, 3] = d_s
    val[:, 9:12] = nrm
    todo = np.ones(n, dtype=bool)
    for axis, side in ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)):
        off = _patch(axis, side)
        for tri in _TRIANGLES:
            if not todo.any():
                break
            q = np.nonzero(todo)[0]
            pts = []
            for m in tri:
                pts.append(np.stack([ii[q] + off[m, 0], jj[q] + off[m, 1], kk[q] + off[m, 2]], axis=1))
            p1, p2, p3 = (cent[t[:, 2], t[:, 1], t[:, 0]] for t in pts)
            e1, e2 = p2 - p1, p3 - p1
            nt = np.cross(e1, e2)
            den = np.einsum("ij,ij->i", nrm[q], nt)
            ok = np.abs(den) > 1e-10
            t = np.where(ok, np.einsum("ij,ij->i", p1 - p[q], nt) / np.where(ok, den, 1.0), -1.0)
            ok &= t > 0
            x = p[q] + t[:, None] * nrm[q]
            # barycentric weights of x in (p1, p2, p3)
            w = x - p1
            d00, d01, d11 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
            d20, d21 = (w * e1).sum(1), (w * e2).sum(1)
            den2 = d00 * d11 - d01 * d01
            b2 = (d11 * d20 - d01 * d21) / den2
            b3 = (d00 * d21 - d01 * d20) / den2
            b1 = 1.0 - b2 - b3
            ok &= (b1 >= -1e-12) & (b2 >= -1e-12) & (b3 >= -1e-12)
            hit = q[ok]
            for s, pt in enumerate(pts):
                idx[hit, 3 + 3 * s:6 + 3 * s] = pt[ok]
            val[hit, 0], val[hit, 1], val[hit, 2] = b1[ok], b2[ok], b3[ok]
            val[hit, 4] = t[ok]
            todo[hit] = False
    if todo.any():
        raise ValueError(f"{int(todo.sum())} IB nodes found no interception triangle")
    return idx, val
