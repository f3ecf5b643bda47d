"""Immersed-boundary node lists: the records cv_set_ib_nodes takes (IBMInfo,
variables.h:36-72), for the analytic sphere of a ChannelCase or for a
triangulated body in the reference's Icem format (formats.read_icem).

The reference writes them with its immersed-body search over a triangulated
surface (ibm_search_advanced, then InterceptionPoint, ibm.c:1952-2420).  The
IB nodes are the cells nvert marks 1.4; for each:
  - foot point pmin: the body point nearest the node's cell centre (the
    exact sphere, or the nearest point over the surface's triangles with its
    element and barycentric weights cs1..cs3), d_s = |centre - pmin| and the
    normal (centre - pmin) / d_s;
  - interception: the ray from the cell centre along the normal, tried
    against the reference's six 3x3 patches of neighbouring cell centres in
    its order (i-1, i+1, j-1, j+1, k-1, k+1) and, in each, its eight
    triangles (ICP, ibm.c:1861-1950).  The first triangle the ray meets at
    d > 0 gives the three cells, their barycentric weights cr1..cr3 and
    d_i, the distance to the interception point;
  - the surface velocity at pmin, ua = sum of the element's vertex
    velocities times cs, and dun = (ua - uaold) . normal
    (ibm_interpolation_advanced, ibm.c:3082-3105); nf the element normal.
The analytic sphere is at rest (ua = 0, dun = 0; nf the sphere normal).
This is a restatement for the synthetic bodies, not of the reference's
search code: its own parity is not pinned against the reference (which
needs PETSc to run), only against the geometry (see tests/).
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


def _cell_centres(case, xyz):
    mx, my, mz = case.storage
    cent = np.zeros((mz, my, mx, 3))
    cent[1:-1, 1:-1, 1:-1] = cases.cell_centres(case, xyz)
    return cent


def _ib_cells(case, nvert):
    """(k, j, i) of the IB nodes (nvert 1.4) in the reference's loop order."""
    mx, my, mz = case.storage
    ibm = np.zeros_like(nvert, dtype=bool)
    ibm[1:-1, 1:-1, 1:-1] = np.abs(nvert[1:-1, 1:-1, 1:-1] - 1.4) < 1e-12
    kk, jj, ii = np.nonzero(ibm)
    if len(kk) and (ii.min() < 2 or jj.min() < 2 or kk.min() < 2 or ii.max() > mx - 3 or jj.max() > my - 3 or
                    kk.max() > mz - 3):
        raise ValueError("IB nodes must keep a cell of fluid between them and the domain boundary")
    return kk, jj, ii


def _intercept(cent, ii, jj, kk, p, nrm, idx, val):
    """The interception triangle of each node's ray p + t nrm (t > 0) among
    the six 3x3 patches of neighbouring cell centres, in InterceptionPoint's
    order: cells into idx[:, 3:12], weights into val[:, 0:3], d_i into
    val[:, 4]."""
    n = len(ii)
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
            b1, b2, b3 = _barycentric(x, p1, p2, p3)
            ok &= (b1 >= -1e-12) & (b2 >= -1e-12) & (b3 >= -1e-12)
            hit = q[ok]
            for s_, pt in enumerate(pts):
                idx[hit, 3 + 3 * s_:6 + 3 * s_] = pt[ok]
            val[hit, 0], val[hit, 1], val[hit, 2] = b1[ok], b2[ok], b3[ok]
            val[hit, 4] = t[ok]
            todo[hit] = False
    if todo.any():
        raise ValueError(f"{int(todo.sum())} IB nodes found no interception triangle")


def _barycentric(x, p1, p2, p3):
    """Weights of x (in the triangles' planes) on the vertices p1, p2, p3."""
    e1, e2, w = p2 - p1, p3 - p1, x - p1
    d00, d01, d11 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    d20, d21 = (w * e1).sum(1), (w * e2).sum(1)
    den = d00 * d11 - d01 * d01
    b2 = (d11 * d20 - d01 * d21) / den
    b3 = (d00 * d21 - d01 * d20) / den
    return 1.0 - b2 - b3, b2, b3


def sphere_ib_nodes(case: cases.ChannelCase, xyz: np.ndarray, nvert: np.ndarray):
    """(idx, val): idx (n, 12) int32 = i j k, i1 j1 k1, i2 j2 k2, i3 j3 k3 and
    val (n, 12) float64 = cr1 cr2 cr3, d_s, d_i, ua[3], dun, nf[3], one row
    per IB node (nvert 1.4) in the reference's k, j, i loop order."""
    if case.sphere is None:
        return np.zeros((0, 12), np.int32), np.zeros((0, 12))
    cent = _cell_centres(case, xyz)
    kk, jj, ii = _ib_cells(case, nvert)
    cx, cy, cz, d = case.sphere
    c0, R = np.array([cx, cy, cz]), 0.5 * d
    p = cent[kk, jj, ii]
    v = p - c0
    r = np.linalg.norm(v, axis=1)
    nrm = v / r[:, None]
    n = len(kk)
    idx = np.zeros((n, 12), np.int32)
    val = np.zeros((n, 12))
    idx[:, 0], idx[:, 1], idx[:, 2] = ii, jj, kk
    val[:, 3] = r - R
    val[:, 9:12] = nrm
    _intercept(cent, ii, jj, kk, p, nrm, idx, val)
    return idx, val


# ------------------------------------------------------------------
# triangulated surfaces (the reference's Icem nlist/elist bodies)
# ------------------------------------------------------------------
def icosphere(center, radius, level=3):
    """A sphere surface of 20 * 4^level triangles (an icosahedron subdivided
    level times, vertices projected onto the sphere), outward-oriented."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(x, dtype=np.float64) / np.linalg.norm(x) for x in v]
    for _ in range(level):
        mid = {}
        nf = []

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    vv = np.asarray(center, dtype=np.float64) + radius * np.array(verts)
    return vv, np.array(f, dtype=np.int64)


def _closest_points(p, a, b, c):
    """The point of each triangle (a, b, c) nearest p (row-paired) and its
    barycentric weights."""
    b1, b2, b3 = _barycentric(p - np.einsum("ij,ij->i", p - a, _unit(np.cross(b - a, c - a)))[:, None] *
                              _unit(np.cross(b - a, c - a)), a, b, c)
    inside = (b1 >= 0) & (b2 >= 0) & (b3 >= 0)
    best_q = b1[:, None] * a + b2[:, None] * b + b3[:, None] * c
    best_w = np.stack([b1, b2, b3], axis=1)
    best_d = np.where(inside, np.linalg.norm(p - best_q, axis=1), np.inf)
    for (u, v, iu, iv) in ((a, b, 0, 1), (b, c, 1, 2), (c, a, 2, 0)):
        e = v - u
        tt = np.clip(np.einsum("ij,ij->i", p - u, e) / np.einsum("ij,ij->i", e, e), 0.0, 1.0)
        q = u + tt[:, None] * e
        dd = np.linalg.norm(p - q, axis=1)
        better = dd < best_d
        w = np.zeros_like(best_w)
        w[:, iu], w[:, iv] = 1.0 - tt, tt
        best_q = np.where(better[:, None], q, best_q)
        best_w = np.where(better[:, None], w, best_w)
        best_d = np.where(better, dd, best_d)
    return best_q, best_w, best_d


def _unit(x):
    return x / np.linalg.norm(x, axis=1)[:, None]


def surface_ib_nodes(case: cases.ChannelCase, xyz: np.ndarray, nvert: np.ndarray, vertices, triangles, u=None,
                     uold=None, candidates=16):
    """IB node records from a triangulated body (as read by
    formats.read_icem): for each IB node the nearest surface element and
    point pmin on it, with its barycentric weights cs1..cs3 (the foot-point
    search of ibm_search_advanced); d_s = |centre - pmin|; the interception
    along (centre - pmin) / d_s as for the sphere; ua = sum of the element's
    vertex velocities u times cs, dun = (ua - uaold) . normal, and nf the
    element normal.  u, uold: vertex velocities (n_v, 3), 0 if None."""
    from scipy.spatial import cKDTree
    v = np.asarray(vertices, dtype=np.float64)
    t = np.asarray(triangles, dtype=np.int64)
    cent = _cell_centres(case, xyz)
    kk, jj, ii = _ib_cells(case, nvert)
    n = len(kk)
    idx = np.zeros((n, 12), np.int32)
    val = np.zeros((n, 12))
    if n == 0:
        return idx, val
    p = cent[kk, jj, ii]
    tc = v[t].mean(axis=1)
    k = min(candidates, len(t))
    _, cand = cKDTree(tc).query(p, k=k)
    cand = cand.reshape(n, k)
    best_d = np.full(n, np.inf)
    best_q = np.zeros((n, 3))
    best_w = np.zeros((n, 3))
    best_e = np.zeros(n, dtype=np.int64)
    for c in range(k):
        e = cand[:, c]
        q, w, d = _closest_points(p, v[t[e, 0]], v[t[e, 1]], v[t[e, 2]])
        better = d < best_d
        best_d = np.where(better, d, best_d)
        best_q = np.where(better[:, None], q, best_q)
        best_w = np.where(better[:, None], w, best_w)
        best_e = np.where(better, e, best_e)
    nrm = (p - best_q) / best_d[:, None]
    idx[:, 0], idx[:, 1], idx[:, 2] = ii, jj, kk
    val[:, 3] = best_d
    vu = np.zeros_like(v) if u is None else np.asarray(u, dtype=np.float64)
    vo = np.zeros_like(v) if uold is None else np.asarray(uold, dtype=np.float64)
    tri = t[best_e]
    ua = np.einsum("nk,nkc->nc", best_w, vu[tri])
    uo = np.einsum("nk,nkc->nc", best_w, vo[tri])
    val[:, 5:8] = ua
    val[:, 8] = np.einsum("nc,nc->n", ua, nrm) - np.einsum("nc,nc->n", uo, nrm)
    from .formats import element_normals
    val[:, 9:12] = element_normals(v, t)[best_e]
    _intercept(cent, ii, jj, kk, p, nrm, idx, val)
    return idx, val




def _pseudo_normals(v, t):
    """Angle-weighted pseudo-normals of a closed, outward-oriented surface
    (Baerentzen and Aanaes, IEEE TVCG 11(3), 2005): per facet its unit
    normal; per facet edge (the edge opposite corner q) the sum of the two
    facets sharing it; per vertex the sum of the incident facets' normals
    weighted by their angles there."""
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    fn = _unit(np.cross(b - a, c - a))
    vn = np.zeros_like(v)
    for q in range(3):
        e1 = v[t[:, (q + 1) % 3]] - v[t[:, q]]
        e2 = v[t[:, (q + 2) % 3]] - v[t[:, q]]
        ang = np.arccos(np.clip(np.einsum("nc,nc->n", _unit(e1), _unit(e2)), -1.0, 1.0))
        np.add.at(vn, t[:, q], ang[:, None] * fn)
    ends = np.stack([np.sort(t[:, [(q + 1) % 3, (q + 2) % 3]], axis=1) for q in range(3)], axis=1)
    key, inv = np.unique(ends.reshape(-1, 2), axis=0, return_inverse=True)
    en = np.zeros((len(key), 3))
    np.add.at(en, inv.ravel(), np.repeat(fn, 3, axis=0))
    return fn, en[inv.reshape(-1, 3)], vn


def surface_nvert(case: cases.ChannelCase, xyz: np.ndarray, vertices, triangles, candidates=16):
    """nvert of a closed, outward-oriented triangulated body, as
    ibm_search_advanced (ibm.c:4197) classifies the cells: a cell whose
    centre lies inside the body is solid (3.4), a non-solid cell with any of
    its 26 neighbours solid is an IB node (1.4, the rule of ibm.c:4410-4433),
    else fluid (0); periodic ghost planes repeat the interior.

    Inside or outside: the reference casts a ray and counts facet crossings
    (point_cell_advanced, ibm.c:1285).  Here the centre is inside when it
    lies behind the surface at its nearest point, judged against that
    point's angle-weighted pseudo-normal (facet, edge or vertex, as the
    nearest point falls); for a closed surface the two tests agree."""
    from scipy.spatial import cKDTree
    v = np.asarray(vertices, dtype=np.float64)
    t = np.asarray(triangles, dtype=np.int64)
    mx, my, mz = case.storage
    cent = cases.cell_centres(case, xyz)
    lo, hi = v.min(axis=0), v.max(axis=0)
    box = np.all((cent >= lo) & (cent <= hi), axis=-1)
    pts = cent[box]
    inside = np.zeros(len(pts), dtype=bool)
    if len(pts):
        fn, en, vn = _pseudo_normals(v, t)
        k = min(candidates, len(t))
        _, cand = cKDTree(v[t].mean(axis=1)).query(pts, k=k)
        cand = cand.reshape(len(pts), k)
        best_d = np.full(len(pts), np.inf)
        best_q = np.zeros((len(pts), 3))
        best_n = np.zeros((len(pts), 3))
        for c in range(k):
            e = cand[:, c]
            q, w, dd = _closest_points(pts, v[t[e, 0]], v[t[e, 1]], v[t[e, 2]])
            live = w > 0
            n_live = live.sum(axis=1)
            nrm = fn[e]
            on_edge = n_live == 2
            nrm = np.where(on_edge[:, None], en[e, np.argmin(live, axis=1)], nrm)
            corner = t[e, np.argmax(w, axis=1)]
            nrm = np.where((n_live == 1)[:, None], vn[corner], nrm)
            better = dd < best_d
            best_d = np.where(better, dd, best_d)
            best_q = np.where(better[:, None], q, best_q)
            best_n = np.where(better[:, None], nrm, best_n)
        inside = np.einsum("nc,nc->n", pts - best_q, best_n) < 0
    core = np.zeros((mz, my, mx), dtype=bool)
    inner = np.zeros(cent.shape[:3], dtype=bool)
    inner[box] = inside
    core[1:-1, 1:-1, 1:-1] = inner
    near = np.zeros_like(core)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if dk == dj == di == 0:
                    continue
                near[1:-1, 1:-1, 1:-1] |= core[1 + dk:mz - 1 + dk, 1 + dj:my - 1 + dj, 1 + di:mx - 1 + di]
    nv = np.zeros((mz, my, mx))
    nv[near & ~core] = 1.4
    nv[core] = 3.4
    cases._wrap_periodic_planes(case, nv)
    return nv
