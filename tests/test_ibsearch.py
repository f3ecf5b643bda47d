"""IB node records (curvib_old_amd.ibsearch) against the geometry: the
analytic sphere's records (foot distance = |centre - c| - R, barycentric
interception weights, interception cells that are fluid or IB nodes), and a
triangulated sphere in Icem form against the exact sphere.  The records
are inputs of the reconstruction; its GPU parity with the oracle is in
test_gpu_ib.py."""
import numpy as np

from curvib_old_amd import cases, ibsearch


def test_sphere_records_are_consistent():
    case = cases.ChannelCase(cells=(40, 32, 24), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.4), seed=3)
    xyz, nv, _ = cases.build(case)
    idx, val = ibsearch.sphere_ib_nodes(case, xyz, nv)
    assert len(idx) > 100
    assert np.array_equal(nv[idx[:, 2], idx[:, 1], idx[:, 0]], np.full(len(idx), 1.4))
    cent = ibsearch._cell_centres(case, xyz)
    p = cent[idx[:, 2], idx[:, 1], idx[:, 0]]
    np.testing.assert_allclose(val[:, 3], np.linalg.norm(p - 0.5, axis=1) - 0.2, rtol=1e-14)
    np.testing.assert_allclose(val[:, :3].sum(1), 1.0, rtol=0, atol=1e-12)
    assert val[:, :3].min() > -1e-12 and (val[:, 4] > 0).all()
    for s in range(3):
        st = idx[:, 3 + 3 * s:6 + 3 * s]
        assert (nv[st[:, 2], st[:, 1], st[:, 0]] < 3).all()      # never inside the body
        # the interception point lies on the ray: centre + d_i * normal
    x = sum(val[:, q:q + 1] * cent[idx[:, 5 + 3 * q], idx[:, 4 + 3 * q], idx[:, 3 + 3 * q]] for q in range(3))
    np.testing.assert_allclose(x, p + val[:, 4:5] * val[:, 9:12], rtol=0, atol=1e-12)



def test_surface_ib_nodes_agree_with_the_exact_sphere():
    """IB records from a triangulated sphere (Icem surface) against the exact
    sphere's: foot distances within the mesh's chord error, normals within
    its facet angle; a rigidly translating surface gives ua = U exactly and
    dun = (U - Uold) . n."""
    from curvib_old_amd import ibsearch
    case = cases.ChannelCase(cells=(32, 24, 16), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.45), seed=21)
    xyz, nv, _ = cases.build(case)
    i1, v1 = ibsearch.sphere_ib_nodes(case, xyz, nv)
    R = 0.225
    vv, tt = ibsearch.icosphere((0.5, 0.5, 0.5), R, 5)
    edge = np.linalg.norm(vv[tt[:, 0]] - vv[tt[:, 1]], axis=1).max()
    U, Uo = np.array([0.3, -0.1, 0.2]), np.array([0.25, -0.1, 0.0])
    i2, v2 = ibsearch.surface_ib_nodes(case, xyz, nv, vv, tt, u=np.tile(U, (len(vv), 1)),
                                       uold=np.tile(Uo, (len(vv), 1)))
    assert np.array_equal(i1[:, :3], i2[:, :3])
    chord = R - np.sqrt(R * R - (edge / np.sqrt(3)) ** 2)
    assert np.abs(v1[:, 3] - v2[:, 3]).max() <= chord
    assert np.einsum("nc,nc->n", v1[:, 9:12], v2[:, 9:12]).min() > np.cos(2 * edge / R)
    np.testing.assert_allclose(v2[:, 5:8], np.broadcast_to(U, (len(v2), 3)), rtol=0, atol=1e-15)
    # the ray normal is (centre - pmin) / d_s; dun = (U - Uo) . n
    cent = ibsearch._cell_centres(case, xyz)
    p = cent[i2[:, 2], i2[:, 1], i2[:, 0]]
    assert (np.abs(v2[:, 8]) <= np.linalg.norm(U - Uo) * (1 + 1e-12)).all()
    assert np.abs(np.abs(v2[:, 8]) - np.abs((p - np.array([0.5, 0.5, 0.5])) @ (U - Uo) /
                                            np.linalg.norm(p - np.array([0.5, 0.5, 0.5]), axis=1))).max() < 0.05
    # interception weights are barycentric: sum to 1, in [0, 1]
    np.testing.assert_allclose(v2[:, :3].sum(1), 1.0, rtol=0, atol=1e-12)
    assert v2[:, :3].min() > -1e-12 and (v2[:, 4] > 0).all()



def _torus(R, r, nu=64, nv=32, c=(0.5, 0.5, 0.5)):
    """Closed, outward-oriented torus about the z axis through c."""
    u = 2 * np.pi * np.arange(nu) / nu
    v = 2 * np.pi * np.arange(nv) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    pts = np.stack([(R + r * np.cos(V)) * np.cos(U), (R + r * np.cos(V)) * np.sin(U), r * np.sin(V)], -1)
    pts = pts.reshape(-1, 3) + np.asarray(c)
    a = np.arange(nu)[:, None] * nv + np.arange(nv)[None, :]
    b = ((np.arange(nu)[:, None] + 1) % nu) * nv + np.arange(nv)[None, :]
    a1 = np.arange(nu)[:, None] * nv + (np.arange(nv)[None, :] + 1) % nv
    b1 = ((np.arange(nu)[:, None] + 1) % nu) * nv + (np.arange(nv)[None, :] + 1) % nv
    tri = np.concatenate([np.stack([a, b, b1], -1).reshape(-1, 3), np.stack([a, b1, a1], -1).reshape(-1, 3)])
    return pts, tri


def _core_centres(case, xyz, nv):
    return cases.cell_centres(case, xyz), nv[1:-1, 1:-1, 1:-1] > 3


def test_surface_nvert_of_a_convex_body_is_exact():
    """Convex icosphere: a centre is inside iff it lies behind every facet
    plane (brute force over all facets), and the IB nodes are the non-solid
    cells with a solid one among their 26 neighbours."""
    case = cases.ChannelCase(cells=(24, 20, 16), amp=0.1, beta=2.0, sphere=None, seed=5)
    xyz, _, _ = cases.build(case)
    vv, tt = ibsearch.icosphere((0.5, 0.5, 0.5), 0.3, 2)
    nv = ibsearch.surface_nvert(case, xyz, vv, tt)
    cent, core = _core_centres(case, xyz, nv)
    a, b, c = vv[tt[:, 0]], vv[tt[:, 1]], vv[tt[:, 2]]
    n = np.cross(b - a, c - a)
    side = np.einsum("fc,xfc->xf", n, cent.reshape(-1, 1, 3) - a[None])
    inside = (side < 0).all(axis=1).reshape(core.shape)
    assert inside.sum() > 100
    assert np.array_equal(core, inside)
    ib = np.abs(nv - 1.4) < 1e-12
    solid = nv > 3
    pad = np.pad(solid, 1)
    nb = np.zeros_like(solid)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if (dk, dj, di) != (0, 0, 0):
                    nb |= pad[1 + dk:pad.shape[0] - 1 + dk, 1 + dj:pad.shape[1] - 1 + dj, 1 + di:pad.shape[2] - 1 + di]
    inner = (slice(1, -1),) * 3
    assert np.array_equal(ib[inner], (nb & ~solid)[inner])


def test_surface_nvert_matches_the_exact_sphere_and_torus():
    """Against the analytic bodies: every centre farther from the exact
    surface than the mesh's chord error is classified as the exact body
    classifies it.  The torus is not convex: its hole must stay fluid."""
    case = cases.ChannelCase(cells=(64, 48, 48), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.45), seed=21)
    xyz, nv_exact, _ = cases.build(case)
    R = 0.225
    vv, tt = ibsearch.icosphere((0.5, 0.5, 0.5), R, 5)
    nv = ibsearch.surface_nvert(case, xyz, vv, tt)
    edge = np.linalg.norm(vv[tt[:, 0]] - vv[tt[:, 1]], axis=1).max()
    chord = R - np.sqrt(R * R - (edge / np.sqrt(3)) ** 2)
    cent, core = _core_centres(case, xyz, nv)
    far = np.abs(np.linalg.norm(cent - 0.5, axis=-1) - R) > chord
    assert np.array_equal(core[far], (nv_exact[1:-1, 1:-1, 1:-1] > 3)[far])
    assert (core != (nv_exact[1:-1, 1:-1, 1:-1] > 3)).sum() <= 8

    Rt, rt = 0.28, 0.1
    vv, tt = _torus(Rt, rt, 128, 48)
    nv = ibsearch.surface_nvert(case, xyz, vv, tt)
    cent, core = _core_centres(case, xyz, nv)
    d = cent - 0.5
    dist = np.hypot(np.hypot(d[..., 0], d[..., 1]) - Rt, d[..., 2]) - rt
    band = 2 * np.pi * (Rt + rt) / 128 + 2 * np.pi * rt / 48
    far = np.abs(dist) > band ** 2 / rt
    assert core[far].sum() > 200
    assert np.array_equal(core[far], (dist < 0)[far])
    hole = np.hypot(d[..., 0], d[..., 1]) < Rt - rt - 0.02
    assert hole.sum() > 100 and not core[hole].any()


def test_surface_nvert_of_a_box_is_exact_at_edges_and_corners():
    """A box of 12 facets has no chord error: every centre, including those
    whose nearest surface point is on an edge or a corner, is classified as
    the box itself classifies it; a rotated copy as well."""
    case = cases.ChannelCase(cells=(40, 32, 32), amp=0.1, beta=2.0, sphere=None, seed=9)
    xyz, _, _ = cases.build(case)
    corners = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], dtype=float)
    faces = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tri = np.array([f for a, b, c, d in faces for f in ((a, b, c), (a, c, d))])
    lo, ext = np.array([0.31, 0.27, 0.33]), np.array([0.37, 0.41, 0.29])
    th = 0.6
    rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    for R in (np.eye(3), rot):
        ctr = lo + ext / 2
        vv = (corners * ext + lo - ctr) @ R.T + ctr
        nv = ibsearch.surface_nvert(case, xyz, vv, tri)
        cent, core = _core_centres(case, xyz, nv)
        local = (cent - ctr) @ R + ctr
        exact = np.all((local > lo) & (local < lo + ext), axis=-1)
        assert exact.sum() > 500
        assert np.array_equal(core, exact)
