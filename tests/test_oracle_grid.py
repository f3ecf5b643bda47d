"""Pins of the oracle's grid routines beyond the closed forms of
test_oracle_pins.py (the reference ships no tests or golden vectors):

  FormMetrics (metrics.c:80-157)   the geometric conservation law that the
                 reference's MetricsDivergence (metrics.c:1551-1625) checks:
                 the six face area vectors of every cell sum to zero, on the
                 warped, stretched channel and on randomly jittered nodes
  MyNvertRestriction (poisson.c:1662-1773)   a direct numpy evaluation of
                 the two passes on random fine masks, bit for bit on every
                 coarse level, with full and semi-coarsening"""
import numpy as np
import pytest

from curvib_old_amd import cases
from oracle import pyoracle as po


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_face_vectors_close_every_cell(jitter):
    """Each face vector is half the cross product of its diagonals, so the
    outward vectors of a hexahedron with any node positions sum to zero.
    The j- face vector of the lowest row is never computed (its eta stays 0,
    metrics.c:107-110), so the check starts at the second row."""
    case = cases.ChannelCase(cells=(10, 12, 8), amp=0.1, beta=2.0, sphere=None, bctype=(1, 1, 1, 1, 1, 1))
    xyz = cases.node_coords(case)
    if jitter:
        h = 1.0 / 12
        rng = np.random.default_rng(3)
        xyz = xyz + jitter * h * rng.uniform(-1, 1, xyz.shape)
    o = po.Oracle(case.nodes, case.bctype, 1, case.re, case.dt)
    o.set_grid(xyz)
    csi, eta, zet = o.get("csi"), o.get("eta"), o.get("zet")
    s = (csi[1:-1, 2:-1, 1:-1] - csi[1:-1, 2:-1, :-2] + eta[1:-1, 2:-1, 1:-1] - eta[1:-1, 1:-2, 1:-1] +
         zet[1:-1, 2:-1, 1:-1] - zet[:-2, 2:-1, 1:-1])
    scale = np.abs(csi).max()
    assert np.abs(s).max() < 1e-13 * scale, np.abs(s).max() / scale
    # the reference's own form, |sum of all components| * aj
    aj = o.get("aj")[1:-1, 2:-1, 1:-1]
    assert np.abs(s.sum(-1) * aj).max() < 1e-11


def nvert_restrict_direct(nf, shape_c, semi):
    """MyNvertRestriction of the reference (poisson.c:1705-1764): a coarse
    interior cell is blocked (1) when the product of nvert over its fine
    children -- the 2 (or 1, on a semi-coarsened axis) children per axis
    ending at fine index 2i (i) -- exceeds 0.1; then a still-open coarse
    cell becomes 1 when on each axis its two neighbours sum to more than
    1.1, neighbours read from the result of the first pass."""
    mzc, myc, mxc = shape_c
    nc = np.zeros(shape_c)
    step = [1 if s else 2 for s in semi]        # i, j, k
    width = [1 if s else 2 for s in semi]
    for k in range(1, mzc - 1):
        kh = step[2] * k
        for j in range(1, myc - 1):
            jh = step[1] * j
            for i in range(1, mxc - 1):
                ih = step[0] * i
                block = nf[kh - width[2] + 1:kh + 1, jh - width[1] + 1:jh + 1, ih - width[0] + 1:ih + 1]
                if np.prod(block) > 0.1:
                    nc[k, j, i] = 1.0
    first = nc.copy()
    c = first[1:-1, 1:-1, 1:-1]
    xs = first[1:-1, 1:-1, 2:] + first[1:-1, 1:-1, :-2]
    ys = first[1:-1, 2:, 1:-1] + first[1:-1, :-2, 1:-1]
    zs = first[2:, 1:-1, 1:-1] + first[:-2, 1:-1, 1:-1]
    fill = (c < 0.1) & (xs > 1.1) & (ys > 1.1) & (zs > 1.1)
    nc[1:-1, 1:-1, 1:-1][fill] = 1.0
    return nc


@pytest.mark.parametrize("bctype,semi,auto,seed", [
    ((7, 7, 1, 1, 7, 7), (0, 0, 0), 0, 1),
    ((1, 1, 1, 1, 1, 1), (0, 1, 0), 0, 2),
    ((7, 7, 1, 1, 7, 7), (0, 1, 0), 1, 3),
    ((1, 1, 1, 1, 1, 1), (1, 0, 0), 0, 4),
])
def test_nvert_restriction_matches_direct_evaluation(bctype, semi, auto, seed):
    cells = (8, 8, 8)
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, bctype=bctype)
    o = po.Oracle(case.nodes, case.bctype, 3, case.re, case.dt, semi, auto)
    o.set_grid(cases.node_coords(case))
    mx, my, mz = o.dims(2)[:3]
    rng = np.random.default_rng(seed)
    nf = np.zeros((mz, my, mx))
    vals = rng.choice([0.0, 1.4, 3.4], p=[0.25, 0.35, 0.4], size=(mz - 2, my - 2, mx - 2))
    nf[1:-1, 1:-1, 1:-1] = vals
    o.set_nvert(nf)
    fine = o.get("nvert", 2)
    blocked_total = 0
    for lev in (1, 0):
        dims_f, dims_c = o.dims(lev + 1)[:3], o.dims(lev)[:3]
        lsemi = tuple(int(dims_c[a] == dims_f[a]) for a in range(3))
        exp = nvert_restrict_direct(fine, dims_c[::-1], lsemi)
        got = o.get("nvert", lev)
        np.testing.assert_array_equal(got, exp, err_msg=f"level {lev} semi {lsemi}")
        blocked_total += int((exp > 0.5).sum())
        fine = got
    assert blocked_total > 0
