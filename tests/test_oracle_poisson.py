"""Closed-form pins of the oracle's VolumeFlux_rev and Divergence (the
reference ships no tests or golden vectors for this path):

  VolumeFlux_rev (poisson.c:3712-3944)   after the call the net flux from
                 fluid cells into immersed-boundary cells is zero to
                 round-off (the reference's own "IBMFlux22 REV" re-check,
                 poisson.c:3933); every interface face moves by one common
                 correction times its face area, and no other face moves
  Divergence (main.c:469-593)   on fluxes with a known divergence the
                 maximum |div| is the planted value, taken at the first
                 cell in natural order on a tie; cells next to a body are
                 skipped; boundary storage contributes 0

Each pin fails a plausible transcription mistake: a correction scaled or
signed wrongly, a face family left out, a divergence missing its aj factor
or a term, the last instead of the first index of a tie."""
import numpy as np
import pytest

from curvib_old_amd import cases
from oracle import pyoracle as po


def oracle_for(case, nlev=1):
    xyz, nv, uc = cases.build(case)
    o = po.Oracle(case.nodes, case.bctype, nlev, case.re, case.dt)
    o.set_grid(xyz)
    o.set_nvert(nv)
    o.put("ucont", uc)
    return o, xyz, nv


def _interface_faces(nv):
    """Faces between an interior fluid cell (nvert < 0.1) and an interior
    immersed-boundary cell (1.1 < nvert < 1.5), per family d: a boolean
    mask on the face's lower cell and the orientation (+1 when the lower
    cell is the fluid one, so that a positive flux leaves the fluid)."""
    fl = nv < 0.1
    ib = (nv > 1.1) & (nv < 1.5)
    mz, my, mx = nv.shape
    inner = np.zeros_like(fl)
    inner[1:-1, 1:-1, 1:-1] = True
    out = []
    for d in range(3):
        sh = [0, 0, 0]
        sh[2 - d] = 1
        up = np.roll(inner, -1, axis=2 - d) & inner
        a_fl = fl & up & np.roll(ib, -1, axis=2 - d)
        a_ib = ib & up & np.roll(fl, -1, axis=2 - d)
        orient = np.zeros(nv.shape)
        orient[a_fl] = 1.0
        orient[a_ib] = -1.0
        out.append(orient)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_volumeflux_rev_cancels_the_net_ib_flux(seed):
    case = cases.ChannelCase(cells=(16, 14, 12), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.45), seed=seed)
    o, xyz, nv = oracle_for(case)
    rng = np.random.default_rng(seed)
    uc0 = o.get("ucont") + rng.uniform(-1e-3, 1e-3, o.get("ucont").shape)
    o.put("ucont", uc0)
    orient = _interface_faces(nv)
    area = [np.linalg.norm(o.get(f), axis=-1) for f in ("csi", "eta", "zet")]
    net0 = sum((orient[d] * uc0[..., d]).sum() for d in range(3))
    atot = sum((np.abs(orient[d]) * area[d]).sum() for d in range(3))
    assert abs(net0) > 1e-6 and atot > 0
    o.call("o_volumeflux_rev")
    uc1 = o.get("ucont")
    net1 = sum((orient[d] * uc1[..., d]).sum() for d in range(3))
    mag = sum((np.abs(orient[d] * uc1[..., d])).sum() for d in range(3))
    assert abs(net1) < 1e-13 * mag, (net0, net1)
    corr = net0 / atot
    for d in range(3):
        on = orient[d] != 0
        np.testing.assert_allclose(uc1[..., d][on], uc0[..., d][on] - orient[d][on] * area[d][on] * corr,
                                   rtol=1e-13, atol=1e-17)
        np.testing.assert_array_equal(uc1[1:-1, 1:-1, 1:-1, d][~on[1:-1, 1:-1, 1:-1]],
                                      uc0[1:-1, 1:-1, 1:-1, d][~on[1:-1, 1:-1, 1:-1]])
    st = o.stats()
    assert st["ibm_flux0"] == pytest.approx(net0, rel=1e-12)
    assert st["ibm_area0"] == pytest.approx(atot, rel=1e-12)


def test_divergence_finds_the_planted_maximum():
    """Uniform grid of dyadic spacing (so the sums below are exact, and
    the ties are true ties), Ucont.x = q x_face dy dz gives div = q in every
    cell.
    Two faces raise the divergence of their lower cell by the same amount
    D (and lower the cell above by D); a third, larger bump sits next to
    immersed-boundary cells, which the rule skips."""
    cells = (8, 8, 4)
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, bctype=(1, 1, 1, 1, 1, 1))
    o, xyz, nv = oracle_for(case)
    mx, my, mz = o.dims()[:3]
    dx, dy, dz = (1.0 / c for c in cells)
    vol = dx * dy * dz
    q, D = 0.25, 1.0
    uc = np.zeros((mz, my, mx, 3))
    uc[1:-1, 1:-1, :-1, 0] = (q * np.arange(mx - 1) * dx * dy * dz)[None, None, :]
    A = (3, 2, 5)            # (k, j, i) of the first planted maximum
    C = (2, 4, 6)            # ties with A, earlier in natural order
    B = (4, 3, 2)            # larger, beside an IB cell
    for (k, j, i), amp in ((A, D), (C, D)):
        uc[k, j, i, 1] += amp * vol
    uc[B[0], B[1], B[2], 1] += 3 * D * vol
    nvb = nv.copy()
    nvb[B[0], B[1]:B[1] + 2, B[2] - 1] = 1.4      # beside B and the cell above it
    o.set_nvert(nvb)
    o.put("ucont", uc)
    val = o.call("o_divergence")
    st = o.stats()
    nat = lambda k, j, i: i + mx * (j + my * k)      # noqa: E731
    assert val == pytest.approx(q + D, rel=1e-12)
    assert int(st["maxdiv_idx"]) == min(nat(*A), nat(*C)) == nat(*C)
    # without the ties, the other planted cell is found
    uc[C[0], C[1], C[2], 1] -= D * vol
    o.put("ucont", uc)
    assert o.call("o_divergence") == pytest.approx(q + D, rel=1e-12)
    assert int(o.stats()["maxdiv_idx"]) == nat(*A)
    # the skipped bump counts once the body is gone
    o.set_nvert(nv)
    assert o.call("o_divergence") == pytest.approx(q + 3 * D, rel=1e-12)
    assert int(o.stats()["maxdiv_idx"]) == nat(*B)


def test_poisson_rhs_of_a_known_divergence():
    """PoissonRHS (poisson.c:1509-1576): b = -div(Ucont) aj 1.5 / (dt st) on
    interior fluid cells, 0 on boundary storage and on cells with
    nvert > 0.1.  With Ucont.x = q x_face dy dz on a uniform grid every
    cell's divergence is q."""
    cells = (8, 8, 4)
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, dt=2e-3, bctype=(1, 1, 1, 1, 1, 1))
    o, xyz, nv = oracle_for(case)
    mx, my, mz = o.dims()[:3]
    dx, dy, dz = (1.0 / c for c in cells)
    q = 0.25
    uc = np.zeros((mz, my, mx, 3))
    uc[1:-1, 1:-1, :-1, 0] = (q * np.arange(mx - 1) * dx * dy * dz)[None, None, :]
    nvb = nv.copy()
    nvb[2, 3, 4] = 1.4
    nvb[3, 5, 2] = 3.4
    o.set_nvert(nvb)
    o.put("ucont", uc)
    o.call("o_rhs_into_b")
    b = o.get("b")
    exp = np.zeros((mz, my, mx))
    exp[1:-1, 1:-1, 1:-1] = -q * 1.5 / case.dt
    exp[nvb > 0.1] = 0.0
    np.testing.assert_allclose(b, exp, rtol=1e-13, atol=0)


@pytest.mark.parametrize("shear", [False, True])
def test_projection_subtracts_the_face_gradient_of_phi(shear):
    """Projection (poisson.c:2281-2906): with phi = g.x at every cell centre
    of the lattice x = M (i dx, j dy, k dz) the face's derivatives of phi
    along the index directions are exactly g.a, g.b, g.c, so each open face
    loses (A_d.g) dt st / 1.5 of flux (the cofactor identity, as in
    FormFunction1's pressure gradient).  Faces touching an immersed-boundary
    or solid cell (nvert_i + nvert_i+1 != 0) keep their flux."""
    cells = (7, 6, 5)
    M = np.array([[1.0, 0.3, 0.1], [0.2, 1.0, 0.05], [0.1, 0.2, 1.0]]) if shear else np.eye(3)
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, bctype=(1, 1, 1, 1, 1, 1), dt=3e-3)
    xyz = cases.node_coords(case) @ M.T
    a, b, c = (M[:, d] / cells[d] for d in range(3))
    o = po.Oracle(case.nodes, case.bctype, 1, case.re, case.dt)
    o.set_grid(xyz)
    mx, my, mz = o.dims()[:3]
    nv = np.zeros((mz, my, mx))
    nv[3, 3, 4] = 1.4
    o.set_nvert(nv)
    g = np.array([0.7, -1.3, 0.4])
    K, J, I = np.meshgrid(np.arange(mz) - 0.5, np.arange(my) - 0.5, np.arange(mx) - 0.5, indexing="ij")
    phi = (I[..., None] * a + J[..., None] * b + K[..., None] * c) @ g
    rng = np.random.default_rng(5)
    uc0 = rng.uniform(-1, 1, (mz, my, mx, 3))
    o.put("ucont", uc0)
    o.put("phi", phi)
    o.call("o_projection")
    uc = o.get("ucont")
    area = (np.cross(b, c), np.cross(c, a), np.cross(a, b))
    sel = (np.s_[1:-1, 1:-1, 1:-2], np.s_[1:-1, 1:-2, 1:-1], np.s_[1:-2, 1:-1, 1:-1])
    for d in range(3):
        sh = [0, 0, 0]
        sh[2 - d] = 1
        touch = (nv + np.roll(nv, -1, axis=2 - d)) != 0
        exp = uc0[..., d] - float(area[d] @ g) * case.dt / 1.5
        exp[touch] = uc0[..., d][touch]
        np.testing.assert_allclose(uc[sel[d] + (d,)], exp[sel[d]], rtol=1e-12, atol=1e-14,
                                   err_msg=f"face family {d}")
        assert touch[sel[d]].sum() == 2


def test_update_pressure_adds_phi():
    """UpdatePressure (poisson.c:2908-3248): p += phi on interior cells; the
    periodic boundary planes then repeat the wrapped interior planes."""
    case = cases.ChannelCase(cells=(8, 6, 5), amp=0.1, sphere=None)
    o, xyz, nv = oracle_for(case)
    mx, my, mz = o.dims()[:3]
    rng = np.random.default_rng(8)
    p0 = rng.standard_normal((mz, my, mx))
    phi = rng.standard_normal((mz, my, mx))
    o.put("p", p0)
    o.put("phi", phi)
    o.call("o_update_pressure")
    p = o.get("p")
    np.testing.assert_array_equal(p[1:-1, 1:-1, 1:-1], (p0 + phi)[1:-1, 1:-1, 1:-1])
    np.testing.assert_array_equal(p[1:-1, 1:-1, 0], p[1:-1, 1:-1, -2])
    np.testing.assert_array_equal(p[1:-1, 1:-1, -1], p[1:-1, 1:-1, 1])
    np.testing.assert_array_equal(p[0, 1:-1, :], p[-2, 1:-1, :])
    np.testing.assert_array_equal(p[-1, 1:-1, :], p[1, 1:-1, :])
