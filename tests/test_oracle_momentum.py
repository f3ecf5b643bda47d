"""Closed-form pins of the oracle's momentum routines (the reference ships no
tests or golden vectors for this path, so each check is fixed by the
mathematics, on grids where the answer is known exactly):

  Convection (QUICK, rhs.c:1636-1897)   on a uniform grid the QUICK face
                 value of a cubic profile is exact for its quadratic part
                 and off by (3/8) e h^3 sign(F) for its cubic part e x^3
                 (the Lagrange remainder of the upwind-biased quadratic),
                 checked through the convective flux differences of a
                 flux that varies along x and changes sign
  FormFunction1 pressure gradient (rhs.c:2799-3026)   with zero velocity
                 and p linear in x on a uniform or sheared (affine) grid,
                 the face right-hand side is -(face area vector).grad p
  Viscous (rhs.c:1947-2565)   a quadratic velocity on a sheared grid: the
                 viscous part of the face right-hand side (the difference of
                 two Reynolds numbers) is (nu1 - nu2) A.(lap u + grad div u)
  FormFunction1 basis (rhs.c:2685-2703, 2869)   on a lattice whose j face
                 areas vary from face to face, the face right-hand side of
                 a divergence-free shear wave carries the mean of the two
                 areas of each cell
  Free stream (metrics.c:80-157, rhs.c:1463-2565)   a uniform stream on a
                 warped, stretched grid gives a right-hand side at the
                 round-off scale away from the walls
  RungeKutta (implicitsolver.c:32-119)   on a shear wave u_x = sin(2 pi z)
                 the step is linear, du/dt = lambda u with lambda the
                 eigenvalue of the viscous second difference, and the four
                 stages with alpha = 1/4, 1/3, 1/2, 1 amplify the mode by
                 1 + s + s^2/2 + s^3/6 + s^4/24, s = lambda dt

Each pin fails a plausible transcription mistake: a QUICK coefficient other
than 1/8, an upwind branch taken the wrong way, the sign of the -G.grad p
term or its central weight 1/4, swapped RK stage coefficients."""
import numpy as np
import pytest

from curvib_old_amd import cases
from oracle import pyoracle as po

WALLS = (1, 1, 1, 1, 1, 1)


def oracle_for(case, nlev=1, xyz=None):
    xyz0, nv, uc = cases.build(case)
    o = po.Oracle(case.nodes, case.bctype, nlev, case.re, case.dt)
    o.set_grid(xyz0 if xyz is None else xyz)
    o.set_nvert(nv)
    o.put("ucont", uc)
    return o


@pytest.mark.parametrize("U", [0.8, -0.6])
def test_quick_face_values_of_a_cubic_profile(U):
    """w(x) = a + b x + c x^2 + e x^3 carried by an x-flux F_i that varies
    along x and changes sign, on a uniform periodic channel.  QUICK takes
    the face value from the quadratic through the two cells upwind and the
    one downwind: u0 + (1/8)(-u_{-1} - 2 u0 + 3 u_{+1}) for F > 0 and the
    mirror image for F < 0.  By the Lagrange remainder, that value is
    w(x_f) - e (x_f - x_{-1})(x_f - x_0)(x_f - x_{+1}) with the three
    upwind-biased cell centres, i.e. w(x_f) + (3/8) e h^3 sign(F): exact for
    the quadratic part, off by the upwind side's sign for the cubic part.
    So on the k faces
      Rhs.z = -(F_i q_i - F_{i-1} q_{i-1}) / dz,
      q_i = w(x_i + h/2) + (3/8) e h^3 sign(F_i),
    with viscosity negligible (Re = 1e15) and p = 0."""
    cells = (16, 8, 6)
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, re=1e15)
    o = oracle_for(case)
    mx, my, mz = o.dims()[:3]
    dx, dy, dz = (1.0 / c for c in cells)
    a, b, c, e = 0.3, -1.1, 2.7, 1.9
    wf = lambda x: a + b * x + c * x ** 2 + e * x ** 3      # noqa: E731
    xc = (np.arange(mx) - 0.5) * dx            # centre of storage cell i
    xf = np.arange(mx) * dx                    # i face i, between cells i and i+1
    F = U * dy * dz * (0.3 + np.cos(2 * np.pi * xf))
    uc = np.zeros((mz, my, mx, 3))
    uc[1:-1, 1:-1, :-1, 0] = F[:-1]
    uc[:-1, 1:-1, 1:-1, 2] = (wf(xc[1:-1]) * dx * dy)[None, None, :]
    o.put("ucont", uc)
    o.put("p", np.zeros((mz, my, mx)))
    o.call("o_contra2cart")
    o.call("o_formbcs")          # periodic and wall ghost velocities
    o.call("o_ff1_into_rhs")
    rhs = o.get("rhs")
    q = wf(xf) + 0.375 * e * dx ** 3 * np.sign(F)
    flux = F * q
    exp = np.zeros(mx)
    exp[1:] = -(flux[1:] - flux[:-1]) / dz
    # cells whose two i faces use stencils inside the periodic seam
    sl = np.s_[1:-1, 1:-1, 3:mx - 3]
    got = rhs[sl + (2,)]
    np.testing.assert_allclose(got, np.broadcast_to(exp[3:mx - 3], got.shape), rtol=1e-11,
                               atol=1e-13 * np.abs(exp).max())


def _affine(cells, M):
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, bctype=WALLS)
    xyz = cases.node_coords(case) @ M.T
    a, b, c = (M[:, d] / cells[d] for d in range(3))
    return case, xyz, (a, b, c)


@pytest.mark.parametrize("shear", [False, True])
def test_ff1_pressure_gradient_is_minus_area_dot_grad_p(shear):
    """Zero velocity (so Visc = Conv = 0) and p = p0 + g.x at every cell
    centre (ghosts included) of the lattice x = M (i dx, j dy, k dz).  The
    face's pressure derivatives along the three index directions are then
    exactly g.a, g.b, g.c (lattice vectors a, b, c), and the face right-hand
    side -(dpdc g11 + dpde g12 + dpdz g13) * aj_face collapses to
    -(face area vector).g: the cofactor identity
    (g.a)(b x c) + (g.b)(c x a) + (g.c)(a x b) = (a.(b x c)) g.  On the
    sheared grid every cross coefficient is nonzero."""
    cells = (7, 6, 5)
    M = np.array([[1.0, 0.3, 0.1], [0.2, 1.0, 0.05], [0.1, 0.2, 1.0]]) if shear else np.eye(3)
    case, xyz, (a, b, c) = _affine(cells, M)
    o = oracle_for(case, xyz=xyz)
    mx, my, mz = o.dims()[:3]
    g = np.array([0.7, -1.3, 0.4])
    K, J, I = np.meshgrid(np.arange(mz) - 0.5, np.arange(my) - 0.5, np.arange(mx) - 0.5, indexing="ij")
    cent = I[..., None] * a + J[..., None] * b + K[..., None] * c
    o.put("p", 2.0 + cent @ g)
    o.put("ucont", np.zeros((mz, my, mx, 3)))
    o.call("o_ff1_into_rhs")
    rhs = o.get("rhs")
    area = (np.cross(b, c), np.cross(c, a), np.cross(a, b))
    # faces between two interior cells
    sel = (np.s_[1:-1, 1:-1, 1:-2], np.s_[1:-1, 1:-2, 1:-1], np.s_[1:-2, 1:-1, 1:-1])
    for d in range(3):
        got = rhs[sel[d] + (d,)]
        exp = -float(area[d] @ g)
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-15, err_msg=f"face family {d}")


def test_viscous_term_of_a_quadratic_field_on_a_sheared_grid():
    """u_m(x) = g_m.x + x.H_m x / 2 on the affine lattice x = M (i, j, k) h.
    With Ucont = (face area vector).u(face centre), Contra2Cart returns u
    plus a constant (the average of two face values of a quadratic exceeds
    the centre value by a constant), so every velocity gradient the viscous
    rule forms -- normal differences and four-cell central tangential
    differences -- is exact at the face centre.  The viscous flux of a face
    is then nu (grad u + grad u^T).A and its sum over a closed cell is
    nu vol (lap u + grad div u), constant.  Convection does not depend on
    nu, so the difference of the right-hand sides at Re = 1 and Re = 2 is
    the viscous part alone:
      Rhs_d(1) - Rhs_d(2) = (1 - 1/2) A_d.L,  L_m = tr H_m + sum_n H_n[m, n].
    The sheared lattice makes every cross term and the transposed-gradient
    term r.n of the rule contribute."""
    cells = (10, 10, 10)
    M = np.array([[1.0, 0.3, 0.1], [0.2, 1.0, 0.05], [0.1, 0.2, 1.0]])
    rng = np.random.default_rng(7)
    g = rng.uniform(-1, 1, (3, 3))
    H = rng.uniform(-1, 1, (3, 3, 3))
    H = 0.5 * (H + H.transpose(0, 2, 1))

    def vel(x):
        return np.einsum("mn,...n->...m", g, x) + 0.5 * np.einsum("...n,mnp,...p->...m", x, H, x)

    rhs = []
    for re in (1.0, 2.0):
        case, xyz, (a, b, c) = _affine(cells, M)
        case.re = re
        o = oracle_for(case, xyz=xyz)
        mx, my, mz = o.dims()[:3]
        K, J, I = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
        uc = np.zeros((mz, my, mx, 3))
        for d, (off, A) in enumerate((((0, -0.5, -0.5), np.cross(b, c)), ((-0.5, 0, -0.5), np.cross(c, a)),
                                      ((-0.5, -0.5, 0), np.cross(a, b)))):
            xf = ((I + off[0])[..., None] * a + (J + off[1])[..., None] * b + (K + off[2])[..., None] * c)
            uc[..., d] = vel(xf) @ A
        o.put("ucont", uc)
        o.put("p", np.zeros((mz, my, mx)))
        o.call("o_contra2cart")
        o.call("o_formbcs")
        o.call("o_ff1_into_rhs")
        rhs.append(o.get("rhs"))
    L = np.einsum("mnn->m", H) + np.einsum("nmn->m", H)
    area = (np.cross(b, c), np.cross(c, a), np.cross(a, b))
    diff = rhs[0] - rhs[1]
    inner = np.s_[3:-4, 3:-4, 3:-4]
    for d in range(3):
        exp = 0.5 * float(area[d] @ L)
        np.testing.assert_allclose(diff[inner + (d,)], exp, rtol=1e-9, atol=1e-12 * abs(exp),
                                   err_msg=f"face family {d}")


def test_face_rhs_where_face_areas_vary_along_the_family():
    """The lattice x = xi (1 + 0.3 y), y = eta, z = zeta, periodic in z:
    the j faces' area eta = (0, L(y_j) dxi dz, 0), L = 1 + 0.3 y, changes from
    one j face to the next.  A shear wave u = (0, v(z), 0), v = sin(2 pi z),
    with Ucont = (face area vector).u: the flow is divergence free, so its
    convective fluxes cancel cell by cell (geometric conservation), and
    its viscous term is nu L(y_c) dxi dy times the second difference of v
    in z, over dz.  rct.y = aj (eta_j + eta_{j-1})/2 . Rc then carries the
    mean of the two j faces' areas, L(y_c), and the j face right-hand side,
    the mean of its two cells' rct.y, is
      Rhs.y = nu dxi L(y_j) (v_{k+1} - 2 v_k + v_{k-1}) / dz."""
    cells = (6, 12, 8)
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, re=1.0)
    xyz = cases.node_coords(case)
    xyz[..., 0] *= 1.0 + 0.3 * xyz[..., 1]
    o = oracle_for(case, xyz=xyz)
    mx, my, mz = o.dims()[:3]
    dxi, dy, dz = (1.0 / c for c in cells)
    v = np.sin(2 * np.pi * (np.arange(mz) - 0.5) * dz)
    uc = np.stack([o.get(f)[..., 1] for f in ("csi", "eta", "zet")], axis=-1) * v[:, None, None, None]
    uc[:, :, 0, 0] = uc[:, :, -2, 0]
    uc[0, :, :, 2] = uc[-2, :, :, 2]
    o.put("ucont", uc)
    o.put("p", np.zeros((mz, my, mx)))
    o.call("o_contra2cart")
    o.call("o_formbcs")
    o.call("o_ff1_into_rhs")
    rhs = o.get("rhs")
    vp = np.roll(v[1:-1], -1)
    vm = np.roll(v[1:-1], 1)
    d2v = vp - 2 * v[1:-1] + vm
    yj = np.arange(my) * dy              # the j face between cells j and j+1 lies at y = j dy
    exp = dxi * (1 + 0.3 * yj)[None, :] * d2v[:, None] / dz
    rows = np.s_[3:my - 4]
    got = rhs[1:-1, rows, 2:-2, 1]          # the tilt of the i faces is not periodic: skip the seam
    np.testing.assert_allclose(got, np.broadcast_to(exp[:, rows][..., None], got.shape), rtol=1e-11,
                               atol=1e-12 * np.abs(exp).max())


def test_free_stream_is_preserved_on_a_warped_grid():
    """A uniform stream U on the warped, stretched periodic channel:
    Ucont = (face area vector).U, Contra2Cart returns U, every velocity
    gradient vanishes and the convective fluxes sum to U (U.A) over each
    closed cell, which is the geometric conservation law.  Away from the
    no-slip walls (whose ghosts reverse U) the right-hand side is at the
    round-off scale."""
    case = cases.ChannelCase(cells=(12, 16, 10), amp=0.1, beta=2.0, sphere=None, re=100.0)
    o = oracle_for(case)
    c = cases._storage_coords(case, cases.node_coords(case))
    csi, eta, zet = cases._face_vectors(c)
    U = np.array([0.9, 0.0, 1.4])
    uc = np.stack([csi @ U, eta @ U, zet @ U], axis=-1)
    uc[:, :, 0, 0] = uc[:, :, -2, 0]
    uc[0, :, :, 2] = uc[-2, :, :, 2]
    o.put("ucont", uc)
    o.put("p", np.zeros(o.get("p").shape))
    o.call("o_contra2cart")
    o.call("o_formbcs")
    o.call("o_ff1_into_rhs")
    rhs = o.get("rhs")
    scale = np.abs(uc).max() * np.abs(U).max()
    inner = rhs[1:-1, 4:-5, 1:-1]
    assert np.abs(inner).max() < 1e-12 * scale, np.abs(inner).max() / scale


def test_rungekutta_amplifies_a_shear_wave_by_the_rk4_polynomial():
    """u_x = sin(2 pi z) on a uniform channel, periodic in x and z: the
    convective fluxes cancel exactly (u_x does not vary along x and there
    is no z flux) and the viscous term is nu times the second difference
    in z, whose eigenvalue on this mode is
      lambda = -nu (2 - 2 cos(2 pi dz)) / dz^2.
    Four stages U = U_o + alpha_s dt Rhs(U) with alpha = 1/4, 1/3, 1/2, 1
    multiply a linear mode by G(s) = 1 + s + s^2/2 + s^3/6 + s^4/24,
    s = lambda dt.  The no-slip walls reach 4 rows in 4 stages; rows in
    the middle of the channel see the pure mode."""
    cells = (4, 16, 8)
    re = 1.0
    dz = 1.0 / cells[2]
    lam = -(1.0 / re) * (2 - 2 * np.cos(2 * np.pi * dz)) / dz ** 2
    dt = -0.6 / lam
    case = cases.ChannelCase(cells=cells, beta=0.0, amp=0.0, sphere=None, re=re, dt=dt)
    o = oracle_for(case)
    mx, my, mz = o.dims()[:3]
    dx, dy = 1.0 / cells[0], 1.0 / cells[1]
    zc = (np.arange(mz) - 0.5) * dz
    u = np.sin(2 * np.pi * zc)
    uc = np.zeros((mz, my, mx, 3))
    uc[1:-1, 1:-1, :-1, 0] = (u[1:-1] * dy * dz)[:, None, None]
    o.put("ucont", uc)
    o.put("ucont_o", uc)
    o.put("p", np.zeros((mz, my, mx)))
    o.call("o_rungekutta")
    s = lam * dt
    G = 1 + s + s * s / 2 + s ** 3 / 6 + s ** 4 / 24
    got = o.get("ucont")
    mid = np.s_[1:-1, 6:12, :-1]
    np.testing.assert_allclose(got[mid + (0,)], G * uc[mid + (0,)], rtol=1e-12, atol=1e-15)
    assert np.abs(got[mid + (1,)]).max() < 1e-15 and np.abs(got[mid + (2,)][1:]).max() < 1e-15
