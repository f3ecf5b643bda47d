"""Synthetic channel cases: node coordinates (grid.dat content), the nvert
body mask and the initial face fluxes.

These play the role of the reference's input files: ``grid.dat`` in its 3-D
format (init.c:339-343, 610-720), ``bcs.dat`` (init.c:478-492) and the
immersed-body classification that ``ibm_search_advanced_rev`` writes into
``Nvert`` (ibm.c:1074-1079: 0 fluid, 1.4 immersed-boundary node, 3.4 solid).

Geometry (by default a periodic channel, bcs.dat ``7 7 1 1 7 7``; other
bcs.dat faces -- slip 3, far field 6, a k- inlet 5 with an analytic -inlet
profile, a k+ outlet 4 -- keep the same grid and initial flux):
  x, z in [0, Lx] x [0, Lz] uniform; y tanh-stretched with parameter beta,
  y0 = (1 + tanh(beta*(2*eta-1))/tanh(beta))/2, eta = j/(JM-1); the lower
  wall is warped by A*sin(2*pi*x/Lx)*sin(2*pi*z/Lz), blended linearly to zero
  at the flat upper wall:  y = y0 + A*sin(..)*sin(..)*(1-y0).
Body: optional sphere (centre, diameter); a cell whose centre lies inside is
  solid (3.4), a non-solid cell with any of its 26 neighbours solid is an IB
  node (1.4).
Initial flux: Cartesian velocity u = (d1, d2, 1.5*(1-(2y-1)^2) + d3) at each
  face centre, d uniform in [-amp, amp] from a seeded Philox stream, dotted
  with the face area vector (csi / eta / zet of metrics.c:80-157).  Wall
  faces and faces touching a non-fluid cell carry zero flux; periodic face 0
  repeats face m-2.

All arrays are float64 in DMDA storage order [k][j][i] (i fastest), storage
extent m = nodes + 1 per axis.
"""
from __future__ import annotations

import dataclasses
import numpy as np

BCTYPE_CHANNEL = (7, 7, 1, 1, 7, 7)


@dataclasses.dataclass
class ChannelCase:
    cells: tuple            # (Cx, Cy, Cz)
    beta: float = 2.0
    amp: float = 0.05
    sphere: tuple | None = None   # (cx, cy, cz, diameter) or None
    re: float = 180.0
    dt: float = 1e-3
    seed: int = 0
    pert: float = 1e-2
    bctype: tuple = BCTYPE_CHANNEL
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    inlet: int = 2                          # -inlet profile of a bctype-5 k- face (main.c:23)
    inlet_center: tuple = (0.0, 0.0, 0.0)   # -x_c -y_c -z_c (main.c:924-926)

    @property
    def nodes(self):
        return tuple(c + 1 for c in self.cells)

    @property
    def storage(self):
        return tuple(c + 2 for c in self.cells)


def node_coords(case: ChannelCase) -> np.ndarray:
    """Node coordinates, shape (KM, JM, IM, 3), the content of grid.dat."""
    IM, JM, KM = case.nodes
    x = case.lx * np.arange(IM) / (IM - 1)
    z = case.lz * np.arange(KM) / (KM - 1)
    eta = np.arange(JM) / (JM - 1)
    b = case.beta
    if b > 0:
        y0 = 0.5 * (1.0 + np.tanh(b * (2 * eta - 1)) / np.tanh(b))
    else:
        y0 = eta.copy()
    y0 = case.ly * y0
    Z, Y0, X = np.meshgrid(z, y0, x, indexing="ij")
    warp = case.amp * np.sin(2 * np.pi * X / case.lx) * np.sin(2 * np.pi * Z / case.lz)
    Y = Y0 + warp * (case.ly - Y0) / case.ly
    out = np.empty((KM, JM, IM, 3))
    out[..., 0] = X
    out[..., 1] = Y
    out[..., 2] = Z
    return out


def _storage_coords(case: ChannelCase, xyz: np.ndarray) -> np.ndarray:
    """Node coordinates on the storage grid (index IM holds 0, init.c:666)."""
    mx, my, mz = case.storage
    c = np.zeros((mz, my, mx, 3))
    c[:-1, :-1, :-1] = xyz
    return c


def cell_centres(case: ChannelCase, xyz: np.ndarray) -> np.ndarray:
    """Cell centres (metrics.c:335-349) on storage indices 1..m-2."""
    c = xyz
    cent = 0.125 * (c[1:, 1:, 1:] + c[1:, :-1, 1:] + c[:-1, 1:, 1:] + c[:-1, :-1, 1:] +
                    c[1:, 1:, :-1] + c[1:, :-1, :-1] + c[:-1, 1:, :-1] + c[:-1, :-1, :-1])
    return cent   # shape (KM-1, JM-1, IM-1, 3) == interior cells


def nvert_field(case: ChannelCase, xyz: np.ndarray) -> np.ndarray:
    """nvert on the storage grid: 0 fluid, 1.4 IB node, 3.4 solid."""
    mx, my, mz = case.storage
    nv = np.zeros((mz, my, mx))
    if case.sphere is None:
        return nv
    cx, cy, cz, d = case.sphere
    cent = cell_centres(case, xyz)
    r2 = (cent[..., 0] - cx) ** 2 + (cent[..., 1] - cy) ** 2 + (cent[..., 2] - cz) ** 2
    inside = r2 < (0.5 * d) ** 2
    core = np.zeros((mz, my, mx), dtype=bool)
    core[1:-1, 1:-1, 1:-1] = inside
    near = np.zeros_like(core)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if dk == dj == di == 0:
                    continue
                near[1:-1, 1:-1, 1:-1] |= core[1 + dk:mz - 1 + dk, 1 + dj:my - 1 + dj, 1 + di:mx - 1 + di]
    nv[near & ~core] = 1.4
    nv[core] = 3.4
    _wrap_periodic_planes(case, nv)
    return nv


def _wrap_periodic_planes(case, a):
    """Ghost planes of periodic axes repeat the interior (plane 0 = m-2,
    plane m-1 = 1), the state UpdatePressure/FormBCS leave them in."""
    if case.bctype[0] == 7:
        a[:, :, 0] = a[:, :, -2]
        a[:, :, -1] = a[:, :, 1]
    if case.bctype[4] == 7:
        a[0] = a[-2]
        a[-1] = a[1]


def _face_vectors(c):
    """csi/eta/zet area vectors from storage node coordinates c (mz,my,mx,3),
    the cross products of metrics.c:80-157.  Returned on the storage grid
    (zeros where the reference does not compute them)."""
    mz, my, mx, _ = c.shape
    csi = np.zeros_like(c)
    eta = np.zeros_like(c)
    zet = np.zeros_like(c)
    # csi at k in [1,mz-1), j in [1,my-1), i in [0,mx-1)
    K, J = slice(1, mz - 1), slice(1, my - 1)
    I0 = slice(0, mx - 1)
    a = c[K, J, I0]; b = c[0:mz - 2, J, I0]; cc = c[K, 0:my - 2, I0]; d = c[0:mz - 2, 0:my - 2, I0]
    de = 0.5 * (a + b - cc - d)
    dz = 0.5 * (cc + a - d - b)
    csi[K, J, I0] = np.cross(de, dz)
    # eta at k in [1,mz-1), j in [1,my-1), i in [1,mx-1)
    I = slice(1, mx - 1)
    a = c[K, J, I]; b = c[0:mz - 2, J, I]; cc = c[K, J, 0:mx - 2]; d = c[0:mz - 2, J, 0:mx - 2]
    dc = 0.5 * (a + b - cc - d)
    dz = 0.5 * (a + cc - b - d)
    eta[K, J, I] = np.cross(dz, dc)
    # zet at k in [0,mz-1), j in [1,my-1), i in [1,mx-1)
    K0 = slice(0, mz - 1)
    a = c[K0, J, I]; b = c[K0, 0:my - 2, I]; cc = c[K0, J, 0:mx - 2]; d = c[K0, 0:my - 2, 0:mx - 2]
    dc = 0.5 * (a + b - cc - d)
    de = 0.5 * (a + cc - b - d)
    zet[K0, J, I] = np.cross(dc, de)
    return csi, eta, zet


def initial_ucont(case: ChannelCase, xyz: np.ndarray, nvert: np.ndarray) -> np.ndarray:
    """Initial contravariant face fluxes on the storage grid, shape
    (mz, my, mx, 3)."""
    mx, my, mz = case.storage
    c = _storage_coords(case, xyz)
    csi, eta, zet = _face_vectors(c)
    rng = np.random.Generator(np.random.Philox(case.seed))
    uc = np.zeros((mz, my, mx, 3))

    def vel(p, d):
        y = p[..., 1] / case.ly
        u = np.zeros_like(p)
        u[..., 2] = 1.5 * (1.0 - (2.0 * y - 1.0) ** 2)
        u += d
        return u

    K, J, I = slice(1, mz - 1), slice(1, my - 1), slice(1, mx - 1)
    # i faces (face centre = average of its 4 nodes), i in [0, mx-2]
    I0 = slice(0, mx - 1)
    fc = 0.25 * (c[K, J, I0] + c[0:mz - 2, J, I0] + c[K, 0:my - 2, I0] + c[0:mz - 2, 0:my - 2, I0])
    d = rng.uniform(-case.pert, case.pert, size=fc.shape)
    uc[K, J, I0, 0] = np.einsum("...c,...c->...", vel(fc, d), csi[K, J, I0])
    # j faces, j in [0, my-2]
    J0 = slice(0, my - 1)
    fc = 0.25 * (c[K, J0, I] + c[0:mz - 2, J0, I] + c[K, J0, 0:mx - 2] + c[0:mz - 2, J0, 0:mx - 2])
    d = rng.uniform(-case.pert, case.pert, size=fc.shape)
    uc[K, J0, I, 1] = np.einsum("...c,...c->...", vel(fc, d), eta[K, J0, I])
    # k faces, k in [0, mz-2]
    K0 = slice(0, mz - 1)
    fc = 0.25 * (c[K0, J, I] + c[K0, 0:my - 2, I] + c[K0, J, 0:mx - 2] + c[K0, 0:my - 2, 0:mx - 2])
    d = rng.uniform(-case.pert, case.pert, size=fc.shape)
    uc[K0, J, I, 2] = np.einsum("...c,...c->...", vel(fc, d), zet[K0, J, I])

    # walls: zero normal flux on bctype-1 faces
    bc = case.bctype
    if bc[0] == 1:
        uc[:, :, 0, 0] = 0.0
    if bc[1] == 1:
        uc[:, :, mx - 2, 0] = 0.0
    if bc[2] == 1:
        uc[:, 0, :, 1] = 0.0
    if bc[3] == 1:
        uc[:, my - 2, :, 1] = 0.0
    if bc[4] == 1:
        uc[0, :, :, 2] = 0.0
    if bc[5] == 1:
        uc[mz - 2, :, :, 2] = 0.0
    # faces touching a non-fluid cell carry no flux
    solid = nvert > 0.1
    uc[:, :, :-1, 0][solid[:, :, :-1] | solid[:, :, 1:]] = 0.0
    uc[:, :-1, :, 1][solid[:, :-1, :] | solid[:, 1:, :]] = 0.0
    uc[:-1, :, :, 2][solid[:-1] | solid[1:]] = 0.0
    # periodic: face 0 repeats face m-2; storage plane m-1 unused
    if bc[0] == 7:
        uc[:, :, 0, 0] = uc[:, :, mx - 2, 0]
    if bc[4] == 7:
        uc[0, :, :, 2] = uc[mz - 2, :, :, 2]
    # tangential fluxes on ghost planes are unused by the reference; keep 0
    uc[:, :, 0, 1:] = 0.0
    uc[:, :, mx - 1, :] = 0.0
    uc[:, 0, :, 0::2] = 0.0
    uc[:, my - 1, :, :] = 0.0
    uc[0, :, :, :2] = 0.0
    uc[mz - 1, :, :, :] = 0.0
    return uc


def stable_dt(case: ChannelCase, safety: float = 0.4) -> float:
    """Largest dt the explicit 4-stage RK momentum step tolerates on the
    case's grid, the viscous limit dt <= safety * dy_min^2 * Re (dy_min: the
    first cell of the tanh-stretched wall-normal spacing), capped at the
    case's own dt.  The reference's RK driver (implicitsolver.c:32-119) has
    no such check."""
    JM = case.nodes[1]
    eta = np.arange(JM) / (JM - 1)
    b = case.beta
    y = 0.5 * (1.0 + np.tanh(b * (2 * eta - 1)) / np.tanh(b)) if b > 0 else eta
    dy = float(np.min(np.diff(y))) * case.ly
    lim = safety * dy * dy * case.re
    e = np.floor(np.log10(lim))
    lim = float(f"{int(np.floor(lim / 10 ** e))}e{int(e)}")   # one significant digit, rounded down
    return float(min(case.dt, lim))


def build(case: ChannelCase):
    """Return (xyz_nodes (KM,JM,IM,3), nvert (mz,my,mx), ucont (mz,my,mx,3))."""
    xyz = node_coords(case)
    nv = nvert_field(case, xyz)
    uc = initial_ucont(case, xyz, nv)
    return xyz, nv, uc


def config_case(n: int, **over) -> ChannelCase:
    """The five benchmark configurations' geometry (sizes in cells)."""
    if n == 0:
        c = ChannelCase(cells=(64, 64, 64), amp=0.05, sphere=None, re=180.0)
    elif n == 1:
        c = ChannelCase(cells=(256, 192, 192), amp=0.05, sphere=(0.5, 0.5, 0.5, 0.25), re=500.0)
    elif n == 2:
        c = ChannelCase(cells=(512, 256, 256), amp=0.1, sphere=(0.5, 0.5, 0.5, 0.25), re=500.0)
    else:
        raise ValueError("configs 0..2 are single cases; ensembles come from curvib_old_amd.serving")
    return dataclasses.replace(c, **over)
