"""The reference's on-disk formats for the fractional-step path.

  grid.dat    text: the block count, then per block "IM JM KM" and the node
              coordinates, 3-D layout: every x (k, j, i order, i fastest),
              then every y, then every z (init.c:339-343, 606-720); 1-D
              layout (-grid1d): IM lines "x * *", JM lines "* y *", KM lines
              "* * z" (init.c:663-686).  Coordinates are multiplied by
              L_dim/cl on input (init.c:679, 728).
  bcs.dat     text: six boundary types per block, i-, i+, j-, j+, k-, k+
              (init.c:478-492).
  control.dat PETSc options, one "-name value" per line, "#" starts a
              comment; the path reads -ren, -dt (init.c:52-58), -mg_level
              (init.c:316-321), -mg_{i,j,k}_semi (init.c:352-354), -cl, and
              the Poisson KSP options -ps_ksp_{rtol,atol,max_it,gmres_restart}
              (poisson.c:1828-1940, PETSc defaults otherwise).
  fields      PETSc binary Vec files, one per field and time step (main.c:
              75-467): vfield%05d_%1d.dat (Ucont), ufield (Ucat), pfield
              (P), nvfield (Nvert).  A PETSc binary Vec is the 32-bit
              big-endian class id 1211214, the 32-bit big-endian global
              length, then the values as big-endian doubles in the DMDA
              natural order (dof-3 fields interleaved); a dof-3 field has a
              companion .info file "-vecload_block_size 3".

  nlist/elist the immersed body's triangulated surface in Icem format
              (ibm_read_Icem, ibm_io.c:3040-3240): nlist%02d holds the
              vertex count, then one "i i i x y z" line per vertex (the three
              integers unused; coordinates divided by -char_length_ibm, then
              offset by -x_c -y_c -z_c and scaled by L_dim); elist%02d holds
              the element count, then one line of 15 integers per triangle,
              whose 12th-14th are its 1-based vertex numbers.  Element
              normals are (p2 - p1) x (p3 - p1), normalised.

The field helpers exchange data with a curvib_old_amd.solver.UserCtx, in
the layout its get/put use.
"""
from __future__ import annotations

import os

import numpy as np

PETSC_VEC_CLASSID = 1211214


# ---------------------------------------------------------------- grid.dat
def read_grid_dat(path, grid1d=False, scale=1.0):
    """Blocks of a grid.dat file: [(IM, JM, KM, xyz)], xyz of shape
    (KM, JM, IM, 3) in node order (k, j, i)."""
    with open(path) as f:
        toks = f.read().split()
    pos = 0

    def take(n, conv=float):
        nonlocal pos
        out = toks[pos:pos + n]
        if len(out) < n:
            raise ValueError(f"{path}: truncated")
        pos += n
        return [conv(t) for t in out]

    nblk = take(1, int)[0]
    blocks = []
    for _ in range(nblk):
        IM, JM, KM = take(3, int)
        if grid1d:
            xi = np.array(take(3 * IM)).reshape(IM, 3)[:, 0]
            yj = np.array(take(3 * JM)).reshape(JM, 3)[:, 1]
            zk = np.array(take(3 * KM)).reshape(KM, 3)[:, 2]
            Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
            xyz = np.stack([X, Y, Z], axis=-1)
        else:
            n = IM * JM * KM
            xyz = np.empty((KM, JM, IM, 3))
            for c in range(3):
                xyz[..., c] = np.array(take(n)).reshape(KM, JM, IM)
        blocks.append((IM, JM, KM, xyz * scale))
    return blocks


def write_grid_dat(path, blocks):
    """Write [(xyz (KM, JM, IM, 3))] in the 3-D grid.dat layout."""
    with open(path, "w") as f:
        f.write(f"{len(blocks)}\n")
        for xyz in blocks:
            KM, JM, IM, _ = xyz.shape
            f.write(f"{IM} {JM} {KM}\n")
            for c in range(3):
                np.savetxt(f, xyz[..., c].reshape(-1, IM), fmt="%.17e")


# ----------------------------------------------------------------- bcs.dat
def read_bcs_dat(path, nblocks=1):
    with open(path) as f:
        toks = f.read().split()
    if len(toks) < 6 * nblocks:
        raise ValueError(f"{path}: expected {6 * nblocks} boundary types")
    return [tuple(int(t) for t in toks[6 * b:6 * b + 6]) for b in range(nblocks)]


def write_bcs_dat(path, bctypes):
    with open(path, "w") as f:
        for bc in bctypes:
            f.write(" ".join(str(int(v)) for v in bc) + "\n")


# ------------------------------------------------------------- control.dat
def read_control_dat(path):
    """Options of a control.dat file: {name: value string, or True for a
    flag}; later lines win, as in PETSc's options database."""
    opts = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            toks = line.split()
            q = 0
            while q < len(toks):
                t = toks[q]
                if not t.startswith("-"):
                    q += 1
                    continue
                name = t[1:]
                if q + 1 < len(toks) and not (toks[q + 1].startswith("-") and not _is_number(toks[q + 1])):
                    opts[name] = toks[q + 1]
                    q += 2
                else:
                    opts[name] = True
                    q += 1
    return opts


def _is_number(t):
    try:
        float(t)
        return True
    except ValueError:
        return False


def run_options(opts):
    """Grid-level options with the reference's defaults (init.c:52-62,
    316-354; main.c:23-24): Re 3000, dt 0.01, St 1, 3 MG levels, no
    semi-coarsening, inlet profile 2, inlet centre offsets 0."""
    return {
        "ren": float(opts.get("ren", 3000.0)),
        "dt": float(opts.get("dt", 0.01)),
        "st": 1.0,
        "mglevels": int(opts.get("mg_level", 3)),
        "semi": tuple(int(opts.get(f"mg_{a}_semi", 0)) for a in ("i", "j", "k")),
        "cl": float(opts.get("cl", 1.0)),
        "inlet": int(opts.get("inlet", 2)),
        "inlet_center": tuple(float(opts.get(f"{a}_c", 0.0)) for a in ("x", "y", "z")),
    }


def solver_options(opts):
    """Poisson KSP options (-ps_ prefix, poisson.c:1828-1940) as
    UserCtx.set_solver keywords; PETSc's KSP defaults where unset."""
    return {
        "rtol": float(opts.get("ps_ksp_rtol", 1e-5)),
        "atol": float(opts.get("ps_ksp_atol", 1e-50)),
        "max_it": int(opts.get("ps_ksp_max_it", 10000)),
        "restart": int(opts.get("ps_ksp_gmres_restart", 30)),
    }


def load_case_dir(directory, device=0, control="control.dat", nvert=None, **ctx_kw):
    """A UserCtx set up from a reference input deck: grid.dat (3-D layout),
    bcs.dat and control.dat in ``directory``.  nvert: the immersed-body
    classification (the reference's IB search output), zero if None."""
    from .solver import UserCtx
    opts = read_control_dat(os.path.join(directory, control)) if os.path.exists(
        os.path.join(directory, control)) else {}
    ro = run_options(opts)
    (IM, JM, KM, xyz), = read_grid_dat(os.path.join(directory, "grid.dat"), grid1d=bool(int(opts.get("grid1d", 0))),
                                      scale=1.0 / ro["cl"])
    bc = read_bcs_dat(os.path.join(directory, "bcs.dat"))[0]
    u = UserCtx((IM, JM, KM), bc, ro["mglevels"], ro["ren"], ro["dt"], st=ro["st"], semi=ro["semi"],
                device=device, inlet=ro["inlet"], inlet_center=ro["inlet_center"], **ctx_kw)
    u.set_solver(**solver_options(opts))
    u.set_grid(xyz)
    u.set_nvert(np.zeros(u._shape("nvert", -1)) if nvert is None else nvert)
    return u, opts


# ---------------------------------------------------------- PETSc binary Vec
def write_petsc_vec(path, a, block_size=1):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    with open(path, "wb") as f:
        f.write(np.array([PETSC_VEC_CLASSID, a.size], dtype=">i4").tobytes())
        f.write(a.astype(">f8").tobytes())
    info = path + ".info"
    if block_size > 1:
        with open(info, "w") as f:
            f.write(f"-vecload_block_size {block_size}\n")
    elif os.path.exists(info):
        os.unlink(info)


def read_petsc_vec(path):
    with open(path, "rb") as f:
        head = np.frombuffer(f.read(8), dtype=">i4")
        if head.size != 2 or head[0] != PETSC_VEC_CLASSID:
            raise ValueError(f"{path}: not a PETSc binary Vec")
        n = int(head[1])
        data = np.frombuffer(f.read(8 * n), dtype=">f8")
    if data.size != n:
        raise ValueError(f"{path}: truncated ({data.size} of {n} values)")
    return data.astype(np.float64)


# ---------------------------------------------------------------- fields
FIELD_FILES = {"ucont": "vfield", "ucat": "ufield", "p": "pfield", "nvert": "nvfield"}


def field_path(directory, field, ti, bi=0):
    return os.path.join(directory, f"{FIELD_FILES[field]}{ti:05d}_{bi:1d}.dat")


def write_fields(u, ti, bi=0, directory="."):
    """Ucont_P_Binary_Output (main.c:285-330): Nvert, Ucont, Ucat, P."""
    for name in ("nvert", "ucont", "ucat", "p"):
        a = u.get(name)
        write_petsc_vec(field_path(directory, name, ti, bi), a, 3 if a.ndim == 4 else 1)


def read_fields(u, ti, bi=0, directory="."):
    """Ucont_P_Binary_Input (main.c:75-110): Ucont and P into the context;
    returns the stored Nvert (the reference loads it into Nvert_o)."""
    for name in ("ucont", "p"):
        shp = u._shape(name, -1)
        a = read_petsc_vec(field_path(directory, name, ti, bi))
        if a.size != int(np.prod(shp)):
            raise ValueError(f"{name}: file holds {a.size} values, the grid needs {int(np.prod(shp))}")
        u.put(name, a.reshape(shp))
    return read_petsc_vec(field_path(directory, "nvert", ti, bi)).reshape(u._shape("nvert", -1))


# ------------------------------------------------------------ Icem surface
def write_icem(directory, vertices, triangles, ibi=0):
    """nlist%02d / elist%02d of a triangulated surface (0-based triangles)."""
    v = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    t = np.asarray(triangles, dtype=np.int64).reshape(-1, 3)
    with open(os.path.join(directory, f"nlist{ibi:02d}"), "w") as f:
        f.write(f"{len(v)}\n")
        for q, (x, y, z) in enumerate(v):
            f.write(f"{q + 1} 0 0 {x:.17e} {y:.17e} {z:.17e}\n")
    with open(os.path.join(directory, f"elist{ibi:02d}"), "w") as f:
        f.write(f"{len(t)}\n")
        for q, (a, b, c) in enumerate(t):
            f.write(f"{q + 1} 0 0 0 0 0 0 0 0 0 0 {a + 1} {b + 1} {c + 1} 0\n")


def read_icem(directory, ibi=0, cl=1.0, center=(0.0, 0.0, 0.0), L_dim=1.0):
    """The surface ibm_read_Icem reads (ibm_io.c:3076-3240): vertices (n_v, 3)
    after the -char_length_ibm, -x_c/-y_c/-z_c and L_dim transform, 0-based
    triangles (n_e, 3) and the unit element normals (n_e, 3)."""
    with open(os.path.join(directory, f"nlist{ibi:02d}")) as f:
        n_v = int(f.readline().split()[0])
        rows = [f.readline().split() for _ in range(n_v)]
    v = np.array([[float(r[3]), float(r[4]), float(r[5])] for r in rows])
    v = (v / cl + np.asarray(center, dtype=np.float64)) * L_dim
    with open(os.path.join(directory, f"elist{ibi:02d}")) as f:
        n_e = int(f.readline().split()[0])
        rows = [f.readline().split() for _ in range(n_e)]
    t = np.array([[int(r[11]) - 1, int(r[12]) - 1, int(r[13]) - 1] for r in rows], dtype=np.int64)
    return v, t, element_normals(v, t)


def element_normals(v, t):
    """Unit normals (p2 - p1) x (p3 - p1) of the triangles (ibm_io.c:3252-3263)."""
    p1, p2, p3 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    nf = np.cross(p2 - p1, p3 - p1)
    return nf / np.linalg.norm(nf, axis=1)[:, None]

