"""Host-side mirror of the reference's fractional-step call sites on top of
the gfx950 library.

``UserCtx`` stands in for the reference's per-block ``UserCtx``
(variables.h:155-304) plus ``UserMG`` (variables.h:306-319): it owns one
device context holding the multigrid hierarchy, metrics and flow fields.
Its methods carry the reference routine names and run them on the GPU:

    FormFunction1   rhs.c:2567          PoissonSolver_MG  poisson.c:1775
    RungeKutta      implicitsolver.c:32 UpdatePressure    poisson.c:2908
    FormBCS         bcs.c:4436          Projection        poisson.c:2281
    Contra2Cart     rhs.c:25            Divergence        main.c:469
    PoissonRHS      poisson.c:1509      Flow_Solver       solvers.c:944

Host arrays use the DMDA global layout: scalar (mz, my, mx), Cmpnts
(mz, my, mx, 3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class UserCtx:
    def __init__(self, nodes, bctype, mglevels, ren, dt, st=1.0, semi=(0, 0, 0), device=0, coarsen_auto=0,
                 inlet=2, inlet_center=(0.0, 0.0, 0.0)):
        L = _lib.load()
        d = _lib.GridDesc()
        d.IM, d.JM, d.KM = (int(v) for v in nodes)
        for q in range(6):
            d.bctype[q] = int(bctype[q])
        d.mglevels = int(mglevels)
        d.isc, d.jsc, d.ksc = (int(v) for v in semi)
        d.ren, d.dt, d.st = float(ren), float(dt), float(st)
        d.coarsen_auto = int(coarsen_auto)
        d.inletprofile = int(inlet)
        for q in range(3):
            d.inlet_center[q] = float(inlet_center[q])
        h = C.c_void_p()
        check(L.cv_ctx_create(C.byref(d), int(device), C.byref(h)), "cv_ctx_create")
        self.h = h
        self.nodes = tuple(int(v) for v in nodes)
        self.mglevels = mglevels
        self.semi = tuple(int(v) for v in semi)
        self.coarsen_auto = int(coarsen_auto)
        self.storage = tuple(v + 1 for v in self.nodes)
        self.opts = _lib.SolverOpts()
        L.cv_default_solver(C.byref(self.opts))

    # -- life cycle ----------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            _lib.load().cv_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_bytes(self):
        return _lib.load().cv_device_bytes(self.h)

    def set_solver(self, **kw):
        for k, v in kw.items():
            if k == "precision" and isinstance(v, str):
                v = {"fp64": _lib.CV_FP64, "mixed": _lib.CV_MIXED, "mixed16": _lib.CV_MIXED16}[v]
            if k == "smoother" and isinstance(v, str):
                v = {"jacobi": _lib.CV_SMOOTH_JACOBI, "rb": _lib.CV_SMOOTH_RB, "yline": _lib.CV_SMOOTH_YLINE}[v]
            setattr(self.opts, k, v)
        check(_lib.load().cv_set_solver(self.h, C.byref(self.opts)), "cv_set_solver")

    def set_momentum(self, scheme="rk", imp_max_it=50, imp_atol=1e-7, imp_rtol=1e-4, cfl=0.2):
        """Momentum scheme of Flow_Solver (solvers.c:1022-1040): "rk" explicit
        RungeKutta or "imprk" dual time stepping, with the -imp_MAX_IT,
        -imp_atol, -imp_rtol and -cfl options."""
        o = _lib.MomentumOpts()
        o.scheme = {"rk": _lib.CV_MOM_RK, "imprk": _lib.CV_MOM_IMPRK}[scheme]
        o.imp_max_it, o.imp_atol, o.imp_rtol, o.cfl = int(imp_max_it), float(imp_atol), float(imp_rtol), float(cfl)
        check(_lib.load().cv_set_momentum(self.h, C.byref(o)), "cv_set_momentum")

    # -- inputs (grid.dat, nvert, restart fields) -----------------------
    def set_grid(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        IM, JM, KM = self.nodes
        if a.size != IM * JM * KM * 3:
            raise ValueError("node coordinate array has the wrong size")
        check(_lib.load().cv_set_geometry(self.h, _dp(a)), "cv_set_geometry")

    def set_nvert(self, nv):
        a = np.ascontiguousarray(nv, dtype=np.float64)
        if a.size != np.prod(self.storage):
            raise ValueError("nvert array has the wrong size")
        check(_lib.load().cv_set_nvert(self.h, _dp(a)), "cv_set_nvert")

    def _shape(self, name, level):
        if level < 0:
            mx, my, mz = self.storage
        else:
            mx, my, mz = self.level_storage(level)
        return (mz, my, mx, 3) if name in _lib.VEC_FIELDS else (mz, my, mx)

    def level_storage(self, level):
        """Storage extents of MG level ``level`` (MGDACreate, init.c:78-99):
        a coarsened axis keeps (nodes + 1) // 2 nodes, a semi-coarsened one
        keeps its size."""
        cells = level_cells(tuple(n - 1 for n in self.nodes), self.semi, self.mglevels, self.coarsen_auto)[level]
        return tuple(c + 2 for c in cells)

    def get(self, name, level=-1):
        out = np.empty(self._shape(name, level))
        check(_lib.load().cv_get_field(self.h, _lib.FIELD_ID[name], level, _dp(out)), f"get {name}")
        return out

    def put(self, name, arr, level=-1):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self._shape(name, level):
            raise ValueError(f"{name}: expected shape {self._shape(name, level)}, got {a.shape}")
        check(_lib.load().cv_put_field(self.h, _lib.FIELD_ID[name], level, _dp(a)), f"put {name}")

    # -- reference routines ----------------------------------------------
    def Contra2Cart(self):
        check(_lib.load().cv_contra2cart(self.h), "Contra2Cart")

    def set_ib_nodes(self, idx, val, mode="noslip"):
        """IB node records (cv_set_ib_nodes): idx (n, 12) ints = i j k, i1 j1 k1,
        i2 j2 k2, i3 j3 k3 and val (n, 12) doubles = cr1 cr2 cr3, d_s, d_i,
        ua[3], dun, nf[3], e.g. from ibsearch.sphere_ib_nodes.  mode "noslip"
        or "freeslip" (the reference's -inv)."""
        idx = np.asarray(idx).reshape(-1, 12)
        rec = np.zeros(len(idx), dtype=_lib.IB_NODE_DTYPE)
        rec["idx"] = idx
        rec["val"] = np.asarray(val, dtype=np.float64).reshape(-1, 12)
        m = {"noslip": _lib.CV_IB_NOSLIP, "freeslip": _lib.CV_IB_FREESLIP}[mode]
        check(_lib.load().cv_set_ib_nodes(self.h, rec.ctypes.data_as(C.c_void_p), len(rec), m), "cv_set_ib_nodes")

    def ibm_interpolation(self):
        """ibm_interpolation_advanced(user, ibm, ibi, 1) (ibm.c:3012-3254)."""
        check(_lib.load().cv_ibm_interpolation(self.h), "ibm_interpolation_advanced")

    def InflowFlux(self):
        """Inlet fluxes and Ubcs of a bctype-5 k- face; returns FluxInSum."""
        out = C.c_double()
        check(_lib.load().cv_inflow_flux(self.h, C.byref(out)), "InflowFlux")
        return out.value

    def OutflowFlux(self):
        """Flux through a bctype-4 k+ face (diagnostic)."""
        out = C.c_double()
        check(_lib.load().cv_outflow_flux(self.h, C.byref(out)), "OutflowFlux")
        return out.value

    def FormBCS(self):
        check(_lib.load().cv_form_bcs(self.h), "FormBCS")

    def FormFunction1(self):
        out = np.empty(self._shape("rhs", -1))
        check(_lib.load().cv_form_function1(self.h, _dp(out)), "FormFunction1")
        return out

    def RungeKutta(self):
        check(_lib.load().cv_rungekutta(self.h), "RungeKutta")

    def ImpRK(self):
        st = _lib.StepStats()
        check(_lib.load().cv_imprk(self.h, C.byref(st)), "ImpRK")
        return st.as_dict()

    def CalcRHS(self):
        out = np.empty(self._shape("rhs", -1))
        check(_lib.load().cv_calcrhs(self.h, _dp(out)), "CalcRHS")
        return out

    def VolumeFlux_rev(self):
        out = np.zeros(4)
        check(_lib.load().cv_volumeflux_rev(self.h, _dp(out)), "VolumeFlux_rev")
        return out

    def PoissonRHS(self):
        out = np.empty(self._shape("b", -1))
        check(_lib.load().cv_poisson_rhs(self.h, _dp(out)), "PoissonRHS")
        return out

    def PoissonApply(self, x, level=-1):
        a = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(a)
        check(_lib.load().cv_poisson_apply(self.h, level, _dp(a), _dp(out)), "PoissonApply")
        return out

    def PoissonSolver_MG(self, allow_unconverged=False):
        st = _lib.StepStats()
        rc = _lib.load().cv_poisson_solve(self.h, C.byref(st))
        if not (rc == _lib.CV_OK or (allow_unconverged and rc == _lib.CV_ERR_NOT_CONVERGED)):
            check(rc, "PoissonSolver_MG")
        return st.as_dict()

    def UpdatePressure(self):
        check(_lib.load().cv_update_pressure(self.h), "UpdatePressure")

    def Projection(self):
        check(_lib.load().cv_projection(self.h), "Projection")

    def Divergence(self):
        v = C.c_double()
        i = C.c_long()
        check(_lib.load().cv_divergence(self.h, C.byref(v), C.byref(i)), "Divergence")
        return v.value, i.value

    def Flow_Solver(self, allow_unconverged=False):
        st = _lib.StepStats()
        rc = _lib.load().cv_flow_step(self.h, C.byref(st))
        if not (rc == _lib.CV_OK or (allow_unconverged and rc == _lib.CV_ERR_NOT_CONVERGED)):
            check(rc, "Flow_Solver")
        return st.as_dict()

    # -- instrumentation ---------------------------------------------------
    def prof_select(self, kernel, level=-1, precision=-1):
        """Time every launch of one kernel class (``_lib.CV_PROF`` name) on
        one level with device events on the context stream."""
        k = _lib.CV_PROF[kernel] if isinstance(kernel, str) else int(kernel)
        check(_lib.load().cv_prof_select(self.h, k, int(level), int(precision)), "cv_prof_select")

    def prof_read(self, reset=False):
        """(launches, device ms summed, algorithmic bytes summed)."""
        n = C.c_long()
        ms = C.c_double()
        by = C.c_double()
        check(_lib.load().cv_prof_read(self.h, C.byref(n), C.byref(ms), C.byref(by), int(reset)), "cv_prof_read")
        return n.value, ms.value, by.value

    # operator plug-in points (tests)
    def MyRestriction(self, coarse_level, xf):
        shp = self._shape("b", coarse_level)
        out = np.empty(shp)
        a = np.ascontiguousarray(xf, dtype=np.float64)
        check(_lib.load().cv_mg_restrict(self.h, coarse_level, _dp(a), _dp(out)), "MyRestriction")
        return out

    def MyInterpolation(self, fine_level, xc):
        shp = self._shape("b", fine_level)
        out = np.empty(shp)
        a = np.ascontiguousarray(xc, dtype=np.float64)
        check(_lib.load().cv_mg_prolong(self.h, fine_level, _dp(a), _dp(out)), "MyInterpolation")
        return out

    def PoissonNullSpaceFunction(self, x, level=-1):
        a = np.ascontiguousarray(x, dtype=np.float64).copy()
        check(_lib.load().cv_nullspace_project(self.h, level, _dp(a)), "PoissonNullSpaceFunction")
        return a


def level_cells(cells, semi, nlev, coarsen_auto=0):
    """Cell counts of every MG level, coarsest first (the hierarchy
    cv_ctx_create builds); None where a level cannot be formed."""
    c = list(cells)
    out = [tuple(c)]
    for _ in range(nlev - 1):
        allowed = [a for a in range(3) if not semi[a]]
        if not allowed:
            return None
        most = max(c[a] for a in allowed)
        for a in allowed:
            if coarsen_auto and c[a] < most:
                continue
            if c[a] % 2:
                return None
            c[a] //= 2
        out.append(tuple(c))
    if min(min(x) for x in out) < 3:
        return None
    return out[::-1]


def from_case(case, mglevels=3, device=0, semi=(0, 0, 0), coarsen_auto=0):
    """UserCtx loaded with a curvib_old_amd.cases.ChannelCase."""
    from . import cases
    xyz, nv, uc = cases.build(case)
    u = UserCtx(case.nodes, case.bctype, mglevels, case.re, case.dt, semi=semi, device=device,
                coarsen_auto=coarsen_auto, inlet=case.inlet, inlet_center=case.inlet_center)
    u.set_grid(xyz)
    u.set_nvert(nv)
    u.put("ucont", uc)
    return u, (xyz, nv, uc)


class Batch:
    """A batch of UserCtx cases advanced together (``cv_batch_*``): one
    kernel launch per multigrid depth and operation for the pressure solve
    of every case, the rest of the step per case on its own stream.  The
    cases must share solver options (y-line smoother) and stay open while
    the batch exists."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        arr = (C.c_void_p * len(self.ctxs))(*[u.h.value if isinstance(u.h, C.c_void_p) else u.h
                                              for u in self.ctxs])
        h = C.c_void_p()
        check(_lib.load().cv_batch_create(arr, len(self.ctxs), C.byref(h)), "cv_batch_create")
        self.h = h

    def _run(self, fn, name, allow_unconverged):
        st = (_lib.StepStats * len(self.ctxs))()
        rc = fn(self.h, st)
        if not (rc == _lib.CV_OK or (allow_unconverged and rc == _lib.CV_ERR_NOT_CONVERGED)):
            check(rc, name)
        return [s.as_dict() for s in st]

    def step(self, allow_unconverged=False):
        """One fractional step of every case; per-case statistics."""
        return self._run(_lib.load().cv_batch_step, "cv_batch_step", allow_unconverged)

    def poisson_solve(self, allow_unconverged=False):
        """VolumeFlux_rev, PoissonRHS and the pressure solve of every case."""
        return self._run(_lib.load().cv_batch_poisson_solve, "cv_batch_poisson_solve", allow_unconverged)

    def close(self):
        if getattr(self, "h", None):
            _lib.load().cv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
