"""Continuous-batching serving of independent channel cases on one GPU.

The server (``cv_server_*`` in the library) admits queued requests when a
slot is free and the case's device context fits the HBM budget beside the
cases in flight.  Each admitted case keeps its fields resident for all its
steps on its own HIP stream, so the cases in flight overlap on the GPU.
This module samples the ensemble and serving workloads, drives the server
with a Poisson arrival stream, and summarises per-step latency and
aggregate cell-updates/s.

Workloads:
  ensemble  n cases submitted at once, grids log-uniform from 32^3 to
            128x96x96 cells, wall amplitude U[0, 0.1], Re U[100, 2000],
            a sphere of diameter U[0.1, 0.3] in half of the cases (its IB
            nodes reconstructed when the inputs carry their records)
  serving   Poisson arrivals (rate lam per second) of cases with grids
            log-uniform from 32^3 to 192^3 cells, geometry and Re as above,
            nsteps fractional steps each, an HBM admission budget
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import multiprocessing as mp
import time

import numpy as np

from . import _lib, cases
from ._lib import check

SEMI = (0, 1, 0)


COARSEN_AUTO = 1


def levels_for(cells, semi=SEMI, max_levels=8, min_cells=4, coarsen_auto=COARSEN_AUTO):
    """Most MG levels the grid allows (at most ``max_levels``): every
    coarsening halves an even cell count and the coarsest level keeps at
    least ``min_cells`` cells on the coarsened axes."""
    from .solver import level_cells
    best = 1
    for lev in range(2, max_levels + 1):
        h = level_cells(cells, semi, lev, coarsen_auto)
        if h is None:
            break
        if any(h[0][a] < min_cells for a in range(3) if not semi[a]):
            break
        best = lev
    return best


def sample_case(rng, lo=(32, 32, 32), hi=(192, 192, 192), round_to=8):
    """One case of the ensemble/serving distributions."""
    t = rng.uniform()
    cells = tuple(int(round_to * max(1, round(l * (h / l) ** t / round_to))) for l, h in zip(lo, hi))
    amp = float(rng.uniform(0.0, 0.1))
    re = float(rng.uniform(100.0, 2000.0))
    sphere = None
    if rng.uniform() < 0.5:
        sphere = (0.5, 0.5, 0.5, float(rng.uniform(0.1, 0.3)))
    seed = int(rng.integers(0, 2 ** 31 - 1))
    c = cases.ChannelCase(cells=cells, amp=amp, sphere=sphere, re=re, seed=seed)
    return dataclasses.replace(c, dt=cases.stable_dt(c))


def _build_arrays(case):
    return cases.build(case)


def _build_arrays_ib(case):
    from . import ibsearch
    xyz, nv, uc = cases.build(case)
    ib = ibsearch.sphere_ib_nodes(case, xyz, nv) if case.sphere is not None else None
    return xyz, nv, uc, ib


def build_inputs(case_list, workers=8, ib=False):
    """Node coordinates, nvert and initial fluxes of every case, built in a
    spawn-started process pool (call before any GPU work in this process).
    With ib, a fourth entry per case holds the sphere's IB node records
    (ibsearch.sphere_ib_nodes; None without a sphere)."""
    fn = _build_arrays_ib if ib else _build_arrays
    if workers <= 1 or len(case_list) <= 1:
        return [fn(c) for c in case_list]
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        return pool.map(fn, case_list, chunksize=1)


def _ib_records(ib):
    """(idx, val) arrays as the cv_ib_node records the library takes."""
    idx, val = ib
    rec = np.zeros(len(idx), dtype=_lib.IB_NODE_DTYPE)
    rec["idx"] = np.asarray(idx).reshape(-1, 12)
    rec["val"] = np.asarray(val, dtype=np.float64).reshape(-1, 12)
    return rec


def grid_desc(case, mglevels, semi=SEMI, coarsen_auto=COARSEN_AUTO):
    d = _lib.GridDesc()
    d.coarsen_auto = int(coarsen_auto)
    d.IM, d.JM, d.KM = case.nodes
    for q in range(6):
        d.bctype[q] = int(case.bctype[q])
    d.mglevels = int(mglevels)
    d.isc, d.jsc, d.ksc = (int(v) for v in semi)
    d.ren, d.dt, d.st = float(case.re), float(case.dt), 1.0
    d.inletprofile = int(case.inlet)
    for q in range(3):
        d.inlet_center[q] = float(case.inlet_center[q])
    return d


def solver_opts(**kw):
    o = _lib.SolverOpts()
    _lib.load().cv_default_solver(C.byref(o))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = {"fp64": _lib.CV_FP64, "mixed": _lib.CV_MIXED, "mixed16": _lib.CV_MIXED16}[v]
        if k == "smoother" and isinstance(v, str):
            v = {"jacobi": _lib.CV_SMOOTH_JACOBI, "rb": _lib.CV_SMOOTH_RB, "yline": _lib.CV_SMOOTH_YLINE}[v]
        setattr(o, k, v)
    return o


def ctx_bytes(case, mglevels, restart=30, semi=SEMI):
    return _lib.load().cv_ctx_bytes(C.byref(grid_desc(case, mglevels, semi)), int(restart))


def geometry_scratch_bytes(case):
    """Device bytes cv_set_geometry holds while it builds the finest level's
    metrics (the server reserves them beside the context at admission):
    55 padded doubles per point (3 ghost layers per side) and the node
    coordinates."""
    IM, JM, KM = case.nodes
    return 55 * (IM + 7) * (JM + 7) * (KM + 7) * 8 + 3 * IM * JM * KM * 8


def admission_bytes(case, mglevels, restart=30, semi=SEMI):
    """The most device bytes a case holds at once: its context while the
    metrics scratch is held beside it."""
    return ctx_bytes(case, mglevels, restart, semi) + geometry_scratch_bytes(case)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Server:
    """ctypes handle of one cv_server."""

    def __init__(self, budget_bytes=200e9, slots=16, device=0, mode="streams"):
        L = _lib.load()
        h = C.c_void_p()
        m = {"streams": _lib.CV_SERVE_STREAMS, "batched": _lib.CV_SERVE_BATCHED}[mode]
        check(L.cv_server_create_mode(int(device), int(budget_bytes), int(slots), m, C.byref(h)),
              "cv_server_create_mode")
        self.h = h
        self._points = {}       # storage points of each submitted case

    def submit(self, case, arrays, nsteps, mglevels=None, semi=SEMI, **solver):
        """arrays: (xyz, nvert, ucont) or (xyz, nvert, ucont, ib) with ib the IB
        node records (idx, val) or None."""
        xyz, nv, uc = (np.ascontiguousarray(a, dtype=np.float64) for a in arrays[:3])
        ib = arrays[3] if len(arrays) > 3 else None
        IM, JM, KM = case.nodes
        ns = (IM + 1) * (JM + 1) * (KM + 1)
        # the library copies exactly these many doubles from each pointer
        for name, a, want in (("xyz", xyz, 3 * IM * JM * KM), ("nvert", nv, ns), ("ucont", uc, 3 * ns)):
            if a.size != want:
                raise ValueError(f"submit: {name} holds {a.size} values, the case needs {want}")
        if mglevels is None:
            mglevels = levels_for(case.cells, semi)
        d = grid_desc(case, mglevels, semi)
        o = solver_opts(**solver)
        cid = C.c_long()
        rec = _ib_records(ib) if ib is not None else np.zeros(0, dtype=_lib.IB_NODE_DTYPE)
        check(_lib.load().cv_server_submit_ib(self.h, C.byref(d), C.byref(o), _dp(xyz), _dp(nv), _dp(uc),
                                              int(nsteps), rec.ctypes.data_as(C.c_void_p), len(rec),
                                              _lib.CV_IB_NOSLIP, C.byref(cid)), "cv_server_submit_ib")
        self._points[cid.value] = ns
        return cid.value

    def wait(self, cid):
        info = _lib.CaseInfo()
        rc = _lib.load().cv_server_wait(self.h, int(cid), C.byref(info))
        return rc, {"status": info.status, "rc": info.rc, "steps_done": info.steps_done, "bytes": info.bytes,
                    "t_submit": info.t_submit, "t_admit": info.t_admit, "t_done": info.t_done,
                    "last": info.last.as_dict()}

    def step_ms(self, cid, nmax=100000):
        ms = np.zeros(nmax)
        its = np.zeros(nmax, dtype=np.int32)
        n = _lib.load().cv_server_step_ms(self.h, int(cid), _dp(ms), its.ctypes.data_as(C.POINTER(C.c_int)), nmax)
        if n < 0:
            check(n, "cv_server_step_ms")
        return ms[:n], its[:n]

    def fetch(self, cid, field, shape):
        """p or ucont of a finished case; shape must hold the case's storage
        points (mz, my, mx) for p, (mz, my, mx, 3) for ucont."""
        out = np.empty(shape)
        want = self._points[int(cid)] * (3 if field == "ucont" else 1)
        if out.size != want:
            raise ValueError(f"fetch: {field} of case {cid} holds {want} values, the buffer {out.size}")
        check(_lib.load().cv_server_fetch(self.h, int(cid), _lib.FIELD_ID[field], _dp(out)), "cv_server_fetch")
        return out

    def release(self, cid):
        check(_lib.load().cv_server_release(self.h, int(cid)), "cv_server_release")

    def status(self):
        s = _lib.ServerInfo()
        check(_lib.load().cv_server_status(self.h, C.byref(s)), "cv_server_status")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self):
        if getattr(self, "h", None):
            _lib.load().cv_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def latency_summary(ms):
    ms = np.asarray(ms, dtype=np.float64)
    if ms.size == 0:
        return {"n": 0}
    return {"n": int(ms.size), "p50_ms": float(np.percentile(ms, 50)), "p99_ms": float(np.percentile(ms, 99)),
            "mean_ms": float(ms.mean()), "max_ms": float(ms.max())}


def run_stream(requests, arrivals, nsteps, precision, budget_bytes=200e9, slots=16, device=0, log=None,
               mode="streams", **solver):
    """Submit ``requests`` [(case, arrays)] at the wall-clock offsets
    ``arrivals`` (seconds, increasing), wait for all, and return per-case
    records plus the aggregate figures.  Every case runs ``nsteps`` steps.
    mode: "streams" (a worker and stream per slot) or "batched" (all
    admitted cases advance together as one batch)."""
    srv = Server(budget_bytes, slots, device, mode)
    ids = []
    t0 = time.perf_counter()
    try:
        for (case, arrays), ta in zip(requests, arrivals):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append((case, srv.submit(case, arrays, nsteps, precision=precision, **solver)))
        recs = []
        for case, cid in ids:
            rc, info = srv.wait(cid)
            ms, its = srv.step_ms(cid)
            recs.append({"cells": int(np.prod(case.cells)), "rc": rc, "info": info, "step_ms": ms, "its": its})
            if log:
                log(f"[serve {precision}] case {cid} {case.cells} rc={rc} steps={len(ms)} "
                    f"p50={np.median(ms) if len(ms) else 0:.1f}ms its={int(np.median(its)) if len(its) else 0}")
        wall = time.perf_counter() - t0
        st = srv.status()
    finally:
        srv.close()
    all_ms = np.concatenate([r["step_ms"] for r in recs]) if recs else np.zeros(0)
    cell_updates = float(sum(r["cells"] * len(r["step_ms"]) for r in recs))
    first = min(r["info"]["t_submit"] for r in recs)
    last = max(r["info"]["t_done"] for r in recs)
    return {
        "precision": precision,
        "mode": mode,
        "cases": len(recs),
        "failed": sum(1 for r in recs if r["rc"] != 0),
        "steps": int(all_ms.size),
        "step_latency": latency_summary(all_ms),
        "queue_wait_s": latency_summary(1e3 * np.array([r["info"]["t_admit"] - r["info"]["t_submit"]
                                                        for r in recs])),
        "makespan_s": last - first,
        "wall_s": wall,
        "cell_updates_per_s": cell_updates / (last - first) if last > first else 0.0,
        "bytes_peak": st["bytes_peak"],
        "records": recs,
    }


def poisson_arrivals(n, lam, rng):
    return np.cumsum(rng.exponential(1.0 / lam, size=n))


def stated_workloads(n_serve=50, n_ens=64, seed=2026):
    """The serving stream and the ensemble of the benchmark, drawn from one
    seeded generator: ``n_serve`` cases with grids log-uniform in
    32^3..192^3 and Poisson arrival times at 50 cases/s, then ``n_ens``
    cases with grids in 32^3..128x96x96 (wall amplitude U[0, 0.1], Re
    U[100, 2000], a sphere of diameter U[0.1, 0.3] in half the cases).
    Returns (serve cases, arrival offsets in s, ensemble cases)."""
    rng = np.random.default_rng(seed)
    serve = [sample_case(rng, lo=(32, 32, 32), hi=(192, 192, 192)) for _ in range(n_serve)]
    arrivals = poisson_arrivals(n_serve, 50.0, rng)
    ens = [sample_case(rng, lo=(32, 32, 32), hi=(128, 96, 96)) for _ in range(n_ens)]
    return serve, arrivals, ens


def run_batched(requests, nsteps, precision, warmup=1, device=0, log=None, max_levels=8, **solver):
    """The ensemble as one cv_batch: every case [(case, arrays)] resident in
    its own context, all advanced together ``warmup`` untimed steps and then
    ``nsteps`` timed steps.  Every case finishes a step when the batch step
    returns, so each step's wall time is every case's step latency."""
    from .solver import Batch, UserCtx
    ctxs = []
    b = None
    try:
        for case, arrays in requests:
            xyz, nv, uc = arrays[:3]
            u = UserCtx(case.nodes, case.bctype, levels_for(case.cells, max_levels=max_levels), case.re, case.dt,
                        semi=SEMI, device=device, coarsen_auto=COARSEN_AUTO, inlet=case.inlet,
                        inlet_center=case.inlet_center)
            u.set_grid(xyz)
            u.set_nvert(nv)
            if len(arrays) > 3 and arrays[3] is not None:
                u.set_ib_nodes(*arrays[3])
            u.put("ucont", uc)
            u.set_solver(precision=precision, **solver)
            ctxs.append(u)
        b = Batch(ctxs)
        for _ in range(warmup):
            b.step()
        ms, its = [], []
        for s in range(nsteps):
            t0 = time.perf_counter()
            st = b.step()
            ms.append(1e3 * (time.perf_counter() - t0))
            its.append(max(x["ksp_its"] for x in st))
            if log:
                log(f"[batch {precision}] step {s}: {ms[-1]:.1f} ms, max its {its[-1]}")
        cells = float(sum(np.prod(c.cells) for c, _ in requests))
        wall = sum(ms) * 1e-3
        return {
            "precision": precision,
            "cases": len(requests),
            "steps": nsteps,
            "step_latency": latency_summary(ms),
            "max_its": its,
            "cell_updates_per_s": cells * nsteps / wall if wall > 0 else 0.0,
            "makespan_s": wall,
            "device_bytes": int(sum(u.device_bytes() for u in ctxs)),
        }
    finally:
        if b is not None:
            b.close()
        for u in ctxs:
            u.close()
