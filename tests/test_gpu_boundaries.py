"""Boundary faces beyond walls and periodic pairs on the GPU against the
oracle: inlet (bctype 5 on the k- face), outlet (4 on the k+ face), slip
(3) and far field (6).  InflowFlux (bcs.c:99-246) with each analytic inlet
profile, OutflowFlux (bcs.c:403-469), FormBCS with its far-field, slip and
outlet sections (bcs.c:4500-5010) and the Ubcs ghost values, then full
fractional steps with RK and ImpRK momentum, and a batch that mixes
inlet/outlet cases with a periodic channel.

Tolerances: single routines 1e-13 relative (the outlet sums are reduced in
another order than the oracle's sequential loops); fields after each
converged step (both sides at rtol 1e-13) 1e-10 relative L2, printed per
step; the batch against each case alone: exact equality."""
import numpy as np
import pytest

from curvib_old_amd import cases, serving
from curvib_old_amd.solver import Batch
from curvib_old_amd.solver import from_case as gpu_case
from oracle.pyoracle import from_case as oracle_case
from conftest import assert_fields_match

pytestmark = pytest.mark.gpu

INOUT = (7, 7, 1, 1, 5, 4)


def inout_case(inlet=7, center=(0.0, 0.5, 0.0), bctype=INOUT, **kw):
    base = dict(cells=(32, 24, 16), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.45), re=500.0, seed=31, dt=2e-4)
    base.update(kw)
    return cases.ChannelCase(bctype=bctype, inlet=inlet, inlet_center=center, **base)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# profile 7: channel parabola about y = 1/2; 10: the square duct put over
# yc in [1, 2]; 4: pipe of radius 1/2 about (y, z) = (1/2, 0); 1: uniform
@pytest.mark.parametrize("inlet,center", [(7, (0.0, 0.5, 0.0)), (10, (0.0, -1.0, 0.0)), (4, (0.0, 0.5, 0.0)),
                                          (1, (0.0, 0.0, 0.0))])
def test_inflow_and_formbcs_match_the_oracle(inlet, center):
    case = inout_case(inlet, center)
    u, _ = gpu_case(case, mglevels=3)
    o, _ = oracle_case(case, nlev=3)
    try:
        u.Contra2Cart()
        o.call("o_contra2cart")
        fin = u.InflowFlux()
        o.call("o_inflowflux")
        fo = o.flux_stats()
        assert fo["flux_in"] > 0.1
        assert abs(fin - fo["flux_in"]) <= 1e-13 * abs(fo["flux_in"])
        assert rel(u.get("ucont"), o.get("ucont")) < 1e-13
        u.FormBCS()
        o.call("o_formbcs")
        for f in ("ucont", "ucat"):
            e = rel(u.get(f), o.get(f))
            assert e < 1e-13, (f, e)
        # the k ghost planes carry 2 Ubcs - Ucat of the inlet and outlet
        ug, og = u.get("ucat"), o.get("ucat")
        for k in (0, -1):
            assert rel(ug[k], og[k]) < 1e-13, k
        fout = u.OutflowFlux()
        o.call("o_outflowflux")
        fo = o.flux_stats()
        assert abs(fout - fo["flux_out_diag"]) <= 1e-13 * abs(fo["flux_out_diag"])
    finally:
        u.close()
        o.close()


@pytest.mark.parametrize("scheme,bctype", [("rk", INOUT), ("imprk", INOUT), ("rk", (1, 1, 1, 1, 1, 4))])
def test_inlet_outlet_steps_track_the_oracle(scheme, bctype):
    """Full fractional steps (Flow_Solver, solvers.c:944-1156) through an
    inlet/outlet box, and a closed box with an outlet only (walls on the
    i faces: the per-face wall flux reset)."""
    case = inout_case(7, bctype=bctype)
    u, _ = gpu_case(case, mglevels=3)
    o, _ = oracle_case(case, nlev=3)
    if scheme == "imprk":
        imp = dict(imp_max_it=30, imp_atol=1e-7, imp_rtol=1e-4, cfl=0.3)
        u.set_momentum("imprk", **imp)
        o.set_momentum("imprk", **imp)
    u.set_solver(rtol=1e-13, max_it=400)
    o.set_solver(rtol=1e-13, max_it=400)
    try:
        for s in range(3 if scheme == "rk" else 2):
            st = u.Flow_Solver()
            o.call("o_flow_step")
            fo = o.flux_stats()
            if scheme == "imprk":
                so = o.imp_stats()
                assert st["imp_its"] == int(so["its"]) and st["imp_lambda"] == so["lambda"], (s, st, so)
            assert_fields_match(f"{scheme} {bctype} step {s}", u, o)
            # flux sums: relative to the inflow, or to the unit flux scale of
            # the outlet cells when the inflow is 0 (a wall at k-)
            scale = max(abs(fo["flux_in"]), 1.0)
            assert abs(st["flux_in"] - fo["flux_in"]) <= 1e-12 * scale, (s, st, fo)
            assert abs(st["flux_out"] - fo["flux_out"]) <= 1e-12 * scale, (s, st, fo)
            # the outlet correction balances the inflow
            assert abs(st["flux_out"] - st["flux_in"]) <= 1e-12 * scale, (s, st)
            assert st["maxdiv"] < 1e-6
    finally:
        u.close()
        o.close()


def test_batch_with_inlet_outlet_matches_alone():
    """A batch holding inlet/outlet cases runs the momentum per case around
    the batched pressure solve (the batched momentum handles periodic and
    wall faces only); every case, the periodic channel included, ends each
    step with the bits it gets alone."""
    solver = dict(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, coarse_sweeps=10, omega=0.9,
                  restart=30)
    cl = [inout_case(7, cells=(32, 32, 32), sphere=(0.5, 0.5, 0.4, 0.3)),
          cases.ChannelCase(cells=(32, 32, 32), amp=0.05, sphere=None, re=300.0, seed=3, dt=2e-4),
          inout_case(1, center=(0.0, 0.0, 0.0), cells=(48, 32, 32), sphere=None, dt=2e-4)]

    def ctx(c):
        u, _ = gpu_case(c, mglevels=serving.levels_for(c.cells), semi=serving.SEMI, coarsen_auto=serving.COARSEN_AUTO)
        u.set_solver(precision="mixed", **solver)
        return u

    inb = [ctx(c) for c in cl]
    alone = [ctx(c) for c in cl]
    b = Batch(inb)
    try:
        for s in range(3):
            sb = b.step()
            for q, u in enumerate(alone):
                sa = u.Flow_Solver()
                assert sb[q]["ksp_its"] == sa["ksp_its"] and sb[q]["maxdiv"] == sa["maxdiv"], (s, q)
                assert sb[q]["flux_in"] == sa["flux_in"] and sb[q]["flux_out"] == sa["flux_out"], (s, q)
            for q, (ua, ub) in enumerate(zip(inb, alone)):
                for f in ("p", "ucont", "ucat"):
                    assert np.array_equal(ua.get(f), ub.get(f)), (s, q, f)
    finally:
        b.close()
        for u in inb + alone:
            u.close()


# slip box (slip on the i faces and j+, k faces); far field on j+ and both k
# faces; slip and far field beside an inlet and an outlet
@pytest.mark.parametrize("bctype", [(3, 3, 1, 3, 3, 3), (7, 7, 1, 6, 6, 6), (3, 6, 1, 3, 5, 4)])
def test_slip_and_farfield_formbcs_match_the_oracle(bctype):
    case = inout_case(7, bctype=bctype)
    u, _ = gpu_case(case, mglevels=3)
    o, _ = oracle_case(case, nlev=3)
    try:
        u.Contra2Cart()
        o.call("o_contra2cart")
        u.InflowFlux()
        o.call("o_inflowflux")
        for rep in range(2):
            u.FormBCS()
            o.call("o_formbcs")
            for f in ("ucont", "ucat"):
                e = rel(u.get(f), o.get(f))
                assert e < 1e-13, (rep, f, e)
    finally:
        u.close()
        o.close()


@pytest.mark.parametrize("scheme,bctype", [("rk", (7, 7, 1, 3, 7, 7)), ("rk", (7, 7, 1, 6, 5, 4)),
                                           ("imprk", (3, 3, 1, 6, 6, 6))])
def test_slip_and_farfield_steps_track_the_oracle(scheme, bctype):
    """Fractional steps of an open channel (slip top), a far-field top over
    an inlet/outlet channel, and a slip/far-field box on ImpRK."""
    case = inout_case(7, bctype=bctype)
    u, _ = gpu_case(case, mglevels=3)
    o, _ = oracle_case(case, nlev=3)
    if scheme == "imprk":
        imp = dict(imp_max_it=30, imp_atol=1e-7, imp_rtol=1e-4, cfl=0.3)
        u.set_momentum("imprk", **imp)
        o.set_momentum("imprk", **imp)
    u.set_solver(rtol=1e-13, max_it=400)
    o.set_solver(rtol=1e-13, max_it=400)
    try:
        for s in range(3 if scheme == "rk" else 2):
            st = u.Flow_Solver()
            o.call("o_flow_step")
            if scheme == "imprk":
                so = o.imp_stats()
                assert st["imp_its"] == int(so["its"]) and st["imp_lambda"] == so["lambda"], (s, st, so)
            assert_fields_match(f"{scheme} {bctype} step {s}", u, o)
            assert st["maxdiv"] < 1e-6
    finally:
        u.close()
        o.close()

