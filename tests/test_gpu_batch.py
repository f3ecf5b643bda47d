"""Batched ensemble (cv_batch_*): every case of a batch of uneven grids --
different multigrid depths and y-line lengths -- ends each fractional step
with the same bits in every field, and the same FGMRES iteration count, as
the same case stepped alone by cv_flow_step (Flow_Solver, solvers.c:944-1156),
in mixed (FP32 and FP16 coefficients) and FP64 precision, with the captured
hipGraph and with direct launches, with the momentum batched and per case,
and with stencil tiles of another row width than the case alone uses.
Tolerance: exact equality."""
import numpy as np
import pytest

from curvib_old_amd import _lib, cases, serving
from curvib_old_amd.solver import Batch, from_case

pytestmark = pytest.mark.gpu

SOLVER = dict(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, coarse_sweeps=10, omega=0.9,
              restart=30)


def _cases():
    rng = np.random.default_rng(77)
    cl = [serving.sample_case(rng, lo=(24, 24, 24), hi=(72, 64, 64)) for _ in range(4)]
    # a long y-line (160 rows: the 16-row chunks of the line solve) beside the short ones
    c = cases.ChannelCase(cells=(32, 160, 24), amp=0.08, sphere=(0.5, 0.5, 0.5, 0.2), re=700.0, seed=5)
    cl.append(cases.ChannelCase(**{**c.__dict__, "dt": cases.stable_dt(c)}))
    return cl


def _ctx(case, precision):
    u, _ = from_case(case, mglevels=serving.levels_for(case.cells), semi=serving.SEMI,
                     coarsen_auto=serving.COARSEN_AUTO)
    u.set_solver(precision=precision, **SOLVER)
    return u


# mom "1": the momentum and the rest of the step batched on the batch stream;
# "0": per case on each case's stream (graph "1": replayed from per-case graphs)
# tw: lanes per stencil tile row in the batch (each case alone uses 32); the
# tile shape must not change any bit
# legs: first depth of the V-cycle run as persistent per-depth launches
# ("off": one launch per operation at every depth); kc: k planes per stencil item
@pytest.mark.parametrize("precision,graph,mom,tw,legs,kc", [
    ("mixed", "1", "1", "32", "0", "4"), ("fp64", "1", "1", "32", "0", "4"),
    ("mixed", "1", "0", "32", "0", "4"), ("mixed", "0", "0", "32", "0", "4"),
    ("mixed16", "1", "1", "32", "0", "4"), ("mixed16", "1", "1", "64", "off", "4"),
    ("fp64", "1", "1", "64", "off", "4"), ("mixed16", "1", "1", "32", "2", "1"),
    ("mixed16", "0", "1", "32", "1", "8"), ("fp64", "0", "1", "32", "3", "2")])
def test_batch_matches_each_case_alone(monkeypatch, precision, graph, mom, tw, legs, kc):
    monkeypatch.setenv("CV_BATCH_GRAPH", graph)
    monkeypatch.setenv("CV_BATCH_MOM", mom)
    monkeypatch.setenv("CV_BATCH_TW", tw)
    monkeypatch.setenv("CV_BATCH_LEGS", legs)
    monkeypatch.setenv("CV_BATCH_LEG_KC", kc)
    print(f"batch vs alone: {precision} graph {graph} mom {mom} tw {tw} legs {legs} kc {kc}", flush=True)
    cl = _cases()
    levels = {serving.levels_for(c.cells) for c in cl}
    assert len(levels) > 1, "the batch must mix multigrid depths"
    inb = [_ctx(c, precision) for c in cl]
    alone = [_ctx(c, precision) for c in cl]
    b = Batch(inb)
    try:
        # three steps: the per-case graphs of the work around the solve are
        # captured on the first eligible step and replayed after it
        for s in range(3):
            sb = b.step()
            for q, u in enumerate(alone):
                sa = u.Flow_Solver()
                assert sb[q]["ksp_its"] == sa["ksp_its"], (s, q, sb[q]["ksp_its"], sa["ksp_its"])
                assert sb[q]["ksp_rnorm"] == sa["ksp_rnorm"] and sb[q]["maxdiv"] == sa["maxdiv"], (s, q)
                assert sb[q]["converged"] == 1
            for q, (ua, ub) in enumerate(zip(inb, alone)):
                for f in ("p", "phi", "ucont", "ucat"):
                    assert np.array_equal(ua.get(f), ub.get(f)), (s, q, f)
    finally:
        b.close()
        for u in inb + alone:
            u.close()


def test_persistent_leg_wait_failure_is_reported(monkeypatch):
    """The persistent V-cycle legs bound every wait; a wait that runs past
    its limit makes every work-group leave and the step returns an error.
    The failure is forced from the host (a negative limit fails every wait at
    once), so every kernel still ends normally."""
    monkeypatch.setenv("CV_BATCH_LEGS", "0")
    monkeypatch.setenv("CV_BATCH_LEG_SPIN", "-1")
    cl = _cases()[:2]
    inb = [_ctx(c, "mixed16") for c in cl]
    b = Batch(inb)
    try:
        with pytest.raises(_lib.CvError) as e:
            b.step()
        assert e.value.code == _lib.CV_ERR_HIP and "wait limit" in str(e.value), str(e.value)
    finally:
        b.close()
        for u in inb:
            u.close()
    # a batch built without the forced failure steps normally afterwards
    monkeypatch.setenv("CV_BATCH_LEG_SPIN", str(1 << 22))
    inb = [_ctx(c, "mixed16") for c in cl]
    b = Batch(inb)
    try:
        assert all(x["converged"] == 1 for x in b.step())
    finally:
        b.close()
        for u in inb:
            u.close()


def test_batch_rejects_mixed_options():
    cl = _cases()[:2]
    a, b = _ctx(cl[0], "mixed"), _ctx(cl[1], "fp64")
    try:
        with pytest.raises(_lib.CvError) as e:
            Batch([a, b])
        assert e.value.code == _lib.CV_ERR_ARG
        b.set_solver(precision="mixed", **{**SOLVER, "smoother": "jacobi"})
        a.set_solver(precision="mixed", **{**SOLVER, "smoother": "jacobi"})
        with pytest.raises(_lib.CvError) as e:
            Batch([a, b])
        assert e.value.code == _lib.CV_ERR_ARG
        a.set_solver(precision="mixed", **SOLVER)
        b.set_solver(precision="mixed", **SOLVER)
        bt = Batch([a, b])
        a.set_solver(precision="mixed", **{**SOLVER, "rtol": 1e-8})
        with pytest.raises(_lib.CvError) as e:
            bt.step()
        assert e.value.code == _lib.CV_ERR_STATE
        bt.close()
    finally:
        a.close()
        b.close()


def test_batch_with_imprk_matches_alone():
    """Cases on ImpRK (implicitsolver.c:7094-7256; pseudo iterations decided
    on the host) take the per-case path around the batched solve and still
    end each step with the bits they get alone."""
    cl = _cases()[:3]
    imp = dict(imp_max_it=30, imp_atol=1e-7, imp_rtol=1e-4, cfl=0.3)
    inb = [_ctx(c, "mixed") for c in cl]
    alone = [_ctx(c, "mixed") for c in cl]
    for u in inb + alone:
        u.set_momentum("imprk", **imp)
    b = Batch(inb)
    try:
        for s in range(2):
            sb = b.step()
            for q, u in enumerate(alone):
                sa = u.Flow_Solver()
                assert sb[q]["ksp_its"] == sa["ksp_its"] and sb[q]["imp_its"] == sa["imp_its"], (s, q)
            for q, (ua, ub) in enumerate(zip(inb, alone)):
                for f in ("p", "ucont", "ucat"):
                    assert np.array_equal(ua.get(f), ub.get(f)), (s, q, f)
    finally:
        b.close()
        for u in inb + alone:
            u.close()
