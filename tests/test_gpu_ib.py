"""Immersed-boundary reconstruction on the GPU against the oracle:
ibm_interpolation_advanced (ibm.c:3012-3254) over the IB nodes of the
analytic sphere (curvib_old_amd.ibsearch), no-slip and free-slip, with a
moving-surface velocity and dU_n/dt term; then fractional steps with RK and
ImpRK momentum, where the reconstruction runs at the reference's call sites;
and a batch holding an IB case.

The GPU sweeps touch only the faces and cells next to IB nodes; the oracle
restates the reference's whole-grid loops and Contra2Cart, so equality also
checks that nothing outside those lists changes.

Tolerances: one call 1e-13 relative; fields after each converged step
(both sides at rtol 1e-13) 1e-10 relative L2, printed per step; the batch
against each case alone: exact equality."""
import numpy as np
import pytest

from curvib_old_amd import cases, ibsearch, serving
from curvib_old_amd.solver import Batch
from curvib_old_amd.solver import from_case as gpu_case
from oracle.pyoracle import from_case as oracle_case
from conftest import assert_fields_match

pytestmark = pytest.mark.gpu

SPHERE = cases.ChannelCase(cells=(32, 24, 16), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.45), re=500.0, seed=21,
                           dt=2e-4)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _nodes(case, moving=False):
    xyz, nv, _ = cases.build(case)
    idx, val = ibsearch.sphere_ib_nodes(case, xyz, nv)
    if moving:
        rng = np.random.default_rng(3)
        val[:, 5:8] = rng.uniform(-0.05, 0.05, (len(val), 3))
        val[:, 8] = rng.uniform(-1e-4, 1e-4, len(val))
    return idx, val


@pytest.mark.parametrize("mode,moving", [("noslip", False), ("noslip", True), ("freeslip", True)])
def test_ibm_interpolation_matches_the_oracle(mode, moving):
    u, _ = gpu_case(SPHERE, mglevels=3)
    o, _ = oracle_case(SPHERE, nlev=3)
    try:
        idx, val = _nodes(SPHERE, moving)
        assert len(idx) > 100
        u.set_ib_nodes(idx, val, mode)
        o.set_ib_nodes(idx, val, {"noslip": 0, "freeslip": 1}[mode])
        mx, my, mz = u.storage
        p = np.random.default_rng(5).uniform(-1, 1, (mz, my, mx))
        u.put("p", p)
        o.put("p", p)
        u.Contra2Cart()
        o.call("o_contra2cart")
        u.ibm_interpolation()
        o.call("o_ibm_interpolation")
        for f in ("ucat", "ucont", "p"):
            e = rel(u.get(f), o.get(f))
            assert e < 1e-13, (f, e)
        # the reconstruction moved the IB nodes
        nv = o.get("nvert")
        assert np.abs(u.get("ucat")[np.abs(nv - 1.4) < 1e-12]).max() > 1e-3
    finally:
        u.close()
        o.close()


@pytest.mark.parametrize("scheme", ["rk", "imprk"])
def test_ib_steps_track_the_oracle(scheme):
    u, _ = gpu_case(SPHERE, mglevels=3)
    o, _ = oracle_case(SPHERE, nlev=3)
    try:
        idx, val = _nodes(SPHERE)
        u.set_ib_nodes(idx, val)
        o.set_ib_nodes(idx, val, 0)
        if scheme == "imprk":
            imp = dict(imp_max_it=30, imp_atol=1e-7, imp_rtol=1e-4, cfl=0.3)
            u.set_momentum("imprk", **imp)
            o.set_momentum("imprk", **imp)
        u.set_solver(rtol=1e-13, max_it=400)
        o.set_solver(rtol=1e-13, max_it=400)
        for s in range(3 if scheme == "rk" else 2):
            st = u.Flow_Solver()
            o.call("o_flow_step")
            if scheme == "imprk":
                so = o.imp_stats()
                assert st["imp_its"] == int(so["its"]) and st["imp_lambda"] == so["lambda"], (s, st, so)
            assert_fields_match(f"ib {scheme} step {s}", u, o)
            assert st["maxdiv"] < 1e-6
    finally:
        u.close()
        o.close()


def test_set_nvert_drops_the_ib_nodes():
    """The IB face and cell lists follow nvert: a new nvert removes the
    nodes, and the reconstruction does nothing until they are set again."""
    u, (xyz, nv, uc) = gpu_case(SPHERE, mglevels=3)
    try:
        idx, val = _nodes(SPHERE)
        u.set_ib_nodes(idx, val)
        u.set_nvert(nv)
        u.Contra2Cart()
        before = u.get("ucat")
        u.ibm_interpolation()
        assert np.array_equal(u.get("ucat"), before)
    finally:
        u.close()


@pytest.mark.parametrize("mom", ["1", "0"])
def test_batch_with_ib_matches_alone(monkeypatch, mom):
    """mom "1": the IB case runs its reconstruction inside the batched
    momentum; "0": per case around the batched solve."""
    monkeypatch.setenv("CV_BATCH_MOM", mom)
    solver = dict(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, coarse_sweeps=10, omega=0.9,
                  restart=30)
    sph = cases.ChannelCase(cells=(32, 32, 32), amp=0.08, sphere=(0.5, 0.5, 0.5, 0.4), re=400.0, seed=4, dt=2e-4)
    cl = [sph, cases.ChannelCase(cells=(32, 32, 32), amp=0.05, sphere=None, re=300.0, seed=3, dt=2e-4)]

    def ctx(c):
        u, _ = gpu_case(c, mglevels=serving.levels_for(c.cells), semi=serving.SEMI, coarsen_auto=serving.COARSEN_AUTO)
        u.set_solver(precision="mixed", **solver)
        if c.sphere is not None:
            u.set_ib_nodes(*_nodes(c))
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
            for q, (ua, ub) in enumerate(zip(inb, alone)):
                for f in ("p", "ucont", "ucat"):
                    assert np.array_equal(ua.get(f), ub.get(f)), (s, q, f)
    finally:
        b.close()
        for u in inb + alone:
            u.close()


def test_batch_refuses_new_ib_nodes_after_create():
    """A batch's captured graphs hold each case's IB buffers: new IB nodes
    after cv_batch_create make the next batch call fail with CV_ERR_STATE
    instead of replaying stale pointers."""
    from curvib_old_amd import _lib
    sph = cases.ChannelCase(cells=(32, 32, 32), amp=0.08, sphere=(0.5, 0.5, 0.5, 0.4), re=400.0, seed=4, dt=2e-4)
    u, _ = gpu_case(sph, mglevels=serving.levels_for(sph.cells), semi=serving.SEMI, coarsen_auto=serving.COARSEN_AUTO)
    u.set_solver(precision="mixed", rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3,
                 coarse_sweeps=10, omega=0.9, restart=30)
    u.set_ib_nodes(*_nodes(sph))
    b = Batch([u])
    try:
        b.step()
        u.set_ib_nodes(*_nodes(sph))
        with pytest.raises(_lib.CvError) as e:
            b.step()
        assert e.value.code == _lib.CV_ERR_STATE
    finally:
        b.close()
        u.close()


def test_icem_surface_steps_track_the_oracle(tmp_path):
    """A triangulated sphere read from Icem files (nlist/elist) translating
    rigidly: nvert from the surface (inside/outside of the facets), its IB
    records (foot points on the facets, surface velocity and dU_n) drive two
    RK steps on the GPU and on the oracle."""
    from curvib_old_amd import formats
    vv, tt = ibsearch.icosphere((0.5, 0.5, 0.5), 0.225, 4)
    formats.write_icem(str(tmp_path), vv, tt)
    v, t, _ = formats.read_icem(str(tmp_path))
    xyz, _, _ = cases.build(SPHERE)
    nv = ibsearch.surface_nvert(SPHERE, xyz, v, t)
    U = np.tile([0.05, 0.0, -0.02], (len(v), 1))
    idx, val = ibsearch.surface_ib_nodes(SPHERE, xyz, nv, v, t, u=U, uold=0.5 * U)
    u, _ = gpu_case(SPHERE, mglevels=3)
    o, _ = oracle_case(SPHERE, nlev=3)
    try:
        u.set_nvert(nv)
        o.set_nvert(nv)
        u.set_ib_nodes(idx, val)
        o.set_ib_nodes(idx, val, 0)
        u.set_solver(rtol=1e-13, max_it=400)
        o.set_solver(rtol=1e-13, max_it=400)
        for s in range(2):
            u.Flow_Solver()
            o.call("o_flow_step")
            assert_fields_match(f"icem rk step {s}", u, o)
    finally:
        u.close()
        o.close()
