"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds the oracle's sources with
-fsanitize=address,undefined (no recovery) into a stand-alone driver.  Each
case runs whole fractional steps on a small immersed-sphere grid through it:
the driver must exit cleanly (a sanitizer report aborts it), and its fields
must agree with the optimised oracle library on the same inputs.

The reference's own undefined behaviour on this path (the stale pressure
derivative carry in FormFunction1, rhs.c:2803-2861, and the stale loop bound
in ImpRK, implicitsolver.c:7101,7198) is restated with defined values in the
oracle; these runs cover both routines."""
import os
import subprocess

import numpy as np
import pytest

from curvib_old_amd import cases, ibsearch
from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(HERE, "..", "oracle")
EXE = os.path.join(ORACLE, "_build", "cv_oracle_sanitize")

INOUT = (7, 7, 1, 1, 5, 4)
CASES = {
    "inlet_outlet_rk": (cases.ChannelCase(cells=(16, 12, 10), amp=0.1, sphere=(0.5, 0.5, 0.5, 0.4), seed=9, dt=2e-4,
                                          re=500.0, bctype=INOUT, inlet=7, inlet_center=(0.0, 0.5, 0.0)), "rk", 2),
    "channel_imprk": (cases.ChannelCase(cells=(12, 14, 10), amp=0.05, sphere=(0.5, 0.5, 0.5, 0.45), seed=3, dt=5e-4,
                                        re=300.0), "imprk", 2),
    "slip_farfield_rk": (cases.ChannelCase(cells=(10, 12, 12), amp=0.0, sphere=(0.5, 0.5, 0.5, 0.4), seed=4, dt=2e-4,
                                           re=200.0, bctype=(3, 3, 1, 6, 6, 6)), "rk", 2),
}
NLEV = 2
RTOL = 1e-11


@pytest.fixture(scope="module")
def exe():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    return EXE


def _write_input(path, case, scheme, nsteps, ib):
    xyz, nv, uc = cases.build(case)
    idx, val = ib if ib is not None else (np.zeros((0, 12), np.int32), np.zeros((0, 12)))
    idx = np.ascontiguousarray(idx, dtype="<i4").reshape(-1, 12)
    val = np.ascontiguousarray(val, dtype="<f8").reshape(-1, 12)
    hdr = np.array(list(case.nodes) + list(case.bctype) + [NLEV, case.inlet, {"rk": 0, "imprk": 1}[scheme], nsteps,
                                                           len(idx), 0, 0], dtype="<i4")
    dh = np.array([case.re, case.dt, *case.inlet_center, RTOL], dtype="<f8")
    with open(path, "wb") as f:
        for a in (hdr, dh, np.asarray(xyz, "<f8"), np.asarray(nv, "<f8"), np.asarray(uc, "<f8"), idx, val):
            f.write(np.ascontiguousarray(a).tobytes())
    return xyz, nv, uc, (idx, val)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", sorted(CASES))
def test_steps_run_clean_under_asan_and_ubsan(exe, tmp_path, name):
    case, scheme, nsteps = CASES[name]
    xyz0, nv0, _ = cases.build(case)
    ib = ibsearch.sphere_ib_nodes(case, xyz0, nv0)
    assert len(ib[0]) > 0
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    xyz, nv, uc, (idx, val) = _write_input(fin, case, scheme, nsteps, ib)
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, fin, fout], env=env, capture_output=True, text=True, timeout=280)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]

    mx, my, mz = case.storage
    ns = mx * my * mz
    out = np.fromfile(fout, dtype="<f8")
    assert out.size == 4 * ns + 2 * nsteps
    p_san, uc_san = out[:ns].reshape(mz, my, mx), out[ns:4 * ns].reshape(mz, my, mx, 3)
    stat = out[4 * ns:].reshape(nsteps, 2)

    # the block-Jacobi smoother takes one block per thread: same count as the driver
    threads = po.get_threads()
    po.set_threads(2)
    o = po.Oracle(case.nodes, case.bctype, NLEV, case.re, case.dt)
    try:
        o.set_inlet(case.inlet, case.inlet_center)
        o.set_grid(xyz)
        o.set_nvert(nv)
        o.set_ib_nodes(idx, val, 0)
        o.put("ucont", uc)
        o.set_solver(rtol=RTOL, max_it=200)
        if scheme == "imprk":
            o.set_momentum("imprk", imp_max_it=20, imp_atol=1e-7, imp_rtol=1e-4, cfl=0.2)
        for s in range(nsteps):
            o.call("o_flow_step")
            st = o.stats()
            assert stat[s, 0] == st["its"], (s, stat[s], st["its"])
            assert abs(stat[s, 1] - st["maxdiv"]) <= 1e-9 * max(abs(st["maxdiv"]), 1e-300), (s, stat[s], st["maxdiv"])
        p, u = o.get("p"), o.get("ucont")
    finally:
        o.close()
        po.set_threads(threads)
    for label, a, b in (("p", p_san, p), ("ucont", uc_san, u)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        print(f"{name} {label}: sanitized vs optimised oracle rel L2 {err:.2e}")
        assert err <= 1e-10, (label, err)
