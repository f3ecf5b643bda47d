"""Serving layer on the GPU: continuous batching over the C ABI.

A case's result is the same alone and in the mix: each admitted case owns
its context and stream and every reduction has a fixed order, so the
tolerance is exact equality (0).  Admission respects the HBM budget, and a
case larger than the budget is refused at submit."""
import numpy as np
import pytest

from curvib_old_amd import _lib, cases, serving

pytestmark = pytest.mark.gpu

OPTS = dict(precision="mixed", smoother="yline", omega=0.9, pre_sweeps=3, post_sweeps=3, coarse_sweeps=10,
            restart=30, rtol=1e-9, max_it=200)


def small_cases(n, seed):
    rng = np.random.default_rng(seed)
    return [serving.sample_case(rng, lo=(16, 16, 16), hi=(48, 40, 40)) for _ in range(n)]


@pytest.fixture(scope="module")
def mix():
    cs = small_cases(6, 11)
    return [(c, cases.build(c)) for c in cs]


def run(reqs, slots, budget=50e9, nsteps=3, mode="streams", stagger=0.0, opts=OPTS):
    import time
    srv = serving.Server(budget, slots, mode=mode)
    try:
        ids = []
        for q, (c, a) in enumerate(reqs):
            if stagger and q:
                time.sleep(stagger)
            ids.append(srv.submit(c, a, nsteps, **opts))
        out = []
        for (c, _), cid in zip(reqs, ids):
            rc, info = srv.wait(cid)
            assert rc == 0, (c.cells, rc, _lib.last_error())
            mx, my, mz = c.storage
            out.append((srv.fetch(cid, "p", (mz, my, mx)), srv.fetch(cid, "ucont", (mz, my, mx, 3)), info))
        return out, srv.status()
    finally:
        srv.close()


def test_alone_equals_mix(mix):
    alone, _ = run(mix[:1], slots=1)
    together, st = run(mix, slots=4)
    p0, u0, i0 = alone[0]
    p1, u1, i1 = together[0]
    assert i0["steps_done"] == i1["steps_done"] == 3
    assert np.array_equal(p0, p1)
    assert np.array_equal(u0, u1)
    # the mix really overlapped: more than one case was resident at once
    assert st["bytes_peak"] > max(i["bytes"] for _, _, i in together)


def test_admission_counts_the_metrics_scratch(mix):
    """A case is admitted when its context and cv_set_geometry's transient
    scratch fit the budget together, and refused at submit one byte below."""
    c, a = mix[0]
    need = serving.admission_bytes(c, serving.levels_for(c.cells), OPTS["restart"])
    srv = serving.Server(budget_bytes=need - 1, slots=1)
    try:
        with pytest.raises(_lib.CvError) as e:
            srv.submit(c, a, 1, **OPTS)
        assert e.value.code == _lib.CV_ERR_OOM
    finally:
        srv.close()
    out, st = run([(c, a)], slots=1, budget=need, nsteps=1)
    assert out[0][2]["steps_done"] == 1 and st["bytes_peak"] == need


def _two_largest_budget(mix):
    """Room for the two largest contexts, and for any one case's admission
    (context plus metrics scratch)."""
    b = sorted(serving.ctx_bytes(c, serving.levels_for(c.cells), OPTS["restart"]) for c, _ in mix)
    adm = max(serving.admission_bytes(c, serving.levels_for(c.cells), OPTS["restart"]) for c, _ in mix)
    return max(b[-1] + b[-2], adm)


def test_budget_bounds_residency(mix):
    budget = _two_largest_budget(mix)
    out, st = run(mix, slots=6, budget=budget, nsteps=1)
    assert st["bytes_peak"] <= budget
    assert all(i["steps_done"] == 1 for _, _, i in out)


def test_oversize_case_refused(mix):
    c, a = mix[0]
    srv = serving.Server(budget_bytes=1 << 20, slots=1)
    try:
        with pytest.raises(_lib.CvError) as e:
            srv.submit(c, a, 1, **OPTS)
        assert e.value.code == _lib.CV_ERR_OOM
    finally:
        srv.close()


def test_ctx_bytes_matches_allocation(mix):
    from curvib_old_amd.solver import UserCtx
    c, (xyz, nv, uc) = mix[1]
    lev = serving.levels_for(c.cells)
    u = UserCtx(c.nodes, c.bctype, lev, c.re, c.dt, semi=serving.SEMI, coarsen_auto=serving.COARSEN_AUTO)
    u.set_solver(**OPTS)
    u.set_grid(xyz)
    u.set_nvert(nv)
    u.put("ucont", uc)
    u.Flow_Solver()
    assert u.device_bytes() == serving.ctx_bytes(c, lev, OPTS["restart"])
    u.close()


def test_batched_mode_equals_streams(mix):
    """The batched server (all admitted cases advance as one cv_batch, rebuilt
    as cases join and leave) gives every case the bits it gets in the
    streams server; staggered arrivals change the batch mid-run."""
    ref, _ = run(mix, slots=6, nsteps=6)
    got, st = run(mix, slots=4, mode="batched", nsteps=6, stagger=0.005)
    for (p0, u0, i0), (p1, u1, i1) in zip(ref, got):
        assert i1["steps_done"] == 6
        assert np.array_equal(p0, p1)
        assert np.array_equal(u0, u1)
    assert st["bytes_peak"] > max(i["bytes"] for _, _, i in got)


def test_batched_mode_budget(mix):
    budget = _two_largest_budget(mix)
    out, st = run(mix, slots=6, budget=budget, nsteps=1, mode="batched")
    assert st["bytes_peak"] <= budget
    assert all(i["steps_done"] == 1 for _, _, i in out)


def test_ib_cases_alone_equal_mix_in_both_modes():
    """Cases with a sphere carry their IB node records through
    cv_server_submit_ib: each ends with the bits cv_flow_step gives it alone
    with the same records, streams and batched alike."""
    from curvib_old_amd import ibsearch
    from curvib_old_amd.solver import from_case
    rng = np.random.default_rng(5)
    cs = []
    while len(cs) < 4:
        c = serving.sample_case(rng, lo=(24, 24, 24), hi=(48, 40, 40))
        if c.sphere is not None or len(cs) % 2:
            cs.append(c)
    reqs = []
    for c in cs:
        xyz, nv, uc = cases.build(c)
        ib = ibsearch.sphere_ib_nodes(c, xyz, nv) if c.sphere is not None else None
        reqs.append((c, (xyz, nv, uc, ib)))
    assert any(a[3] is not None for _, a in reqs)
    ref = []
    for c, (xyz, nv, uc, ib) in reqs:
        u, _ = from_case(c, mglevels=serving.levels_for(c.cells), semi=serving.SEMI,
                         coarsen_auto=serving.COARSEN_AUTO)
        try:
            u.set_solver(**OPTS)
            if ib is not None:
                u.set_ib_nodes(*ib)
            for _ in range(3):
                u.Flow_Solver()
            ref.append((u.get("p"), u.get("ucont")))
        finally:
            u.close()
    for mode in ("streams", "batched"):
        out, _ = run(reqs, slots=4, mode=mode)
        for q, ((p, u, info), (pr, ur)) in enumerate(zip(out, ref)):
            assert info["steps_done"] == 3
            assert np.array_equal(p, pr), (mode, q)
            assert np.array_equal(u, ur), (mode, q)


def test_mixed16_budget_and_batched_equals_streams(mix):
    """CV_MIXED16 allocates its FP16 packs (32 bytes per point and level) when
    the option is set; the server counts them in a case's bytes, and the
    batched server still gives each case the streams server's bits."""
    from curvib_old_amd.solver import UserCtx
    opts16 = {**OPTS, "precision": "mixed16"}
    c, (xyz, nv, uc) = mix[1]
    lev = serving.levels_for(c.cells)
    u = UserCtx(c.nodes, c.bctype, lev, c.re, c.dt, semi=serving.SEMI, coarsen_auto=serving.COARSEN_AUTO)
    try:
        u.set_grid(xyz)
        u.set_nvert(nv)
        u.put("ucont", uc)
        u.set_solver(**opts16)
        u.Flow_Solver()
        half = 32 * sum(int(np.prod(u.level_storage(q))) for q in range(lev))
        assert u.device_bytes() == serving.ctx_bytes(c, lev, OPTS["restart"]) + half
    finally:
        u.close()
    ref, _ = run(mix, slots=6, nsteps=3, opts=opts16)
    got, _ = run(mix, slots=4, mode="batched", nsteps=3, stagger=0.005, opts=opts16)
    for (cc, _), (p0, u0, i0), (p1, u1, i1) in zip(mix, ref, got):
        assert i1["steps_done"] == 3
        assert np.array_equal(p0, p1) and np.array_equal(u0, u1)
        lv = serving.levels_for(cc.cells)
        assert i0["bytes"] > serving.ctx_bytes(cc, lv, OPTS["restart"])


def _with_nan(arrays):
    xyz, nv, uc = arrays[:3]
    uc = uc.copy()
    mz, my, mx, _ = uc.shape
    uc[mz // 2, my // 2, mx // 2, 0] = np.nan
    return (xyz, nv, uc)


def test_nonfinite_step_is_reported(mix):
    """A NaN in the face fluxes reaches the Poisson right-hand side: the
    step returns CV_ERR_NONFINITE instead of a converged-looking status."""
    from curvib_old_amd.solver import UserCtx
    c, a = mix[0]
    xyz, nv, uc = _with_nan(a)
    u = UserCtx(c.nodes, c.bctype, serving.levels_for(c.cells), c.re, c.dt, semi=serving.SEMI,
                coarsen_auto=serving.COARSEN_AUTO)
    try:
        u.set_grid(xyz)
        u.set_nvert(nv)
        u.put("ucont", uc)
        u.set_solver(**OPTS)
        with pytest.raises(_lib.CvError) as e:
            u.Flow_Solver()
        assert e.value.code == _lib.CV_ERR_NONFINITE, e.value
    finally:
        u.close()


@pytest.mark.parametrize("mode", ["streams", "batched"])
def test_server_fails_only_the_nonfinite_case(mix, mode):
    """A case that goes non-finite fails with CV_ERR_NONFINITE; the cases
    beside it (in the batched server: in the same batch) finish with the
    bits they get alone."""
    reqs = [mix[0], (mix[1][0], _with_nan(mix[1][1])), mix[2]]
    ref, _ = run([mix[0], mix[2]], slots=2, nsteps=2)
    srv = serving.Server(50e9, 4, mode=mode)
    try:
        ids = [srv.submit(c, a, 2, **OPTS) for c, a in reqs]
        res = [srv.wait(cid) for cid in ids]
        assert res[1][0] == _lib.CV_ERR_NONFINITE and res[1][1]["status"] == 3, res[1]
        for q, r in ((0, 0), (2, 1)):
            assert res[q][0] == 0 and res[q][1]["steps_done"] == 2, (q, res[q])
            c = reqs[q][0]
            mx, my, mz = c.storage
            assert np.array_equal(srv.fetch(ids[q], "p", (mz, my, mx)), ref[r][0]), (mode, q)
    finally:
        srv.close()


def _hip_used_bytes():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value


def test_device_memory_stays_within_the_budget(mix):
    """hipMemGetInfo sampled every millisecond while the server admits,
    builds (metrics scratch included) and steps cases concurrently: the
    device memory beyond what was in use before stays below the budget plus
    a fixed allowance for the runtime's own state."""
    import threading
    import time
    run(mix[:1], slots=1, nsteps=1)                 # load the code objects first
    base = _hip_used_bytes()
    peak = [0]
    done = threading.Event()

    def sample():
        while not done.is_set():
            peak[0] = max(peak[0], _hip_used_bytes() - base)
            time.sleep(1e-3)

    budget = int(2.5 * max(serving.admission_bytes(c, serving.levels_for(c.cells), OPTS["restart"])
                           for c, _ in mix))
    th = threading.Thread(target=sample)
    th.start()
    try:
        out, st = run(mix, slots=6, budget=budget, nsteps=2)
    finally:
        done.set()
        th.join()
    allowance = 256 << 20
    print(f"device memory: peak {peak[0] / 2**20:.0f} MiB above the start, budget {budget / 2**20:.0f} MiB, "
          f"admitted peak {st['bytes_peak'] / 2**20:.0f} MiB", flush=True)
    assert st["bytes_peak"] <= budget
    assert peak[0] <= budget + allowance, (peak[0], budget)
