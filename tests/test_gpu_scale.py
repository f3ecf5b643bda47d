"""The ensemble and serving workloads at their stated scale on the GPU.

- Ensemble: the 64 seeded cases of the benchmark (grids log-uniform in
  32^3..128x96x96, wall amplitude U[0, 0.1], Re U[100, 2000], a sphere with
  its IB nodes in half the cases) in one cv_batch, CV_MIXED16 and FP64,
  three steps.  Every case converges to rtol 1e-9 on its FP64 residual,
  its divergence sits at the residual scale and its fields stay finite;
  the largest case and three random cases end every step with the bits
  they get stepped alone.
- Serving: the benchmark's stream of 50 cases (grids 32^3..192^3, Poisson
  arrivals at 50 cases/s) through the streams server with a 200 GB HBM
  budget, 20 steps each: no case fails, the admitted bytes never exceed the
  budget, and a sampled case ends with the bits it gets alone.

The inputs are built in this process (no worker processes: the GPU is
already in use here)."""
import time

import numpy as np
import pytest

from curvib_old_amd import serving
from curvib_old_amd.solver import Batch, UserCtx

pytestmark = pytest.mark.gpu

SOLVER = dict(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, coarse_sweeps=10, omega=0.9,
              restart=30)


def _ctx(case, arrays, precision):
    xyz, nv, uc = arrays[:3]
    u = UserCtx(case.nodes, case.bctype, serving.levels_for(case.cells), case.re, case.dt, semi=serving.SEMI,
                coarsen_auto=serving.COARSEN_AUTO)
    u.set_grid(xyz)
    u.set_nvert(nv)
    if arrays[3] is not None:
        u.set_ib_nodes(*arrays[3])
    u.put("ucont", uc)
    u.set_solver(precision=precision, **SOLVER)
    return u


@pytest.fixture(scope="module")
def ensemble():
    _, _, ens = serving.stated_workloads()
    t0 = time.perf_counter()
    arrays = serving.build_inputs(ens, workers=1, ib=True)
    print(f"ensemble: {len(ens)} cases, {sum(np.prod(c.cells) for c in ens) / 1e6:.1f} M cells, inputs built in "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    return list(zip(ens, arrays))


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["mixed16", "fp64"])
def test_ensemble_of_64_in_one_batch(ensemble, precision):
    assert len(ensemble) == 64
    sizes = [int(np.prod(c.cells)) for c, _ in ensemble]
    rng = np.random.default_rng(7)
    picks = [int(np.argmax(sizes))] + [int(q) for q in rng.choice(len(ensemble), 3, replace=False)]
    ctxs = [_ctx(c, a, precision) for c, a in ensemble]
    alone = {q: _ctx(*ensemble[q], precision) for q in picks}
    b = Batch(ctxs)
    try:
        for s in range(3):
            t0 = time.perf_counter()
            st = b.step()
            ms = 1e3 * (time.perf_counter() - t0)
            for q, x in enumerate(st):
                assert x["converged"] == 1 and x["ksp_rnorm"] <= 1e-9 * x["ksp_bnorm"] * 1.0001, (s, q, x)
                assert np.isfinite(x["maxdiv"]) and x["maxdiv"] < 1e-6, (s, q, x["maxdiv"])
            for q, u in alone.items():
                sa = u.Flow_Solver()
                assert sa["ksp_its"] == st[q]["ksp_its"] and sa["maxdiv"] == st[q]["maxdiv"], (s, q)
                for f in ("p", "ucont", "ucat"):
                    assert np.array_equal(u.get(f), ctxs[q].get(f)), (s, q, f)
            print(f"ensemble {precision} step {s}: {ms:.0f} ms, its {min(x['ksp_its'] for x in st)}.."
                  f"{max(x['ksp_its'] for x in st)}, max maxdiv {max(x['maxdiv'] for x in st):.2e}", flush=True)
        for u in ctxs:
            for f in ("p", "ucont"):
                assert np.isfinite(u.get(f)).all()
    finally:
        b.close()
        for u in ctxs + list(alone.values()):
            u.close()


@pytest.mark.timeout(900)
def test_serving_stream_of_50_with_a_200gb_budget():
    serve, arrivals, _ = serving.stated_workloads()
    t0 = time.perf_counter()
    arrays = serving.build_inputs(serve, workers=1, ib=True)
    print(f"serving: {len(serve)} cases, inputs built in {time.perf_counter() - t0:.1f} s", flush=True)
    reqs = list(zip(serve, arrays))
    budget = 200e9
    srv = serving.Server(budget, 16, mode="streams")
    ids = []
    try:
        t0 = time.perf_counter()
        for (c, a), ta in zip(reqs, arrivals):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(c, a, 20, precision="mixed16", **SOLVER))
        infos = []
        for (c, _), cid in zip(reqs, ids):
            rc, info = srv.wait(cid)
            assert rc == 0 and info["steps_done"] == 20, (c.cells, rc, info)
            infos.append(info)
        st = srv.status()
        print(f"serving: {len(ids)} cases x 20 steps in {time.perf_counter() - t0:.1f} s, peak admitted "
              f"{st['bytes_peak'] / 1e9:.1f} GB of {budget / 1e9:.0f} GB", flush=True)
        assert st["bytes_peak"] <= budget
        q = int(np.random.default_rng(3).integers(len(reqs)))
        c, a = reqs[q]
        mx, my, mz = c.storage
        p_mix = srv.fetch(ids[q], "p", (mz, my, mx))
        u_mix = srv.fetch(ids[q], "ucont", (mz, my, mx, 3))
    finally:
        srv.close()
    u = _ctx(c, a, "mixed16")
    try:
        for _ in range(20):
            u.Flow_Solver()
        assert np.array_equal(u.get("p"), p_mix) and np.array_equal(u.get("ucont"), u_mix), c.cells
    finally:
        u.close()
