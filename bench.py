#!/usr/bin/env python3
"""Fractional-step benchmark of the gfx950 CURVIB core on one MI355X.

Workload: the warped periodic channel with an immersed sphere at
512x256x256 cells (33.5 M cells, 34.2 M DMDA storage points), wall
amplitude 0.1, sphere diameter 0.25, Re 500, tanh wall stretching beta 2,
seeded +-1e-2 perturbations (curvib_old_amd.cases.config_case(2)), with the
explicit-RK stable dt of that grid.  One step is the reference's
Flow_Solver (solvers.c:944-1156): 4-stage RK momentum with FormFunction1,
PoissonRHS, the Poisson solve to relative residual 1e-9, UpdatePressure,
Projection and Divergence, with the immersed-boundary reconstruction
(ibm_interpolation_advanced) of the sphere's IB nodes where the reference
runs it with -imm 1 (--no-ib leaves the IB nodes held).

Three contexts run the same steps from the same initial state: the FP64
path (FP64 V-cycle), the mixed path (the FP32 V-cycle with FP16-stored face
coefficients inside the FP64 FGMRES, CV_MIXED16) and, under "mixed_fp32",
the FP32 V-cycle with FP32 coefficients (CV_MIXED); under
"mixed_fp32_4level_rb" a fourth context runs CV_MIXED with the
preconditioner the mixed workload is stated with (4 fully coarsened
levels, 2 red-black pre- and post-sweeps).  ``value`` is the mixed
path's steps/s; the FP64 figures, the ratios, the per-step pressure error of
each mixed path against the FP64 path and the per-step Poisson residuals
sit beside it.  All inputs are resident in HBM before the timed region
starts; every step call returns after a synchronisation of the library's
stream.

``roofline``: the fine-level operator residual kernel of the V-cycle (the
most launched stencil of the step), its algorithmic bytes per launch over
its average device-event duration on the library stream, measured during
the timed steps.  ``traffic``: HBM bytes per launch of the same kernel from
the rocprofv3 FETCH_SIZE/WRITE_SIZE summary under profiles/, when present.

``step_roofline``: the whole step against the same peak, from the byte
model of the surveyed design (per storage point 4 x 428 + 65 +
sum_j (173 + V + 16 j) + 8 N + 16 + 307 bytes with the counted FGMRES
iterations N) over the measured step time.

``shard_share_8way``: one GPU's share of an 8-way k-slab decomposition of
the workload (512x256x32 cells at the same spacing, no halo exchange), its
momentum and per-FGMRES-iteration times to set beside 1/8 of the full
step's (a slab solved alone needs more iterations than the coupled solve,
so its whole step is not the shard's).

``cpu_baseline``: the C oracle (the reference algorithm restated: assembled
19-point CSR operator, FGMRES(30) with a 3-level V-cycle and block-Jacobi
ILU(0) Richardson smoothing, one block per thread, RK momentum) timed for
one fractional step of the same workload at 256x128x128 cells on every host
thread, scaled by cell count; ``cpu_baseline_1core`` the same on one thread
at 128x64x64 cells.  Both run before any GPU work.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "fractional steps/s at 512x256x256 (FP64 vs mixed) + p50/p99 ensemble latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cells_of(s):
    v = tuple(int(x) for x in s.lower().split("x"))
    if len(v) != 3:
        raise ValueError(s)
    return v


def workload_case(cells):
    from curvib_old_amd import cases
    full = cases.config_case(2)
    dt = cases.stable_dt(full)
    return cases.config_case(2, cells=tuple(cells), dt=dt)


SOLVER = dict(smoother="yline", omega=0.9, pre_sweeps=3, post_sweeps=3, coarse_sweeps=10,
              restart=30, max_it=200, rtol=1e-9, atol=0.0)


# the preconditioner the mixed-precision workload is stated with: an FP32
# V-cycle of 4 fully coarsened levels, 2 red-black pre- and post-sweeps
# (omega 1; 40 sweeps on the coarsest level, 64x32x32 at 512x256x256)
STATED_PC = dict(levels=4, semi=(0, 0, 0), coarsen_auto=0,
                 solver=dict(smoother="rb", omega=1.0, pre_sweeps=2, post_sweeps=2, coarse_sweeps=40, max_it=400))


def make_ctx(case, arrays, levels, semi, precision, coarsen_auto, ib=None, solver_over=None):
    from curvib_old_amd import solver
    xyz, nv, uc = arrays
    u = solver.UserCtx(case.nodes, case.bctype, levels, case.re, case.dt, semi=semi, coarsen_auto=coarsen_auto)
    u.set_solver(precision=precision, **{**SOLVER, **(solver_over or {})})
    u.set_grid(xyz)
    u.set_nvert(nv)
    if ib is not None:
        u.set_ib_nodes(*ib)
    u.put("ucont", uc)
    return u


def reset(u, arrays):
    """The state of a fresh context: initial fluxes, zero p, phi and Ucat
    (Ucat at immersed-boundary nodes carries over from step to step, and the
    first reconstruction of a step starts from it)."""
    _, _, uc = arrays
    u.put("ucont", uc)
    z = np.zeros(u._shape("p", -1))
    u.put("p", z)
    u.put("phi", z)
    u.put("ucat", np.zeros(u._shape("ucat", -1)))


def timed_steps(u, tag, warmup, steps, prof=None, stats_out=None):
    for s in range(warmup):
        st = u.Flow_Solver()
        log(f"[{tag}] warmup {s}: its={st['ksp_its']} t_p={st['t_p']:.4f}s")
    if prof is not None:
        u.prof_select(*prof)
    its, stats = [], []
    t0 = time.perf_counter()
    for s in range(steps):
        st = u.Flow_Solver()
        its.append(st["ksp_its"])
        stats.append(st)
    t1 = time.perf_counter()
    if stats_out is not None:
        stats_out.extend(stats)
    pr = u.prof_read() if prof is not None else None
    if prof is not None:
        u.prof_select("none")
    log(f"[{tag}] {steps} steps in {t1 - t0:.3f}s its={its}")
    return t1 - t0, its, pr


def accuracy_pass(u64, umx, arrays, steps):
    reset(u64, arrays)
    reset(umx, arrays)
    rows = []
    for s in range(steps):
        a = u64.Flow_Solver()
        b = umx.Flow_Solver()
        p64 = u64.get("p")
        pmx = umx.get("p")
        den = float(np.linalg.norm(p64))
        err = float(np.linalg.norm(pmx - p64)) / den if den > 0 else 0.0
        rows.append(dict(step=s, p_rel_l2=err,
                         fp64_its=a["ksp_its"], mixed_its=b["ksp_its"],
                         fp64_rel_res=a["ksp_rnorm"] / a["ksp_bnorm"] if a["ksp_bnorm"] else 0.0,
                         mixed_rel_res=b["ksp_rnorm"] / b["ksp_bnorm"] if b["ksp_bnorm"] else 0.0,
                         fp64_maxdiv=a["maxdiv"], mixed_maxdiv=b["maxdiv"]))
        log(f"[accuracy] step {s}: p rel L2 {err:.3e} its {a['ksp_its']}/{b['ksp_its']}")
    return rows


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch recorded by the rocprofv3 counter summary of this
    workload's kernel, or None."""
    p = os.path.join(HERE, "profiles", "traffic_stencil_resid_512x256x256.json")
    try:
        with open(p) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    e = d.get(kernel_tag)
    return float(e["hbm_bytes_per_launch"]) if e and "hbm_bytes_per_launch" in e else None


def cpu_baseline(sample_cells, full_cells, threads, ib=True):
    """One fractional step of the oracle on a reduced version of the
    workload on ``threads`` host threads (the block-Jacobi smoother takes
    one block per thread); steps/s at the full size scaled by cell count."""
    from curvib_old_amd import ibsearch
    import oracle.pyoracle as po
    po.set_threads(threads)
    case = workload_case(sample_cells)
    o, (xyz, nv, _) = po.from_case(case, nlev=3)
    if ib:
        o.set_ib_nodes(*ibsearch.sphere_ib_nodes(case, xyz, nv), 0)
    o.set_solver(restart=30, max_it=500, rtol=SOLVER["rtol"], atol=0.0, smooth_it=2)
    t0 = time.perf_counter()
    o.call("o_flow_step")
    dt = time.perf_counter() - t0
    st = o.stats()
    o.close()
    ncell = float(np.prod(sample_cells))
    nfull = float(np.prod(full_cells))
    cups = ncell / dt
    log(f"[cpu] oracle step at {sample_cells}: {dt:.2f}s its={int(st['its'])} threads={int(st['threads'])}")
    return {
        "value": cups / nfull,
        "unit": "steps/s",
        "cores": int(st["threads"]),
        "kind": "port",
        "sample": (f"one fractional step of the {'x'.join(map(str, sample_cells))} version of the workload "
                   f"({ncell / 1e6:.2f} M cells, {int(st['its'])} FGMRES its to rtol {SOLVER['rtol']:g}) "
                   f"on the C oracle in {dt:.2f} s = {cups / 1e6:.3f} M cell-updates/s, scaled to "
                   f"{nfull / 1e6:.1f} M cells"),
        "cell_updates_per_s": cups,
    }


def step_bytes(its, vcycle):
    """Algorithmic HBM bytes per storage point of one fractional step, the
    byte model of the surveyed design with the counted FGMRES iterations N:
    4 x 428 (FormFunction1 fused with the RK stage update) + 65 (PoissonRHS,
    VolumeFlux_rev) + sum_j (173 + V + 16 j) (apply, V-cycle,
    orthogonalisation, normalisation, nullspace of iteration j)
    + (8 N + 16) (x += Z y) + 307 (Projection, UpdatePressure, Divergence,
    Contra2Cart), V = 652 B for the FP64 V-cycle, 363 B for the FP32 one
    with its conversions."""
    n = int(its)
    return 4 * 428 + 65 + sum(173 + vcycle + 16 * j for j in range(1, n + 1)) + (8 * n + 16) + 307


def step_roofline(npts, its, seconds, vcycle):
    b = float(sum(step_bytes(n, vcycle) for n in its)) * npts
    achieved = b / seconds / 1e9
    return {"bound": "hbm", "model": "4*428 + 65 + sum_j(173 + V + 16j) + (8N + 16) + 307 B per storage point",
            "V": vcycle, "its": list(its), "bytes_per_step": b / len(its), "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS}


def shard_share(full_case, steps, warmup, mixed):
    """One GPU's share of an 8-way k-slab decomposition of the workload: the
    512x256x32 slab at the same grid spacing (z length 1/8, periodic; the
    wall warp scaled by 1/8 with its z period, so its slope is the full
    grid's), no halo exchange and no body, stepped alone in the mixed
    precision."""
    from curvib_old_amd import cases, serving
    cells = (full_case.cells[0], full_case.cells[1], full_case.cells[2] // 8)
    case = cases.config_case(2, cells=cells, lz=full_case.lz / 8, amp=full_case.amp / 8, sphere=None)
    case = cases.config_case(2, cells=cells, lz=full_case.lz / 8, amp=full_case.amp / 8, sphere=None,
                             dt=min(full_case.dt, cases.stable_dt(case)))
    arrays = cases.build(case)
    levels = serving.levels_for(cells, semi=(0, 1, 0), coarsen_auto=1)
    u = make_ctx(case, arrays, levels, (0, 1, 0), mixed, 1)
    stats = []
    try:
        t, its, _ = timed_steps(u, f"{mixed} shard share", warmup, steps, stats_out=stats)
    finally:
        u.close()
    return {"cells": list(cells), "mg_levels": levels, "precision": mixed, "dt": case.dt, "ms_per_step": 1e3 * t / steps,
            "its": its, "cell_updates_per_s": float(np.prod(cells)) * steps / t, **phase_ms(stats)}


def phase_ms(stats):
    """Mean momentum and pressure-solve milliseconds per step, and the solve's
    milliseconds per FGMRES iteration, from the library's step statistics."""
    n = max(len(stats), 1)
    t_mom = 1e3 * sum(x["t_mom"] for x in stats) / n
    t_p = 1e3 * sum(x["t_p"] for x in stats) / n
    its = sum(x["ksp_its"] for x in stats) / n
    return {"momentum_ms": t_mom, "poisson_ms": t_p, "poisson_ms_per_iteration": t_p / its if its else None}


def serving_workloads(n_serve, n_ens, seed, log_workers, ib=True):
    """Requests of the serving stream (Poisson arrivals, lam 50/s, grids
    32^3..192^3) and of the ensemble (grids 32^3..128x96x96), inputs built
    in spawned worker processes before any GPU work."""
    from curvib_old_amd import serving
    serve, arrivals, ens = serving.stated_workloads(n_serve, n_ens, seed)
    t0 = time.perf_counter()
    arrays = serving.build_inputs(serve + ens, workers=log_workers, ib=ib)
    log(f"[serving] built {len(arrays)} case inputs in {time.perf_counter() - t0:.1f}s")
    return list(zip(serve, arrays[:n_serve])), arrivals, list(zip(ens, arrays[n_serve:]))


def batch_ensemble_only(args):
    """Only the batched-ensemble leg (the same 64 seeded cases as the full
    run): one untimed step, then --ensemble-steps timed steps; one JSON line
    with its latency and throughput (for profiling that leg alone)."""
    from curvib_old_amd import serving
    _, _, ens = serving.stated_workloads(args.serving_requests, args.ensemble_cases, 2026)
    arrays = serving.build_inputs(ens, workers=args.build_workers, ib=not args.no_ib)
    r = serving.run_batched(list(zip(ens, arrays)), args.ensemble_steps, args.batch_ensemble_only, log=log,
                            max_levels=args.ensemble_max_levels, **SOLVER)
    print(json.dumps({"workload": f"{len(ens)}-case ensemble, one cv_batch", **r}), flush=True)


def run_serving(serve_reqs, arrivals, ens_reqs, nsteps, ens_steps, mixed="mixed16", batch_steps=100):
    """The serving and ensemble legs in FP64 and in the mixed leg's precision
    (reported under the "mixed" keys)."""
    from curvib_old_amd import serving
    opts = {k: v for k, v in SOLVER.items()}
    out = {}
    precs = (("fp64", "fp64"), ("mixed", mixed))
    for key, prec in precs:
        r = serving.run_stream(serve_reqs, arrivals, nsteps, prec, budget_bytes=200e9, slots=16, log=None, **opts)
        r.pop("records")
        out[f"serving_{key}"] = r
        log(f"[serving {prec}] {r['cases']} cases x {nsteps} steps: step p50 {r['step_latency'].get('p50_ms', 0):.1f} ms "
            f"p99 {r['step_latency'].get('p99_ms', 0):.1f} ms, {r['cell_updates_per_s'] / 1e6:.1f} M cell-updates/s, "
            f"peak HBM {r['bytes_peak'] / 2**30:.1f} GiB, failed {r['failed']}")
    for key, prec in precs:
        r = serving.run_stream(ens_reqs, np.zeros(len(ens_reqs)), ens_steps, prec, budget_bytes=200e9, slots=16,
                               **opts)
        r.pop("records")
        out[f"ensemble_{key}"] = r
        log(f"[ensemble {prec}] {r['cases']} cases x {ens_steps} steps: makespan {r['makespan_s']:.2f} s, step p50 "
            f"{r['step_latency'].get('p50_ms', 0):.1f} ms p99 {r['step_latency'].get('p99_ms', 0):.1f} ms, "
            f"{r['cell_updates_per_s'] / 1e6:.1f} M cell-updates/s")
    # the serving stream again through the batched server: every admitted case
    # advances in one cv_batch, rebuilt as cases arrive and finish
    for key, prec in precs:
        r = serving.run_stream(serve_reqs, arrivals, nsteps, prec, budget_bytes=200e9, slots=64, log=None,
                               mode="batched", **opts)
        r.pop("records")
        out[f"serving_batched_{key}"] = r
        log(f"[serving batched {prec}] {r['cases']} cases x {nsteps} steps: step p50 "
            f"{r['step_latency'].get('p50_ms', 0):.1f} ms p99 {r['step_latency'].get('p99_ms', 0):.1f} ms, "
            f"{r['cell_updates_per_s'] / 1e6:.1f} M cell-updates/s, peak HBM {r['bytes_peak'] / 2**30:.1f} GiB, "
            f"failed {r['failed']}")
    # the same ensemble as one batch: one launch per multigrid depth for all cases
    for key, prec in precs:
        r = serving.run_batched(ens_reqs, batch_steps, prec, **opts)
        out[f"ensemble_batched_{key}"] = r
        log(f"[ensemble batched {prec}] {r['cases']} cases x {batch_steps} steps: {r['makespan_s']:.2f} s, step p50 "
            f"{r['step_latency'].get('p50_ms', 0):.1f} ms p99 {r['step_latency'].get('p99_ms', 0):.1f} ms, "
            f"{r['cell_updates_per_s'] / 1e6:.1f} M cell-updates/s, max its {r['max_its']}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", default="512x256x256")
    ap.add_argument("--levels", type=int, default=8)
    ap.add_argument("--coarsen-auto", type=int, default=1)
    ap.add_argument("--semi", default="0,1,0")
    ap.add_argument("--cpu-sample", default="256x128x128")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: OMP_NUM_THREADS, else every core")
    ap.add_argument("--cpu1-sample", default="128x64x64", help="the one-core CPU sample")
    ap.add_argument("--no-fp32-leg", action="store_true", help="skip the CV_MIXED (FP32 coefficient) leg")
    ap.add_argument("--no-stated-leg", action="store_true",
                    help="skip the leg with the stated preconditioner (FP32, 4 levels, 2+2 red-black sweeps)")
    ap.add_argument("--no-shard-leg", action="store_true",
                    help="skip the one-GPU share of an 8-way k-slab decomposition (512x256x32)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-ib", action="store_true", help="no immersed-boundary reconstruction")
    ap.add_argument("--restart", type=int, default=SOLVER["restart"])
    ap.add_argument("--pre", type=int, default=SOLVER["pre_sweeps"])
    ap.add_argument("--post", type=int, default=SOLVER["post_sweeps"])
    ap.add_argument("--coarse", type=int, default=SOLVER["coarse_sweeps"])
    ap.add_argument("--omega", type=float, default=SOLVER["omega"])
    ap.add_argument("--smoother", default=SOLVER["smoother"], choices=("jacobi", "rb", "yline"))
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--mixed", default="mixed16", choices=("mixed", "mixed16"),
                    help="precision of the mixed leg: FP32 V-cycle, or with FP16 face coefficients")
    ap.add_argument("--ensemble-max-levels", type=int, default=8)
    ap.add_argument("--batch-ensemble-only", choices=("mixed", "mixed16", "fp64"), default=None,
                    help="run only the batched-ensemble leg in this precision")
    ap.add_argument("--serving-requests", type=int, default=50)
    ap.add_argument("--serving-steps", type=int, default=20)
    ap.add_argument("--ensemble-cases", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--ensemble-batch-steps", type=int, default=100,
                    help="timed steps of the batched ensemble (its p99 needs about 100 samples)")
    ap.add_argument("--build-workers", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    SOLVER.update(restart=args.restart, pre_sweeps=args.pre, post_sweeps=args.post, coarse_sweeps=args.coarse,
                  omega=args.omega, smoother=args.smoother)
    if args.batch_ensemble_only:
        batch_ensemble_only(args)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; runs on more than one GPU are not supported "
              "(the single-case path would shard by k-slabs with halo exchange, not built here)",
              file=sys.stderr)
        sys.exit(2)

    serve_in = None
    if not args.no_serving:
        serve_in = serving_workloads(args.serving_requests, args.ensemble_cases, 2026, args.build_workers,
                                     ib=not args.no_ib)

    cells = cells_of(args.cells)
    # the CPU legs run before any GPU work: all host threads, then one
    cpu = cpu1 = None
    if not args.no_cpu:
        nthr = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        cpu = cpu_baseline(cells_of(args.cpu_sample), cells, nthr, ib=not args.no_ib)
        cpu1 = cpu_baseline(cells_of(args.cpu1_sample), cells, 1, ib=not args.no_ib)
    semi = tuple(int(v) for v in args.semi.split(","))
    from curvib_old_amd import cases
    case = workload_case(cells)
    t0 = time.perf_counter()
    arrays = cases.build(case)
    ib = None
    if not args.no_ib:
        from curvib_old_amd import ibsearch
        ib = ibsearch.sphere_ib_nodes(case, arrays[0], arrays[1])
    log(f"[setup] case {cells} dt={case.dt:g} built in {time.perf_counter() - t0:.1f}s, "
        f"{0 if ib is None else len(ib[0])} IB nodes")

    u64 = make_ctx(case, arrays, args.levels, semi, "fp64", args.coarsen_auto, ib)
    umx = make_ctx(case, arrays, args.levels, semi, args.mixed, args.coarsen_auto, ib)
    log(f"[setup] device bytes per context: {u64.device_bytes() / 2**30:.1f} GiB")

    t64, its64, _ = timed_steps(u64, "fp64", args.warmup, args.steps)
    stats_mx = []
    tmx, itsmx, pr = timed_steps(umx, args.mixed, args.warmup, args.steps, prof=("resid", -1, 1), stats_out=stats_mx)
    # the north star's FP32 V-cycle (FP32 coefficients) as a leg of its own
    leg32 = None
    if args.mixed != "mixed" and not args.no_fp32_leg:
        u32 = make_ctx(case, arrays, args.levels, semi, "mixed", args.coarsen_auto, ib)
        t32, its32, _ = timed_steps(u32, "mixed", args.warmup, args.steps)
        acc32 = [] if args.no_accuracy else accuracy_pass(u64, u32, arrays, args.steps)
        u32.close()
        leg32 = {"precision": "mixed", "steps_per_s": args.steps / t32, "ms_per_step": 1e3 * t32 / args.steps,
                 "its": its32, "cell_updates_per_s": float(np.prod(cells)) * args.steps / t32,
                 "over_fp64": t64 / t32, "accuracy": acc32}

    # the stated preconditioner point: FP32 V-cycle, 4 levels, 2+2 red-black sweeps
    leg_pc = None
    if not args.no_stated_leg:
        upc = make_ctx(case, arrays, STATED_PC["levels"], STATED_PC["semi"], "mixed", STATED_PC["coarsen_auto"], ib,
                       STATED_PC["solver"])
        tpc, itspc, _ = timed_steps(upc, "mixed 4-level rb", args.warmup, args.steps)
        accpc = [] if args.no_accuracy else accuracy_pass(u64, upc, arrays, args.steps)
        upc.close()
        leg_pc = {"precision": "mixed", "mg_levels": STATED_PC["levels"], "semi_coarsening": list(STATED_PC["semi"]),
                  **STATED_PC["solver"], "steps_per_s": args.steps / tpc, "ms_per_step": 1e3 * tpc / args.steps,
                  "its": itspc, "cell_updates_per_s": float(np.prod(cells)) * args.steps / tpc,
                  "over_fp64": t64 / tpc, "accuracy": accpc}

    acc = [] if args.no_accuracy else accuracy_pass(u64, umx, arrays, args.steps)
    shard = None if args.no_shard_leg else shard_share(case, args.steps, args.warmup, args.mixed)
    # host-array exchange of a drop-in caller that hands Ucont in and takes
    # Ucont and P back every step (pageable numpy buffers, PCIe)
    t0 = time.perf_counter()
    uc = umx.get("ucont")
    umx.put("ucont", uc)
    pp = umx.get("p")
    xchg_ms = 1e3 * (time.perf_counter() - t0)
    del uc, pp
    u64.close()
    umx.close()
    del arrays

    srv = None
    if serve_in is not None:
        srv = run_serving(*serve_in, args.serving_steps, args.ensemble_steps, args.mixed, args.ensemble_batch_steps)
        serve_in = None

    ncell = float(np.prod(cells))
    npts = float(np.prod([c + 2 for c in cells]))
    sps_mx = args.steps / tmx
    sps_64 = args.steps / t64

    launches, ms, byts = pr
    achieved = (byts / (ms * 1e-3)) / 1e9 if ms > 0 else 0.0
    roof = {
        "bound": "hbm",
        "kernel": ("k_stencil<float, 1, 4, true, true> (fine-level operator residual r = b - A x of the FP32 "
                   "V-cycle, FP16 face coefficients: 36 B per point)" if args.mixed == "mixed16" else
                   "k_stencil<float, 1, 4, true> (fine-level operator residual r = b - A x of the FP32 V-cycle: "
                   "56 B per point)"),
        "launches": launches,
        "avg_us": 1e3 * ms / launches if launches else None,
        "bytes_per_launch": byts / launches if launches else None,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        # the committed counter summary is of the default (FP16 coefficient) kernel
        "traffic": traffic_from_profiles("k_stencil<float, 1>") if args.mixed == "mixed16" else None,
    }

    step_roof = {"fp64": step_roofline(npts, its64, t64, 652), "mixed": step_roofline(npts, itsmx, tmx, 363)}
    if leg32:
        step_roof["mixed_fp32"] = step_roofline(npts, leg32["its"], args.steps / leg32["steps_per_s"], 363)

    out = {
        "metric": METRIC,
        "value": sps_mx,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32",
        "data": "synthetic",
        "config": {
            "workload": f"warped channel + immersed sphere {'x'.join(map(str, cells))} cells, Re {case.re:g}, "
                        f"dt {case.dt:g}, Poisson rtol {SOLVER['rtol']:g}",
            "cells": int(ncell), "storage_points": int(npts), "ib_nodes": 0 if ib is None else len(ib[0]),
            "mg_levels": args.levels, "semi_coarsening": list(semi), "coarsen_auto": args.coarsen_auto,
            "mixed_precision": args.mixed, **SOLVER,
        },
        "fp64": {"steps_per_s": sps_64, "ms_per_step": 1e3 * t64 / args.steps, "its": its64,
                 "cell_updates_per_s": ncell * sps_64},
        "mixed": {"precision": args.mixed, "steps_per_s": sps_mx, "ms_per_step": 1e3 * tmx / args.steps,
                  "its": itsmx, "cell_updates_per_s": ncell * sps_mx, **phase_ms(stats_mx)},
        "mixed_over_fp64": sps_mx / sps_64,
        "mixed_fp32": leg32,
        "mixed_fp32_4level_rb": leg_pc,
        "shard_share_8way": shard,
        "host_exchange_ms_per_step": xchg_ms,
        "pcie_inclusive_mixed_ms_per_step": 1e3 * tmx / args.steps + xchg_ms,
        "accuracy": acc,
        "ensemble_latency_ms": (None if srv is None else {
            "serving_mixed_p50": srv["serving_mixed"]["step_latency"].get("p50_ms"),
            "serving_mixed_p99": srv["serving_mixed"]["step_latency"].get("p99_ms"),
            "serving_fp64_p50": srv["serving_fp64"]["step_latency"].get("p50_ms"),
            "serving_fp64_p99": srv["serving_fp64"]["step_latency"].get("p99_ms"),
            "ensemble_mixed_p50": srv["ensemble_mixed"]["step_latency"].get("p50_ms"),
            "ensemble_mixed_p99": srv["ensemble_mixed"]["step_latency"].get("p99_ms"),
            "serving_batched_mixed_p50": srv["serving_batched_mixed"]["step_latency"].get("p50_ms"),
            "serving_batched_mixed_p99": srv["serving_batched_mixed"]["step_latency"].get("p99_ms"),
            "serving_batched_fp64_p50": srv["serving_batched_fp64"]["step_latency"].get("p50_ms"),
            "serving_batched_fp64_p99": srv["serving_batched_fp64"]["step_latency"].get("p99_ms"),
            "ensemble_batched_mixed_p50": srv["ensemble_batched_mixed"]["step_latency"].get("p50_ms"),
            "ensemble_batched_mixed_p99": srv["ensemble_batched_mixed"]["step_latency"].get("p99_ms"),
            "ensemble_batched_fp64_p50": srv["ensemble_batched_fp64"]["step_latency"].get("p50_ms"),
            "ensemble_batched_fp64_p99": srv["ensemble_batched_fp64"]["step_latency"].get("p99_ms"),
        }),
        "serving": srv,
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "gpu_mixed_over_cpu": (sps_mx / cpu["value"]) if cpu else None,
        "gpu_mixed_over_cpu_1core": (sps_mx / cpu1["value"]) if cpu1 else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
