"""GPU parity of single routines and multi-step runs against the oracle,
the mixed-precision path against the FP64 path, the smoother options and
the not-converged status.

Tolerances: FP64 routine outputs 1e-12 relative; fields after each
converged step (both sides solved to rtol 1e-13) 1e-10 relative L2, the
achieved errors printed per step; mixed against FP64 pressure at rtol 1e-9:
1e-8 relative L2 (also with the FP16 coefficient V-cycle)."""
import numpy as np
import pytest

from curvib_old_amd import _lib, cases
from curvib_old_amd.solver import from_case as gpu_case
from oracle.pyoracle import from_case as oracle_case
from conftest import assert_fields_match

pytestmark = pytest.mark.gpu

SPHERE = cases.ChannelCase(cells=(32, 24, 16), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.45), re=500.0, seed=21,
                           dt=2e-4)
CHANNEL = cases.ChannelCase(cells=(24, 16, 16), amp=0.05, beta=2.0, sphere=None, re=180.0, seed=22, dt=4e-4)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_rhs_volumeflux_divergence():
    u, _ = gpu_case(SPHERE, mglevels=3)
    o, _ = oracle_case(SPHERE, nlev=3)
    try:
        # Divergence of the initial fluxes (main.c:469-593): value and first argmax
        dv, di = u.Divergence()
        do = o.call("o_divergence")
        assert abs(dv - do) <= 1e-12 * abs(do)
        assert di == int(o.stats()["maxdiv_idx"])
        # VolumeFlux_rev (poisson.c:3712-3944): corrected fluxes and the IB flux/area sums
        fa = u.VolumeFlux_rev()
        o.call("o_volumeflux_rev")
        so = o.stats()
        assert rel(u.get("ucont"), o.get("ucont")) < 1e-13
        for q, k in enumerate(("ibm_flux0", "ibm_area0", "ibm_flux1", "ibm_area1")):
            assert abs(fa[q] - so[k]) <= 1e-12 * max(abs(so[k]), 1e-12), (k, fa[q], so[k])
        # PoissonRHS (poisson.c:1509-1576)
        b = u.PoissonRHS()
        o.call("o_rhs_into_b")
        assert rel(b, o.get("b")) < 1e-12
    finally:
        u.close()
        o.close()


def test_three_steps_track_the_oracle():
    u, _ = gpu_case(CHANNEL, mglevels=3)
    o, _ = oracle_case(CHANNEL, nlev=3)
    u.set_solver(rtol=1e-13, max_it=400, smoother="yline", pre_sweeps=3, post_sweeps=3, omega=0.9)
    o.set_solver(rtol=1e-13, max_it=400)
    try:
        for s in range(3):
            u.Flow_Solver()
            o.call("o_flow_step")
            assert_fields_match(f"rk channel step {s}", u, o)
    finally:
        u.close()
        o.close()


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
def test_mixed_pressure_matches_fp64(precision):
    """mixed16: the V-cycle's stencil reads FP16 face coefficients; the
    FP64 FGMRES still drives the FP64 residual to rtol, so the pressure
    meets the same bound, in at most two more iterations per step."""
    case = cases.ChannelCase(cells=(64, 48, 32), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.25), re=500.0,
                             seed=23)
    case = cases.ChannelCase(**{**case.__dict__, "dt": cases.stable_dt(case)})
    opts = dict(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, omega=0.9)
    a, _ = gpu_case(case, mglevels=4, semi=(0, 1, 0), coarsen_auto=1)
    b, _ = gpu_case(case, mglevels=4, semi=(0, 1, 0), coarsen_auto=1)
    a.set_solver(precision="fp64", **opts)
    b.set_solver(precision=precision, **opts)
    try:
        for s in range(3):
            sa, sb = a.Flow_Solver(), b.Flow_Solver()
            assert sa["ksp_rnorm"] <= 1e-9 * sa["ksp_bnorm"] and sb["ksp_rnorm"] <= 1e-9 * sb["ksp_bnorm"]
            assert sb["ksp_its"] <= sa["ksp_its"] + 2, (s, sa["ksp_its"], sb["ksp_its"])
            e = rel(b.get("p"), a.get("p"))
            assert e < 1e-8, (s, e)
    finally:
        a.close()
        b.close()


def test_mixed16_takes_its_own_path_and_follows_nvert():
    """The FP16 coefficients change the preconditioner (a residual history
    unlike mixed's); they are made when the option is set and again when
    nvert is set, so both orders give the same solve."""
    case = cases.ChannelCase(cells=(48, 32, 32), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.3), re=500.0, seed=24,
                             dt=2e-4)
    opts = dict(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, omega=0.9)
    ctx = [gpu_case(case, mglevels=4, semi=(0, 1, 0), coarsen_auto=1) for _ in range(3)]
    (m, _), (h1, _), (h2, (_, nv, _)) = ctx
    try:
        m.set_solver(precision="mixed", **opts)
        h1.set_solver(precision="mixed16", **opts)
        h2.set_solver(precision="mixed16", **opts)
        h2.set_nvert(nv)
        for u in (m, h1, h2):
            u.FormBCS()
        sm, s1, s2 = m.PoissonSolver_MG(), h1.PoissonSolver_MG(), h2.PoissonSolver_MG()
        assert sm["converged"] == s1["converged"] == s2["converged"] == 1
        assert s1["ksp_rnorm"] != sm["ksp_rnorm"]
        assert s1["ksp_its"] == s2["ksp_its"] and s1["ksp_rnorm"] == s2["ksp_rnorm"]
        assert np.array_equal(h1.get("phi"), h2.get("phi"))
        assert rel(h1.get("phi"), m.get("phi")) < 1e-7
    finally:
        for u, _ in ctx:
            u.close()


@pytest.mark.parametrize("smoother", ["jacobi", "rb", "yline"])
def test_mixed16_smoothers_reach_the_oracle_solution(smoother):
    """The FP16-coefficient V-cycle with each smoother (Jacobi and red-black
    run the damped-Jacobi stencil mode on the FP16 packs) still converges
    FGMRES to the oracle's solution."""
    u, _ = gpu_case(SPHERE, mglevels=3)
    o, _ = oracle_case(SPHERE, nlev=3)
    u.set_solver(precision="mixed16", rtol=1e-12, max_it=500, smoother=smoother, pre_sweeps=2, post_sweeps=2,
                 omega={"jacobi": 0.8, "rb": 1.0, "yline": 0.9}[smoother])
    o.set_solver(rtol=1e-12, max_it=500)
    try:
        u.FormBCS()
        o.call("o_formbcs")
        st = u.PoissonSolver_MG()
        assert o.call("o_poisson_solve") == 0
        assert st["converged"] == 1
        assert rel(u.get("phi"), o.get("phi")) < 1e-8
    finally:
        u.close()
        o.close()


def test_mixed16_long_y_lines():
    """Y-lines of more than 256 rows keep the FP32 line factors (the FP16
    pack serves 8 or 16 rows per thread) beside the FP16 stencil.  Both
    solves meet rtol on the FP64 residual with the same iteration counts
    (+2 allowed); the tall grid is ill-conditioned, so the two converged
    pressures differ by about cond(A) x rtol, and the check uses rtol 1e-12
    (measured difference 5e-10 over three steps, 1e-11 with FP32
    coefficients)."""
    case = cases.ChannelCase(cells=(16, 300, 16), amp=0.05, beta=2.0, sphere=None, re=180.0, seed=25, dt=2e-4)
    opts = dict(rtol=1e-12, max_it=400, smoother="yline", pre_sweeps=3, post_sweeps=3, omega=0.9)
    a, _ = gpu_case(case, mglevels=3, semi=(0, 1, 0))
    b, _ = gpu_case(case, mglevels=3, semi=(0, 1, 0))
    try:
        a.set_solver(precision="fp64", **opts)
        b.set_solver(precision="mixed16", **opts)
        for s in range(3):
            sa, sb = a.Flow_Solver(), b.Flow_Solver()
            assert sa["converged"] == sb["converged"] == 1
            assert sb["ksp_its"] <= sa["ksp_its"] + 2, (s, sa["ksp_its"], sb["ksp_its"])
            assert rel(b.get("p"), a.get("p")) < 1e-8, s
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("smoother", ["jacobi", "rb", "yline"])
def test_smoothers_reach_the_same_solution(smoother):
    u, _ = gpu_case(SPHERE, mglevels=3)
    o, _ = oracle_case(SPHERE, nlev=3)
    u.set_solver(rtol=1e-12, max_it=500, smoother=smoother, pre_sweeps=2, post_sweeps=2,
                 omega={"jacobi": 0.8, "rb": 1.0, "yline": 0.9}[smoother])
    o.set_solver(rtol=1e-12, max_it=500)
    try:
        u.FormBCS()
        o.call("o_formbcs")
        st = u.PoissonSolver_MG()
        assert o.call("o_poisson_solve") == 0
        assert st["converged"] == 1
        assert rel(u.get("phi"), o.get("phi")) < 1e-8
    finally:
        u.close()
        o.close()


def test_not_converged_status():
    u, _ = gpu_case(SPHERE, mglevels=3)
    u.set_solver(rtol=1e-14, max_it=2)
    try:
        with pytest.raises(_lib.CvError) as e:
            u.Flow_Solver()
        assert e.value.code == _lib.CV_ERR_NOT_CONVERGED
        st = u.Flow_Solver(allow_unconverged=True)
        assert st["ksp_its"] == 2 and st["converged"] == 0
    finally:
        u.close()


def test_restart_from_field_files(tmp_path):
    """Ucont_P_Binary_Output then Ucont_P_Binary_Input (main.c:75-330):
    a run restarted from the written fields continues bit for bit."""
    from curvib_old_amd import formats
    a, _ = gpu_case(CHANNEL, mglevels=3)
    b, _ = gpu_case(CHANNEL, mglevels=3)
    try:
        a.Flow_Solver()
        formats.write_fields(a, 1, 0, str(tmp_path))
        a.Flow_Solver()
        nv = formats.read_fields(b, 1, 0, str(tmp_path))
        assert np.array_equal(nv, a.get("nvert"))
        b.Flow_Solver()
        for f in ("ucont", "p", "ucat"):
            assert np.array_equal(a.get(f), b.get(f)), f
    finally:
        a.close()
        b.close()


def test_input_deck_matches_direct_setup(tmp_path):
    """grid.dat + bcs.dat + control.dat give the same context as the
    in-memory setup."""
    from curvib_old_amd import formats
    xyz, nv, uc = cases.build(CHANNEL)
    formats.write_grid_dat(str(tmp_path / "grid.dat"), [xyz])
    formats.write_bcs_dat(str(tmp_path / "bcs.dat"), [CHANNEL.bctype])
    (tmp_path / "control.dat").write_text(f"-ren {CHANNEL.re!r}\n-dt {CHANNEL.dt!r}\n-mg_level 3\n"
                                          "-ps_ksp_rtol 1e-12\n-ps_ksp_max_it 300\n")
    d, opts = formats.load_case_dir(str(tmp_path), nvert=nv)
    u, _ = gpu_case(CHANNEL, mglevels=3)
    u.set_solver(rtol=1e-12, max_it=300)
    try:
        d.put("ucont", uc)
        for f in ("aj", "csi", "ieta", "kzet"):
            assert np.array_equal(d.get(f), u.get(f)), f
        d.Flow_Solver()
        u.Flow_Solver()
        assert np.array_equal(d.get("p"), u.get("p"))
    finally:
        d.close()
        u.close()


@pytest.mark.parametrize("name,cfl", [("channel", 0.3), ("sphere", 0.3), ("channel", 4.0)])
def test_imprk_tracks_the_oracle(name, cfl):
    """ImpRK (implicitsolver.c:7094-7256) as the momentum scheme of two
    fractional steps (BDF1, then BDF2): same pseudo-iteration counts and
    lambda as the oracle, fields to 1e-10.  A pseudo-time factor of 4 makes
    the pseudo iterations of the first step grow, so the lambda-halving
    branch runs (the oracle's final lambda is below the initial one); that
    case runs the one step (its BDF2 step overflows in the oracle too)."""
    case = {"channel": CHANNEL, "sphere": SPHERE}[name]
    u, _ = gpu_case(case, mglevels=3)
    o, _ = oracle_case(case, nlev=3)
    imp = dict(imp_max_it=30, imp_atol=1e-7, imp_rtol=1e-4, cfl=cfl)
    u.set_momentum("imprk", **imp)
    o.set_momentum("imprk", **imp)
    u.set_solver(rtol=1e-13, max_it=400)
    o.set_solver(rtol=1e-13, max_it=400)
    try:
        for s in range(2 if cfl < 1 else 1):
            st = u.Flow_Solver()
            o.call("o_flow_step")
            so = o.imp_stats()
            if cfl > 1 and s == 0:
                assert so["lambda"] < cfl, so
            assert st["imp_its"] == int(so["its"]), (s, st["imp_its"], so)
            assert st["imp_lambda"] == so["lambda"]
            assert abs(st["imp_normdU"] - so["normdU"]) <= 1e-9 * so["normdU"]
            assert_fields_match(f"imprk {name} cfl {cfl} step {s}", u, o)
    finally:
        u.close()
        o.close()


@pytest.mark.parametrize("walls", [False, True])
def test_face_flux_variants_agree(monkeypatch, walls):
    """The LDS-tiled face-flux pass and the global-gather pass of
    FormFunction1 (rhs.c:1636-2703) give identical bits, and both match the
    oracle; the grid spans several tiles in i, j and k so that every tile
    seam and the periodic ghost map of the tile loads are exercised."""
    kw = dict(bctype=(1, 1, 1, 1, 1, 1)) if walls else {}
    case = cases.ChannelCase(cells=(136, 22, 40), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.35), re=400.0,
                             seed=31, **kw)
    u, _ = gpu_case(case, mglevels=2)
    o, _ = oracle_case(case, nlev=2)
    try:
        u.FormBCS()
        o.call("o_formbcs")
        o.call("o_ff1_into_rhs")
        monkeypatch.setenv("CV_RC_KERNEL", "0")
        r0 = u.FormFunction1()
        monkeypatch.setenv("CV_RC_KERNEL", "1")
        r1 = u.FormFunction1()
        assert np.array_equal(r0, r1)
        assert rel(r1, o.get("rhs")) < 1e-12
    finally:
        u.close()
        o.close()


@pytest.mark.parametrize("kernel,per_step", [("resid", None), ("apply", None), ("line", None), ("rct", 4),
                                             ("ff1", 4), ("orth", None)])
def test_kernel_timing_counts_launches(kernel, per_step):
    """cv_prof_select / cv_prof_read time one kernel class with device
    events while the context runs normally: launches, time and algorithmic
    bytes are recorded (four RK stages give four face-flux passes)."""
    u, _ = gpu_case(SPHERE, mglevels=3)
    u.set_solver(rtol=1e-9, max_it=200, smoother="yline", pre_sweeps=3, post_sweeps=3, omega=0.9)
    try:
        u.prof_select(kernel, level=-1)
        u.Flow_Solver()
        launches, ms, nbytes = u.prof_read(reset=True)
        assert launches > 0 and ms > 0 and nbytes > 0, (launches, ms, nbytes)
        if per_step is not None:
            assert launches == per_step, launches
        u.prof_select("none")
        u.Flow_Solver()
        assert u.prof_read()[0] == 0
    finally:
        u.close()


@pytest.mark.parametrize("walls", [False, True])
def test_fused_rk_stage_keeps_the_bits(monkeypatch, walls):
    """On faces without an inlet, far-field, slip or outlet, each RK stage
    runs the update U = U_o + a Rhs, the wall-normal flux reset and
    Contra2Cart as one k-march (implicitsolver.c:32-119, rhs.c:25-144;
    CV_RK_FUSED, on by default).  Two steps with it off and on end with the
    same bits in every field, on a periodic channel and on a box of walls
    (wall-flux masks on all three axes), both around an immersed sphere."""
    kw = dict(bctype=(1, 1, 1, 1, 1, 1)) if walls else {}
    case = cases.ChannelCase(cells=(72, 20, 36), amp=0.1, beta=2.0, sphere=(0.5, 0.5, 0.5, 0.35), re=400.0,
                             seed=17, **kw)
    u0, _ = gpu_case(case, mglevels=2)
    u1, _ = gpu_case(case, mglevels=2)
    try:
        for u in (u0, u1):
            u.set_solver(rtol=1e-10, max_it=300, smoother="yline", pre_sweeps=3, post_sweeps=3, omega=0.9)
        for s in range(2):
            monkeypatch.setenv("CV_RK_FUSED", "0")
            s0 = u0.Flow_Solver()
            monkeypatch.setenv("CV_RK_FUSED", "1")
            s1 = u1.Flow_Solver()
            assert s0["ksp_its"] == s1["ksp_its"], (s, s0["ksp_its"], s1["ksp_its"])
            for f in ("ucont", "ucat", "p", "phi"):
                assert np.array_equal(u0.get(f), u1.get(f)), (s, f)
    finally:
        u0.close()
        u1.close()
