"""C ABI: the gfx950 library loads (no GPU needed) and exports every entry
point include/curvib_hip.h declares; the Python binding lists the same set.
No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "curvib_hip.h")
SO = os.path.join(ROOT, "curvib_old_amd", "libcurvib_hip.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|long)\s+(cv_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("cv_ctx_create", "cv_form_function1", "cv_poisson_solve", "cv_flow_step",
                 "cv_projection", "cv_divergence", "cv_update_pressure", "cv_poisson_rhs"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(SO), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(SO)
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_symbol_list_matches_header():
    from curvib_old_amd import _lib
    assert sorted(_lib.SYMBOLS) == declared()


def test_binding_constants_match_the_header():
    """Every integer #define CV_* of the header that the binding names has
    the header's value (precision modes, smoothers, error codes, ...)."""
    from curvib_old_amd import _lib
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    consts = {k: int(v) for k, v in re.findall(r"^\s*#define\s+(CV_\w+)\s+\(?(-?\d+)\)?\s*$", txt, flags=re.M)}
    shared = [k for k in consts if hasattr(_lib, k)]
    for must in ("CV_FP64", "CV_MIXED", "CV_MIXED16", "CV_OK", "CV_ERR_ARG", "CV_ERR_STATE"):
        assert must in shared, must
    bad = {k: (consts[k], getattr(_lib, k)) for k in shared if getattr(_lib, k) != consts[k]}
    assert not bad, bad


@pytest.mark.skipif(not os.path.exists(SO), reason="library not built")
def test_error_paths_without_gpu():
    """Argument errors are reported before any device call."""
    from curvib_old_amd import _lib
    L = _lib.load()
    out = ctypes.c_void_p()
    assert L.cv_ctx_create(None, 0, ctypes.byref(out)) == _lib.CV_ERR_ARG
    d = _lib.GridDesc()
    d.IM = d.JM = d.KM = 9
    for q, b in enumerate((7, 7, 7, 7, 7, 7)):       # j-periodic is rejected
        d.bctype[q] = b
    d.mglevels, d.ren, d.dt, d.st = 2, 100.0, 1e-3, 1.0
    assert L.cv_ctx_create(ctypes.byref(d), 0, ctypes.byref(out)) == _lib.CV_ERR_ARG
    assert "unsupported" in _lib.last_error()
    assert L.cv_flow_step(None, None) == _lib.CV_ERR_ARG
    o = _lib.SolverOpts()
    L.cv_default_solver(ctypes.byref(o))
    assert o.restart == 30 and o.precision == _lib.CV_FP64


@pytest.mark.skipif(not os.path.exists(SO), reason="library not built")
def test_ctx_bytes_estimate_without_gpu():
    """cv_ctx_bytes is host arithmetic: grows with the grid and the restart
    length, rejects unsupported descriptions."""
    from curvib_old_amd import cases, serving
    a = cases.ChannelCase(cells=(32, 32, 32))
    b = cases.ChannelCase(cells=(64, 32, 32))
    la = serving.levels_for(a.cells)
    assert serving.ctx_bytes(a, la, 30) > serving.ctx_bytes(a, la, 10) > 0
    assert serving.ctx_bytes(b, la, 30) > serving.ctx_bytes(a, la, 30)
    # one Krylov vector more costs 2 fine-level vectors (V and Z)
    n = 34 * 34 * 34
    assert serving.ctx_bytes(a, la, 11) - serving.ctx_bytes(a, la, 10) == 2 * 8 * n
    bad = cases.ChannelCase(cells=(32, 32, 32), bctype=(7, 7, 7, 7, 7, 7))
    assert serving.ctx_bytes(bad, la, 30) == -1


def test_levels_for_semi_coarsening():
    from curvib_old_amd import serving
    from curvib_old_amd.solver import level_cells
    fixed = dict(coarsen_auto=0)
    assert serving.levels_for((32, 32, 32), **fixed) == 4          # x, z: 32 -> 4 cells
    assert serving.levels_for((512, 256, 256), **fixed) == 7
    assert serving.levels_for((96, 40, 96), **fixed) == 5
    # auto coarsening halves only the longest coarsenable axes first
    h = level_cells((512, 256, 256), (0, 1, 0), 8, 1)
    assert h[-1] == (512, 256, 256) and h[-2] == (256, 256, 256) and h[0] == (4, 256, 4)
    assert serving.levels_for((512, 256, 256), coarsen_auto=1) == 8


def _c_layout(tmp_path):
    """sizeof / offsetof of the ABI structs as a C compiler lays them out."""
    src = tmp_path / "layout.c"
    fields = {
        "cv_grid_desc": ["IM", "bctype", "mglevels", "ren", "coarsen_auto", "inletprofile", "inlet_center"],
        "cv_step_stats": ["ksp_its", "maxdiv_idx", "t_mom", "imp_its", "imp_reldU", "flux_in", "flux_out"],
        "cv_solver_opts": ["precision", "rtol", "smoother", "omega"],
        "cv_momentum_opts": ["scheme", "imp_atol", "cfl"],
        "cv_ib_node": ["i", "k3", "cr1", "d_i", "ua", "dun", "nf"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'  printf("{st} size %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'  printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("  return 0;\n}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return fields, {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


def test_binding_structs_match_the_c_layout(tmp_path):
    """The ctypes mirrors of the ABI structs have the C compiler's sizes and
    field offsets."""
    from curvib_old_amd import _lib
    fields, lay = _c_layout(tmp_path)
    py = {"cv_grid_desc": _lib.GridDesc, "cv_step_stats": _lib.StepStats, "cv_solver_opts": _lib.SolverOpts,
          "cv_momentum_opts": _lib.MomentumOpts, "cv_ib_node": _lib.IbNode}
    for st, fs in fields.items():
        cls = py[st]
        assert ctypes.sizeof(cls) == lay[(st, "size")], st
        for f in fs:
            assert getattr(cls, f).offset == lay[(st, f)], (st, f)
    # the numpy record the Python mirror hands to cv_set_ib_nodes
    assert _lib.IB_NODE_DTYPE.itemsize == lay[("cv_ib_node", "size")]
    assert _lib.IB_NODE_DTYPE.fields["val"][1] == lay[("cv_ib_node", "cr1")]


@pytest.mark.skipif(not os.path.exists(SO), reason="library not built")
def test_inlet_outlet_descriptions_without_gpu():
    """Inlet (k-, bctype 5), outlet (k+, bctype 4), slip (3) and far field
    (6) faces are accepted with an analytic inlet profile; file-based
    profiles, inlets and outlets on other faces, slip and far field on the
    j- face (whose eta the reference's FormMetrics leaves at 0) and unpaired
    periodic faces are rejected before any device call."""
    from curvib_old_amd import _lib, cases, serving
    la = 3
    walls = cases.ChannelCase(cells=(16, 16, 16), bctype=(7, 7, 1, 1, 1, 1))
    # Ubcs of the six boundary planes: 6 * (my*mz + mx*mz + mx*my) doubles
    for bc in ((7, 7, 1, 1, 5, 4), (3, 6, 1, 3, 6, 3), (7, 7, 1, 6, 5, 4)):
        ok = cases.ChannelCase(cells=(16, 16, 16), bctype=bc, inlet=7)
        assert serving.ctx_bytes(ok, la, 30) - serving.ctx_bytes(walls, la, 30) == 6 * 3 * 18 * 18 * 8, bc
    for bad in (dict(inlet=0), dict(inlet=3), dict(bctype=(7, 7, 1, 1, 4, 5)), dict(bctype=(5, 4, 1, 1, 1, 1)),
                dict(bctype=(7, 7, 1, 1, 5, 7)), dict(bctype=(7, 7, 3, 1, 1, 1)), dict(bctype=(7, 7, 6, 1, 1, 1)),
                dict(bctype=(7, 7, 1, 2, 1, 1))):
        kw = dict(cells=(16, 16, 16), bctype=(7, 7, 1, 1, 5, 4), inlet=7)
        kw.update(bad)
        assert serving.ctx_bytes(cases.ChannelCase(**kw), la, 30) == -1, bad
    L = _lib.load()
    d = serving.grid_desc(cases.ChannelCase(cells=(16, 16, 16), bctype=(7, 7, 1, 1, 5, 4), inlet=0), 3)
    out = ctypes.c_void_p()
    assert L.cv_ctx_create(ctypes.byref(d), 0, ctypes.byref(out)) == _lib.CV_ERR_ARG
    assert "inlet profile" in _lib.last_error()
