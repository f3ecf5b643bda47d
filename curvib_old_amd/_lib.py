"""ctypes binding of libcurvib_hip.so (the gfx950 C-ABI declared in
include/curvib_hip.h).  There is no CPU fallback: if the library is missing
the import of this module raises."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(_HERE, "libcurvib_hip.so")

CV_OK = 0
CV_ERR_ARG = -1
CV_ERR_OOM = -2
CV_ERR_NOT_CONVERGED = -3
CV_ERR_NONFINITE = -4
CV_ERR_HIP = -5
CV_ERR_STATE = -6
CV_FP64 = 0
CV_MIXED = 1
CV_MIXED16 = 2
CV_SMOOTH_JACOBI = 0
CV_SMOOTH_RB = 1
CV_SMOOTH_YLINE = 2

FIELDS = ["ucont", "ucat", "p", "phi", "rhs", "b", "nvert", "aj", "csi", "eta", "zet",
          "icsi", "ieta", "izet", "iaj", "jcsi", "jeta", "jzet", "jaj",
          "kcsi", "keta", "kzet", "kaj", "ucont_o"]
FIELD_ID = {n: q for q, n in enumerate(FIELDS)}
VEC_FIELDS = {"ucont", "ucat", "rhs", "csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta",
              "jzet", "kcsi", "keta", "kzet", "ucont_o"}

# exported symbols (must match include/curvib_hip.h)
SYMBOLS = [
    "cv_ctx_create", "cv_ctx_destroy", "cv_set_solver", "cv_default_solver", "cv_last_error",
    "cv_device_bytes", "cv_set_geometry", "cv_set_nvert", "cv_get_field", "cv_put_field",
    "cv_contra2cart", "cv_form_bcs", "cv_form_function1", "cv_rungekutta", "cv_volumeflux_rev",
    "cv_poisson_rhs", "cv_poisson_apply", "cv_poisson_solve", "cv_update_pressure",
    "cv_projection", "cv_divergence", "cv_flow_step", "cv_mg_restrict", "cv_mg_prolong",
    "cv_nullspace_project", "cv_prof_select", "cv_prof_read",
    "cv_set_momentum", "cv_default_momentum", "cv_imprk", "cv_calcrhs",
    "cv_ctx_bytes", "cv_server_create", "cv_server_submit", "cv_server_wait", "cv_server_step_ms",
    "cv_server_fetch", "cv_server_release", "cv_server_status", "cv_server_destroy",
    "cv_batch_create", "cv_batch_step", "cv_batch_poisson_solve", "cv_batch_destroy", "cv_server_create_mode",
    "cv_inflow_flux", "cv_outflow_flux", "cv_set_ib_nodes", "cv_ibm_interpolation", "cv_server_submit_ib",
    "cv_divergence_noalloc",
]
CV_SERVE_STREAMS, CV_SERVE_BATCHED = 0, 1
CV_CASE_QUEUED, CV_CASE_RUNNING, CV_CASE_DONE, CV_CASE_FAILED = 0, 1, 2, 3

CV_PROF = {"none": 0, "resid": 1, "apply": 2, "line": 3, "rct": 4, "ff1": 5, "orth": 6}


class GridDesc(C.Structure):
    _fields_ = [("IM", C.c_int), ("JM", C.c_int), ("KM", C.c_int), ("bctype", C.c_int * 6),
                ("mglevels", C.c_int), ("isc", C.c_int), ("jsc", C.c_int), ("ksc", C.c_int),
                ("ren", C.c_double), ("dt", C.c_double), ("st", C.c_double), ("coarsen_auto", C.c_int),
                ("inletprofile", C.c_int), ("inlet_center", C.c_double * 3)]


class SolverOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("restart", C.c_int), ("max_it", C.c_int),
                ("rtol", C.c_double), ("atol", C.c_double), ("pre_sweeps", C.c_int),
                ("post_sweeps", C.c_int), ("coarse_sweeps", C.c_int), ("smoother", C.c_int),
                ("omega", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("ksp_its", C.c_int), ("converged", C.c_int), ("ksp_rnorm", C.c_double),
                ("ksp_bnorm", C.c_double), ("rhs_sum", C.c_double), ("maxdiv", C.c_double),
                ("maxdiv_idx", C.c_long), ("ibm_flux0", C.c_double), ("ibm_area0", C.c_double),
                ("ibm_flux1", C.c_double), ("ibm_area1", C.c_double), ("t_mom", C.c_double),
                ("t_p", C.c_double), ("t_pr", C.c_double), ("imp_its", C.c_int), ("imp_lambda", C.c_double),
                ("imp_normdU", C.c_double), ("imp_reldU", C.c_double), ("flux_in", C.c_double),
                ("flux_out", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


CV_MOM_RK, CV_MOM_IMPRK = 0, 1


class IbNode(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("i1", C.c_int), ("j1", C.c_int), ("k1", C.c_int),
                ("i2", C.c_int), ("j2", C.c_int), ("k2", C.c_int), ("i3", C.c_int), ("j3", C.c_int), ("k3", C.c_int),
                ("cr1", C.c_double), ("cr2", C.c_double), ("cr3", C.c_double), ("d_s", C.c_double),
                ("d_i", C.c_double), ("ua", C.c_double * 3), ("dun", C.c_double), ("nf", C.c_double * 3)]


# the same record as a numpy dtype: 12 ints, then the 12 doubles in field order
IB_NODE_DTYPE = np.dtype([("idx", np.int32, 12), ("val", np.float64, 12)])
CV_IB_NOSLIP, CV_IB_FREESLIP = 0, 1


class MomentumOpts(C.Structure):
    _fields_ = [("scheme", C.c_int), ("imp_max_it", C.c_int), ("imp_atol", C.c_double), ("imp_rtol", C.c_double),
                ("cfl", C.c_double)]


class CaseInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("rc", C.c_int), ("steps_done", C.c_int), ("bytes", C.c_long),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("last", StepStats)]


class ServerInfo(C.Structure):
    _fields_ = [("queued", C.c_long), ("running", C.c_long), ("done", C.c_long), ("bytes_in_use", C.c_long),
                ("bytes_peak", C.c_long), ("budget", C.c_long), ("time", C.c_double)]


_lib = None


def load():
    """Load the library (raises FileNotFoundError / OSError if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO_PATH):
        raise FileNotFoundError(
            f"{SO_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(SO_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.cv_ctx_create.argtypes = [C.POINTER(GridDesc), C.c_int, C.POINTER(C.c_void_p)]
    L.cv_ctx_destroy.argtypes = [vp]
    L.cv_set_solver.argtypes = [vp, C.POINTER(SolverOpts)]
    L.cv_default_solver.argtypes = [C.POINTER(SolverOpts)]
    L.cv_default_solver.restype = None
    L.cv_last_error.argtypes = [C.c_char_p, C.c_int]
    L.cv_device_bytes.argtypes = [vp]
    L.cv_device_bytes.restype = C.c_long
    L.cv_set_geometry.argtypes = [vp, dp]
    L.cv_set_nvert.argtypes = [vp, dp]
    L.cv_get_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.cv_put_field.argtypes = [vp, C.c_int, C.c_int, dp]
    for nm in ("cv_contra2cart", "cv_form_bcs", "cv_rungekutta", "cv_update_pressure", "cv_projection"):
        getattr(L, nm).argtypes = [vp]
    L.cv_form_function1.argtypes = [vp, dp]
    L.cv_inflow_flux.argtypes = [vp, dp]
    L.cv_set_ib_nodes.argtypes = [vp, C.c_void_p, C.c_int, C.c_int]
    L.cv_ibm_interpolation.argtypes = [vp]
    L.cv_outflow_flux.argtypes = [vp, dp]
    L.cv_volumeflux_rev.argtypes = [vp, dp]
    L.cv_poisson_rhs.argtypes = [vp, dp]
    L.cv_poisson_apply.argtypes = [vp, C.c_int, dp, dp]
    L.cv_poisson_solve.argtypes = [vp, C.POINTER(StepStats)]
    L.cv_flow_step.argtypes = [vp, C.POINTER(StepStats)]
    L.cv_divergence.argtypes = [vp, dp, C.POINTER(C.c_long)]
    L.cv_divergence_noalloc.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp]
    L.cv_mg_restrict.argtypes = [vp, C.c_int, dp, dp]
    L.cv_mg_prolong.argtypes = [vp, C.c_int, dp, dp]
    L.cv_nullspace_project.argtypes = [vp, C.c_int, dp]
    L.cv_prof_select.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.cv_prof_read.argtypes = [vp, C.POINTER(C.c_long), dp, dp, C.c_int]
    L.cv_set_momentum.argtypes = [vp, C.POINTER(MomentumOpts)]
    L.cv_default_momentum.argtypes = [C.POINTER(MomentumOpts)]
    L.cv_default_momentum.restype = None
    L.cv_imprk.argtypes = [vp, C.POINTER(StepStats)]
    L.cv_calcrhs.argtypes = [vp, dp]
    L.cv_ctx_bytes.argtypes = [C.POINTER(GridDesc), C.c_int]
    L.cv_ctx_bytes.restype = C.c_long
    L.cv_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
    L.cv_batch_step.argtypes = [vp, C.POINTER(StepStats)]
    L.cv_batch_poisson_solve.argtypes = [vp, C.POINTER(StepStats)]
    L.cv_batch_destroy.argtypes = [vp]
    L.cv_server_create.argtypes = [C.c_int, C.c_long, C.c_int, C.POINTER(C.c_void_p)]
    L.cv_server_create_mode.argtypes = [C.c_int, C.c_long, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.cv_server_submit.argtypes = [vp, C.POINTER(GridDesc), C.POINTER(SolverOpts), dp, dp, dp, C.c_int,
                                   C.POINTER(C.c_long)]
    L.cv_server_submit_ib.argtypes = [vp, C.POINTER(GridDesc), C.POINTER(SolverOpts), dp, dp, dp, C.c_int, C.c_void_p,
                                      C.c_int, C.c_int, C.POINTER(C.c_long)]
    L.cv_server_wait.argtypes = [vp, C.c_long, C.POINTER(CaseInfo)]
    L.cv_server_step_ms.argtypes = [vp, C.c_long, dp, C.POINTER(C.c_int), C.c_int]
    L.cv_server_fetch.argtypes = [vp, C.c_long, C.c_int, dp]
    L.cv_server_release.argtypes = [vp, C.c_long]
    L.cv_server_status.argtypes = [vp, C.POINTER(ServerInfo)]
    L.cv_server_destroy.argtypes = [vp]
    for s in SYMBOLS:
        f = getattr(L, s)
        if s not in ("cv_default_solver", "cv_default_momentum", "cv_device_bytes", "cv_last_error",
                     "cv_ctx_bytes"):
            f.restype = C.c_int
    L.cv_last_error.restype = C.c_int
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(512)
    load().cv_last_error(buf, 512)
    return buf.value.decode(errors="replace")


class CvError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where} failed with status {code}: {last_error()}")
        self.code = code


def check(code, where):
    if code != CV_OK:
        raise CvError(code, where)
    return code
