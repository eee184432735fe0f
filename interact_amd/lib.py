"""ctypes bindings of librsfg (include/rsfg.h).

The shared library is built in-tree by ``interact_amd.build``; importing a
function that needs it fails loudly when it is missing -- there is no CPU
fallback in the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librsfg.so")

RSFG_OK = 0
RSFG_EINVAL, RSFG_ENOMEM, RSFG_EHIP, RSFG_ESTATE, RSFG_ESTORM, RSFG_EDOMAIN = -1, -2, -3, -4, -5, -6
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESTATE", -5: "ESTORM",
          -6: "EDOMAIN"}

SLIP_STRIKE, SLIP_DIP, SLIP_RAKE, SLIP_VECTOR = 0, 1, 3, 8
PREC_FP64, PREC_FP32, PREC_BF16X2, PREC_BF16, PREC_BF16NF = 1, 2, 4, 8, 16
PREC_NAMES = {PREC_FP64: "fp64", PREC_FP32: "fp32", PREC_BF16X2: "bf16x2",
              PREC_BF16: "bf16", PREC_BF16NF: "bf16nf"}
PREC_BY_NAME = {v: k for k, v in PREC_NAMES.items()}
# bytes per entry; bf16nf: 2 + 2 x (fraction of chunks with lo), see FaultSystem.nf_fraction
PREC_BYTES = {PREC_FP64: 8, PREC_FP32: 4, PREC_BF16X2: 4, PREC_BF16: 2, PREC_BF16NF: 2}
RK5DP, RK3BS, ARKIMEX_L2 = 0, 1, 2

# every symbol include/rsfg.h declares
EXPORTS = [
    "rsfg_last_error", "rsfg_version", "rsfg_device_count", "rsfg_create",
    "rsfg_destroy", "rsfg_read_patch_file", "rsfg_assemble",
    "rsfg_operator_dims", "rsfg_operator_bytes", "rsfg_entries", "rsfg_traction_pairs",
    "rsfg_set_triangle_tolerance",
    "rsfg_matvec", "rsfg_matvec_dev", "rsfg_matvec_dev_noalloc", "rsfg_matvec_t",
    "rsfg_matvec_bench", "rsfg_read_bench",
    "rsfg_set_friction", "rsfg_set_loading", "rsfg_rhs",
    "rsfg_solver_defaults", "rsfg_set_state", "rsfg_get_state", "rsfg_solve",
    "rsfg_solve_continue",
    "rsfg_events", "rsfg_monitor", "rsfg_catalog", "rsfg_stage_bench",
    "rsfg_read_table", "rsfg_write_checkpoint", "rsfg_read_checkpoint",
    "rsfg_create_member", "rsfg_ensemble_create", "rsfg_ensemble_member",
    "rsfg_ensemble_solve", "rsfg_ensemble_tune", "rsfg_ensemble_config", "rsfg_ensemble_select",
    "rsfg_ensemble_destroy", "rsfg_ensemble_bench", "rsfg_ensemble_set_operator", "rsfg_set_step_log", "rsfg_get_step_log", "rsfg_replay",
    "rsfg_adjoint_gradient", "rsfg_adjoint_bench",
]


class Patch(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("strike", C.c_float), ("dip", C.c_float),
                ("l", C.c_double), ("w", C.c_double),
                ("rake", C.c_double * 2), ("group", C.c_int),
                ("vert", C.c_double * 9)]


class Assembly(C.Structure):
    _fields_ = [("slip_mode", C.c_int), ("with_normal", C.c_int),
                ("full_space", C.c_int), ("store", C.c_int),
                ("scale", C.c_double), ("nf_tol", C.c_double)]


class Friction(C.Structure):
    _fields_ = [("law", C.c_int), ("v0", C.c_double), ("f0", C.c_double),
                ("eta", C.c_double), ("vmin_state", C.c_double),
                ("sato_beta", C.c_double), ("kt_vc", C.c_double),
                ("calc_sigma_dot", C.c_int), ("limit_sigma", C.c_int),
                ("min_sigma", C.c_double), ("max_sigma", C.c_double)]


class SolverOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("prec", C.c_int), ("rtol", C.c_double),
                ("atol", C.c_double * 6), ("dt_init", C.c_double),
                ("dt_max", C.c_double), ("dt_monitor", C.c_double),
                ("rdx_monitor", C.c_double), ("adx_monitor", C.c_double),
                ("vel_event", C.c_double), ("vel_event_hyst", C.c_double),
                ("rupture_vth", C.c_double), ("max_steps", C.c_long),
                ("max_domain_reject", C.c_int),
                ("steps_per_launch", C.c_int), ("max_events", C.c_int),
                ("max_monitor", C.c_int)]


class AdjointProblem(C.Structure):
    _fields_ = [("method", C.c_int), ("prec", C.c_int), ("checkpoint_every", C.c_int),
                ("nrecv", C.c_int), ("recv", C.c_void_p), ("target", C.c_void_p),
                ("weight", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_long), ("rejects", C.c_long),
                ("domain_rejects", C.c_long), ("rhs_evals", C.c_long),
                ("launches", C.c_long), ("t", C.c_double), ("dt", C.c_double),
                ("status", C.c_int), ("nevent", C.c_int),
                ("nmonitor", C.c_int), ("wall_s", C.c_double),
                ("events_dropped", C.c_int), ("monitor_dropped", C.c_int),
                ("ncatalog", C.c_int), ("catalog_dropped", C.c_int)]


_LIB = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class RsfgError(RuntimeError):
    pass


def load(path=None):
    """Load librsfg.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RsfgError(
            f"{path} is missing: build it with `python -m interact_amd.build`"
            " (the HIP path has no CPU fallback)")
    L = C.CDLL(path)
    h = C.c_void_p
    L.rsfg_last_error.restype = C.c_char_p
    L.rsfg_version.restype = C.c_char_p
    L.rsfg_device_count.argtypes = [C.POINTER(C.c_int)]
    L.rsfg_create.argtypes = [C.c_int, C.POINTER(Patch), C.c_int,
                              C.POINTER(C.c_void_p)]
    L.rsfg_destroy.argtypes = [h]
    L.rsfg_create_member.argtypes = [h, C.POINTER(C.c_void_p)]
    L.rsfg_ensemble_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p)]
    L.rsfg_ensemble_member.argtypes = [h, C.c_int, C.POINTER(C.c_void_p)]
    L.rsfg_ensemble_solve.argtypes = [h, C.POINTER(SolverOpts), C.c_double,
                                      C.POINTER(Stats)]
    L.rsfg_ensemble_tune.argtypes = [h, C.POINTER(SolverOpts), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
    L.rsfg_ensemble_config.argtypes = [h, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
    L.rsfg_ensemble_select.argtypes = [h, C.c_int]
    L.rsfg_ensemble_set_operator.argtypes = [h, C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.rsfg_ensemble_destroy.argtypes = [h]
    L.rsfg_ensemble_bench.argtypes = [h, C.POINTER(SolverOpts), C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    L.rsfg_set_step_log.argtypes = [h, C.c_longlong]
    L.rsfg_get_step_log.argtypes = [h, C.c_void_p, C.c_longlong, C.POINTER(C.c_longlong)]
    L.rsfg_replay.argtypes = [h, _dp, C.c_longlong, C.POINTER(AdjointProblem), C.c_void_p,
                              C.c_void_p, C.c_void_p]
    L.rsfg_adjoint_gradient.argtypes = [h, _dp, C.c_longlong, C.POINTER(AdjointProblem),
                                        C.POINTER(C.c_double), _dp, _dp, _dp]
    L.rsfg_adjoint_bench.argtypes = [h, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.rsfg_read_patch_file.argtypes = [C.c_char_p, C.POINTER(Patch),
                                       C.POINTER(C.c_int)]
    L.rsfg_assemble.argtypes = [h, C.POINTER(Assembly), C.POINTER(C.c_double)]
    L.rsfg_operator_bytes.argtypes = [h, C.c_int, C.POINTER(C.c_double)]
    L.rsfg_operator_dims.argtypes = [h, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.rsfg_entries.argtypes = [h, C.c_int, _ip, C.c_int, _ip, _dp]
    L.rsfg_traction_pairs.argtypes = [h, C.c_long, _ip, _ip, C.c_int, _dp]
    L.rsfg_set_triangle_tolerance.argtypes = [h, C.c_double]
    L.rsfg_read_bench.argtypes = [h, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.rsfg_matvec.argtypes = [h, C.c_int, _dp, _dp]
    L.rsfg_matvec_dev.argtypes = [h, C.c_int, C.c_void_p, C.c_void_p]
    L.rsfg_matvec_dev_noalloc.argtypes = [h, C.c_int, C.c_void_p, C.c_void_p]
    L.rsfg_matvec_t.argtypes = [h, C.c_int, _dp, _dp]
    L.rsfg_matvec_bench.argtypes = [h, C.c_int, C.c_int,
                                    C.POINTER(C.c_double)]
    L.rsfg_set_friction.argtypes = [h, C.POINTER(Friction), _dp, _dp, _dp]
    L.rsfg_set_loading.argtypes = [h, C.c_double, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
    L.rsfg_rhs.argtypes = [h, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
    L.rsfg_solver_defaults.argtypes = [C.POINTER(SolverOpts)]
    L.rsfg_set_state.argtypes = [h, C.c_double, _dp]
    L.rsfg_get_state.argtypes = [h, C.POINTER(C.c_double), _dp]
    L.rsfg_read_table.argtypes = [C.c_char_p, C.c_int, C.c_int, _dp]
    L.rsfg_write_checkpoint.argtypes = [h, C.c_char_p, C.c_int, C.c_int, C.c_long,
                                        C.c_double]
    L.rsfg_read_checkpoint.argtypes = [h, C.c_char_p, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_long),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.rsfg_solve.argtypes = [h, C.POINTER(SolverOpts), C.c_double,
                             C.POINTER(Stats)]
    L.rsfg_solve_continue.argtypes = [h, C.POINTER(SolverOpts), C.c_double,
                                      C.POINTER(Stats)]
    L.rsfg_events.argtypes = [h, _dp, C.POINTER(C.c_int)]
    L.rsfg_monitor.argtypes = [h, _dp, C.POINTER(C.c_int)]
    L.rsfg_catalog.argtypes = [h, _dp, C.POINTER(C.c_int)]
    L.rsfg_stage_bench.argtypes = [h, C.c_int, C.c_int, C.POINTER(C.c_double)]
    for name in EXPORTS:
        if name not in ("rsfg_last_error", "rsfg_version"):
            getattr(L, name).restype = C.c_int
    _LIB = L
    return L


def check(rc, what=""):
    if rc != RSFG_OK:
        msg = load().rsfg_last_error().decode(errors="replace")
        raise RsfgError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def device_count():
    c = C.c_int(0)
    load().rsfg_device_count(C.byref(c))
    return c.value


def read_patch_file(fname):
    """(ctypes Patch array, count) of a reference geometry file: counts the
    patches, then reads them (rsfg_read_patch_file)."""
    n = C.c_int(0)
    check(load().rsfg_read_patch_file(os.fsencode(fname), None, C.byref(n)),
          "rsfg_read_patch_file")
    arr = (Patch * max(n.value, 1))()
    check(load().rsfg_read_patch_file(os.fsencode(fname), arr, C.byref(n)),
          "rsfg_read_patch_file")
    return arr, n.value


def patches_from_arrays(xyz, strike, dip, l, w, rake=None, group=None, vert=None):
    """ctypes patch array; l, w < 0 with vert[i] (9 coordinates) marks a
    triangle."""
    n = len(l)
    arr = (Patch * n)()
    for i in range(n):
        p = arr[i]
        p.x, p.y, p.z = (float(v) for v in xyz[i])
        p.strike = float(strike[i])
        p.dip = float(dip[i])
        p.l = float(l[i])
        p.w = float(w[i])
        if rake is not None:
            p.rake[0], p.rake[1] = float(rake[i][0]), float(rake[i][1])
        else:
            p.rake[0], p.rake[1] = 1.0, 0.0
        p.group = int(group[i]) if group is not None else 0
        if vert is not None:
            for k in range(9):
                p.vert[k] = float(vert[i][k])
    return arr
