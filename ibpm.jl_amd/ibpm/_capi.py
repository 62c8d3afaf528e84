"""ctypes binding of libibpm_hip.so (declarations in include/ibpm_hip.h).

This is the Python counterpart of the `ccall` layer a Julia host would use.
The library must exist: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "csrc", "libibpm_hip.so")

IBPM_OK = 0
IBPM_E_ARG = -1
IBPM_E_HIP = -2
IBPM_E_UNSUPPORTED = -3
IBPM_E_SINGULAR = -4
IBPM_E_NONFINITE = -5
IBPM_NOALLOC_MAX_POINTS = 1 << 20  # doubles in the fixed scratch of ibpm_op_lapinv_noalloc

MOTION_STATIC = 0
MOTION_MOVING_GRID = 1
MOTION_FUNCTION = 2

FIELD_Q, FIELD_Q0, FIELD_GAMMA, FIELD_PSI, FIELD_NONLIN1, FIELD_NONLIN2 = 0, 1, 2, 3, 4, 5
FIELD_FB_TILDE, FIELD_FB, FIELD_CD, FIELD_CL = 6, 7, 8, 9

OPT_YTILE_BATCH, OPT_YTILE_SINGLE, OPT_PROFILE, OPT_GRAPHS, OPT_XWAVES_BATCH, OPT_PRECISION = 0, 1, 2, 3, 4, 5
OPT_MX_SWEEPS = 6
OPT_PIPE = 7
OPT_YSPLIT = 8
OPT_XREUSE = 9
PREC_FP64, PREC_FP32_MFMA, PREC_BF16X3_MFMA = 0, 1, 2

EXPORTED = [
    "ibpm_ctx_create", "ibpm_ctx_destroy", "ibpm_last_error", "ibpm_problem_create", "ibpm_problem_destroy",
    "ibpm_problem_linearize", "ibpm_advance_async", "ibpm_advance_wait",
    "ibpm_problem_get_info", "ibpm_problem_get_binv", "ibpm_problem_get_body_idx", "ibpm_state_upload",
    "ibpm_state_download", "ibpm_state_init", "ibpm_base_flux", "ibpm_vort2flux", "ibpm_advance",
    "ibpm_advance_n", "ibpm_get_forces", "ibpm_compute_cfl", "ibpm_op_lapinv", "ibpm_op_dst2",
    "ibpm_op_nonlinear", "ibpm_op_vort2flux", "ibpm_op_regT", "ibpm_op_reg", "ibpm_op_rot", "ibpm_set_option",
    "ibpm_get_option", "ibpm_time_kernel", "ibpm_profile_stats", "ibpm_get_body_state", "ibpm_cg_stats",
    "ibpm_op_update_bodies", "ibpm_group_create", "ibpm_group_destroy", "ibpm_group_set_option",
    "ibpm_group_advance_async", "ibpm_group_wait", "ibpm_op_lapinv_noalloc", "ibpm_tangent_n", "ibpm_adjoint_n",
    "ibpm_groups_advance_async",
]


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("mg", ctypes.c_int32), ("h", ctypes.c_double),
                ("offx", ctypes.c_double), ("offy", ctypes.c_double), ("Re", ctypes.c_double),
                ("dt", ctypes.c_double), ("beta", ctypes.c_double * 2)]


class BodyDesc(ctypes.Structure):
    _fields_ = [("nbodies", ctypes.c_int32), ("nb", ctypes.POINTER(ctypes.c_int32)),
                ("xb", ctypes.POINTER(ctypes.c_double)), ("ds", ctypes.POINTER(ctypes.c_double)),
                ("motion", ctypes.c_int32)]


class StepScalars(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("Ux", "Uy", "inclination", "U", "V", "theta", "theta_dot")] + [
        ("body_kin", ctypes.POINTER(ctypes.c_double))]


class ProblemInfo(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("mg", ctypes.c_int32), ("nbodies", ctypes.c_int32),
                ("nb_total", ctypes.c_int32), ("nbc", ctypes.c_int32), ("nG", ctypes.c_int64), ("nq", ctypes.c_int64),
                ("h", ctypes.c_double), ("offx", ctypes.c_double), ("offy", ctypes.c_double),
                ("b_asymmetry", ctypes.c_double), ("setup_seconds", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_LIB = None


class IBPMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libibpm_hip error {code}: {msg}")
        self.code = code


def lib():
    """Load libibpm_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python ibpm.jl_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        L.ibpm_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        L.ibpm_ctx_destroy.argtypes = [_vp]
        L.ibpm_last_error.argtypes = [_vp]
        L.ibpm_last_error.restype = ctypes.c_char_p
        L.ibpm_problem_create.argtypes = [_vp, ctypes.POINTER(GridDesc), ctypes.POINTER(BodyDesc), _dp,
                                          ctypes.POINTER(_vp)]
        L.ibpm_problem_destroy.argtypes = [_vp]
        L.ibpm_problem_linearize.argtypes = [_vp, _dp, _dp, _dp]
        L.ibpm_advance_async.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(StepScalars)]
        L.ibpm_advance_wait.argtypes = [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
        L.ibpm_problem_get_info.argtypes = [_vp, ctypes.POINTER(ProblemInfo)]
        L.ibpm_problem_get_binv.argtypes = [_vp, _dp, ctypes.c_size_t]
        L.ibpm_problem_get_body_idx.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t]
        L.ibpm_state_upload.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_size_t]
        L.ibpm_state_download.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_size_t]
        L.ibpm_state_init.argtypes = [_vp, ctypes.POINTER(StepScalars)]
        L.ibpm_base_flux.argtypes = [_vp, ctypes.POINTER(StepScalars)]
        L.ibpm_vort2flux.argtypes = [_vp]
        L.ibpm_advance.argtypes = [_vp, ctypes.c_double, ctypes.POINTER(StepScalars)]
        L.ibpm_advance_n.argtypes = [_vp, ctypes.c_int, _dp, ctypes.POINTER(StepScalars), _dp, _dp,
                                     ctypes.POINTER(ctypes.c_int)]
        L.ibpm_get_forces.argtypes = [_vp, _dp, _dp, _dp]
        L.ibpm_compute_cfl.argtypes = [_vp, _dp]
        L.ibpm_op_lapinv.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.ibpm_op_lapinv_noalloc.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_size_t]
        L.ibpm_op_dst2.argtypes = [_vp, _dp, _dp]
        L.ibpm_op_nonlinear.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.ibpm_op_vort2flux.argtypes = [_vp, _dp, _dp, _dp]
        L.ibpm_op_regT.argtypes = [_vp, _dp, _dp]
        L.ibpm_op_reg.argtypes = [_vp, _dp, _dp]
        L.ibpm_op_rot.argtypes = [_vp, _dp, _dp]
        L.ibpm_time_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp]
        L.ibpm_profile_stats.argtypes = [_vp, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.ibpm_get_body_state.argtypes = [_vp, _dp, _dp, ctypes.c_size_t]
        L.ibpm_cg_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), _dp, _dp]
        L.ibpm_op_update_bodies.argtypes = [_vp, _dp]
        L.ibpm_group_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]
        L.ibpm_group_destroy.argtypes = [_vp]
        L.ibpm_group_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.ibpm_group_advance_async.argtypes = [_vp, ctypes.c_int]
        L.ibpm_groups_advance_async.argtypes = [ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_int]
        L.ibpm_group_wait.argtypes = [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int32)]
        L.ibpm_tangent_n.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(StepScalars), _dp, _dp, _dp, _dp, _dp, _dp]
        L.ibpm_adjoint_n.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(StepScalars), _dp, _dp, ctypes.c_int, _dp, _dp,
                                     _dp, ctypes.POINTER(ctypes.c_int64)]
        L.ibpm_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.ibpm_get_option.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        _LIB = L
    return _LIB


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


class Context:
    """One HIP device (ibpm_ctx)."""

    _by_device: dict = {}

    def __init__(self, device: int = 0):
        h = _vp()
        rc = lib().ibpm_ctx_create(int(device), ctypes.byref(h))
        if rc != IBPM_OK:
            raise IBPMError(rc, f"ibpm_ctx_create(device={device}) failed (no HIP device?)")
        self.h = h
        self.device = device

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        if device not in cls._by_device:
            cls._by_device[device] = Context(device)
        return cls._by_device[device]

    def check(self, rc: int, what: str = ""):
        if rc != IBPM_OK:
            msg = lib().ibpm_last_error(self.h)
            raise IBPMError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc
