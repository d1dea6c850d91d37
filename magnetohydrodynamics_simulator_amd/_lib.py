"""ctypes binding of libmhd_mi355x.so (the C ABI in include/mhd_mi355x.h).

There is no fallback: if the gfx950 library is missing or cannot be loaded,
load() raises.  Build it with ``python -m magnetohydrodynamics_simulator_amd._build``.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmhd_mi355x.so")

MHD_OK, MHD_EINVAL, MHD_EHIP, MHD_ENONFINITE, MHD_ECG, MHD_ENOMEM = range(6)
ERROR_NAMES = {0: "MHD_OK", 1: "MHD_EINVAL", 2: "MHD_EHIP", 3: "MHD_ENONFINITE", 4: "MHD_ECG", 5: "MHD_ENOMEM"}

KERNEL_IDS = {"stage1": 0, "stage2": 1, "cg_start": 2, "cg_pass_a": 3, "cg_pass_b": 4,
              "implicit_finalize1": 5, "implicit_finalize2": 6, "cfl": 7, "other": 8}


class Opts(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("cg_tol", C.c_double),
                ("cg_max_iter", C.c_longlong), ("cg_fail_is_error", C.c_int), ("march_rows", C.c_int),
                ("bc_x", C.c_int), ("bc_y", C.c_int)]


BC_PERIODIC, BC_OUTFLOW = 0, 1
BC_NAMES = {"periodic": BC_PERIODIC, "outflow": BC_OUTFLOW}


class StepStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int * 2), ("cg_spmv", C.c_int * 2), ("cg_error", C.c_double * 2),
                ("dt", C.c_double)]


# (restype, argtypes) of every exported symbol of include/mhd_mi355x.h
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
# mhd_snapshot_fn: int (*)(mhd_ctx*, int frame, double t, double dt, void* user)
SNAPSHOT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p)
SIGNATURES = {
    "mhd_default_opts": (C.c_int, [C.POINTER(Opts)]),
    "mhd_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Opts), C.POINTER(_vp)]),
    "mhd_destroy": (C.c_int, [_vp]),
    "mhd_last_error": (C.c_char_p, [_vp]),
    "mhd_upload": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mhd_download": (C.c_int, [_vp] + [_dp] * 11),
    "mhd_cfl_dt": (C.c_int, [_vp, _dp]),
    "mhd_build_operator": (C.c_int, [_vp, C.c_double]),
    "mhd_imex_step": (C.c_int, [_vp, C.c_double, C.POINTER(StepStats)]),
    "mhd_explicit_rhs": (C.c_int, [_vp] + [_dp] * 6),
    "mhd_resistive_term": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip]),
    "mhd_cg_solve": (C.c_int, [_vp, _dp, _dp, _ip, _dp]),
    "mhd_divb": (C.c_int, [_vp, _dp, _dp]),
    "mhd_run": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, _ip, _dp, _dp, _ip]),
    "mhd_run_resident": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_int, _vp, _vp, _ip, _dp, _dp, _ip]),
    "mhd_run_resident_files": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_char_p, C.c_int, _ip,
                                         _dp, _dp, _ip]),
    "mhd_resident_tune": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mhd_init_fourier": (C.c_int, [_vp, C.c_int, _dp]),
    "mhd_init_column": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "mhd_batch_init_fourier": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "mhd_batch_init_column": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "mhd_init_setup": (C.c_int, [_vp, C.c_int]),
    "mhd_batch_init_setup": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mhd_profile_enable": (C.c_int, [_vp, C.c_int]),
    "mhd_profile_get": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_longlong), _dp]),
    "mhd_calibration_copy": (C.c_int, [_vp, C.c_longlong]),
    "mhd_calibration_mix": (C.c_int, [_vp, C.c_int, C.c_int, C.c_longlong, C.c_int, _dp]),
    "mhd_synchronize": (C.c_int, [_vp]),
    "mhd_adjoint_forward": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "mhd_adjoint_backward": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(_dp), _dp]),
    "mhd_adjoint_backward_device": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(_dp), _dp]),
    "mhd_upload_device": (C.c_int, [_vp, C.POINTER(_dp)]),
    "mhd_write_snapshot": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_double]),
    "mhd_write_restart": (C.c_int, [_vp, C.c_char_p, C.c_double]),
    "mhd_read_restart": (C.c_int, [_vp, C.c_char_p, _dp]),
    "mhd_download_device": (C.c_int, [_vp, C.POINTER(_dp)]),
    "mhd_cfl_dt_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.POINTER(_dp), _dp, _vp]),
    "mhd_adjoint_record_branches": (C.c_int, [_vp, C.c_int]),
    "mhd_adjoint_use_graph": (C.c_int, [_vp, C.c_int]),
    "mhd_adjoint_forward_frozen": (C.c_int, [_vp, C.c_int, C.c_double]),
    "mhd_explicit_rhs_vjp": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(_dp)]),
    "mhd_explicit_rhs_vjp_unfused": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(_dp)]),
    "mhd_batch_create": (C.c_int, [C.c_int, _ip, _ip, _dp, _dp, _dp, C.POINTER(Opts), C.POINTER(_vp)]),
    "mhd_batch_destroy": (C.c_int, [_vp]),
    "mhd_batch_last_error": (C.c_char_p, [_vp]),
    "mhd_batch_set_schedule": (C.c_int, [_vp, _dp, _ip, _dp]),
    "mhd_batch_upload": (C.c_int, [_vp, C.c_int] + [_dp] * 6),
    "mhd_batch_download": (C.c_int, [_vp, C.c_int] + [_dp] * 6),
    "mhd_batch_step": (C.c_int, [_vp, _ip]),
    "mhd_batch_run": (C.c_int, [_vp, C.c_int, _ip]),
    "mhd_batch_run_info": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _ip]),
    "mhd_batch_autotune": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "mhd_batch_tune": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mhd_batch_get_tuning": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhd_batch_synchronize": (C.c_int, [_vp]),
    "mhd_batch_adjoint_forward": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mhd_batch_adjoint_set_lambda": (C.c_int, [_vp, C.c_int, C.POINTER(_dp)]),
    "mhd_batch_adjoint_backward": (C.c_int, [_vp]),
    "mhd_batch_adjoint_result": (C.c_int, [_vp, C.c_int, C.POINTER(_dp), _dp]),
}

_LIB: C.CDLL | None = None


class MhdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def load(path: str | None = None) -> C.CDLL:
    """Load the library once; raises if it is absent (no CPU fallback)."""
    global _LIB
    if _LIB is None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise ImportError(f"{p} is missing: build it with python -m magnetohydrodynamics_simulator_amd._build")
        lib = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(rc: int, ctx=None) -> None:
    if rc != MHD_OK:
        lib = load()
        msg = lib.mhd_last_error(ctx)
        raise MhdError(rc, msg.decode() if msg else "")


def check_batch(rc: int, b=None) -> None:
    if rc != MHD_OK:
        lib = load()
        msg = lib.mhd_batch_last_error(b)
        raise MhdError(rc, msg.decode() if msg else "")
