"""ctypes binding of libseiscl_amd.so (the C-ABI declared in include/seiscl_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C seiscl_amd/csrc``).  Loading fails loudly when it is missing: there
is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseiscl_amd.so")
MAXL = 8


class SeisCsts(ctypes.Structure):
    """Mirror of ``seis_csts`` (include/seiscl_amd.h)."""

    _fields_ = [
        ("ND", ctypes.c_int),
        ("N", ctypes.c_int * 3),
        ("NT", ctypes.c_int),
        ("dt", ctypes.c_float),
        ("dh", ctypes.c_float),
        ("FDORDER", ctypes.c_int),
        ("MAXRELERROR", ctypes.c_int),
        ("freesurf", ctypes.c_int),
        ("abs_type", ctypes.c_int),
        ("nab", ctypes.c_int),
        ("abpc", ctypes.c_float),
        ("VPPML", ctypes.c_float),
        ("NPOWER", ctypes.c_float),
        ("FPML", ctypes.c_float),
        ("K_MAX_CPML", ctypes.c_float),
        ("L", ctypes.c_int),
        ("f0", ctypes.c_float),
        ("FL", ctypes.c_float * MAXL),
        ("seisout", ctypes.c_int),
        ("param_type", ctypes.c_int),
        ("tmax", ctypes.c_int),
        ("tmin", ctypes.c_int),
        ("gradout", ctypes.c_int),
        ("FP16", ctypes.c_int),
        ("gradfix", ctypes.c_int),
        ("restype", ctypes.c_int),
        ("scalerms", ctypes.c_int),
        ("scaleshot", ctypes.c_int),
        ("scalermsnorm", ctypes.c_int),
        ("fmin", ctypes.c_float),
        ("fmax", ctypes.c_float),
        ("Hout", ctypes.c_int),
    ]


# (name, restype, argtypes) of every exported entry point.
_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_I = ctypes.c_int
SIGNATURES = [
    ("seis_last_error", ctypes.c_char_p, []),
    ("seis_csts_default", None, [ctypes.POINTER(SeisCsts)]),
    ("seis_csts_size", ctypes.c_size_t, []),
    ("seis_ctx_create", _I, [_I, ctypes.POINTER(_P)]),
    ("seis_ctx_destroy", _I, [_P]),
    ("seis_ctx_sync", _I, [_P]),
    ("seis_ctx_footprint", ctypes.c_size_t, [_P]),
    ("seis_buf_create", _I, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    ("seis_buf_free", _I, [_P, _P]),
    ("seis_buf_send", _I, [_P, _P, _P, ctypes.c_size_t]),
    ("seis_buf_read", _I, [_P, _P, _P, ctypes.c_size_t]),
    ("seis_model_create", _I, [_P, ctypes.POINTER(SeisCsts), _FP, _FP, _FP, _FP, _FP,
                               ctypes.POINTER(_P)]),
    ("seis_model_destroy", _I, [_P]),
    ("seis_model_set_params", _I, [_P, _FP, _FP, _FP, _FP, _FP]),
    ("seis_model_set_params_device", _I, [_P, _P, _P, _P, _P, _P]),
    ("seis_model_footprint", ctypes.c_size_t, [_P]),
    ("seis_out_count", _I, [_P]),
    ("seis_out_name", ctypes.c_char_p, [_P, _I]),
    ("seis_shot_bind", _I, [_P, _I, _FP, _FP, _I, _FP]),
    ("seis_launch_init", _I, [_P]),
    ("seis_launch_sources", _I, [_P, _I, _I]),
    ("seis_launch_update_v", _I, [_P]),
    ("seis_launch_update_s", _I, [_P]),
    ("seis_launch_surface", _I, [_P]),
    ("seis_launch_varsout", _I, [_P, _I]),
    ("seis_shot_read", _I, [_P, _FP]),
    ("seis_shot_read_device", _I, [_P, _P]),
    ("seis_shot_read_f64", _I, [_P, ctypes.POINTER(ctypes.c_double)]),
    ("seis_hess_read", _I, [_P, _FP]),
    ("seis_grad_set_memory", _I, [_P, ctypes.c_size_t]),
    ("seis_grad_plan", _I, [_P] + [ctypes.POINTER(_I)] * 4 + [ctypes.POINTER(ctypes.c_size_t)]),
    ("seis_run_steps", _I, [_P, _I, _I]),
    ("seis_forward_shot", _I, [_P, _I, _FP, _FP, _I, _FP, _FP]),
    ("seis_forward_shot_static", _I, [_P, _I, _FP, _FP, _I, _FP, _FP]),
    ("seis_read_field", _I, [_P, ctypes.c_char_p, _FP]),
    ("seis_grad_zero", _I, [_P]),
    ("seis_forward_save", _I, [_P, _I, _FP, _FP, _I, _FP, _FP]),
    ("seis_grad_saved_size", _I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    ("seis_grad_saved_read", _I, [_P, _FP]),
    ("seis_grad_saved_write", _I, [_P, _FP]),
    ("seis_residuals", _I, [_P, _I, _FP, _FP, _FP, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double)]),
    ("seis_misfit", _I, [ctypes.POINTER(SeisCsts), _I, _I, _FP, _FP, _FP,
                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    ("seis_res_scale", _I, [_P, _I, _FP, _FP]),
    ("seis_adjoint", _I, [_P, _FP]),
    ("seis_adjoint_device", _I, [_P, _P, _I]),
    ("seis_adjoint_movie", _I, [_P, _FP, _I, _I, ctypes.POINTER(ctypes.c_char_p), _FP]),
    ("seis_grad_dft", _I, [_P, _I, _FP]),
    ("seis_grad_read", _I, [_P, _FP]),
    ("seis_grad_read_tau", _I, [_P, _FP]),
    ("seis_grad_read_device", _I, [_P, _P, _P]),
    ("seis_gradient_shot", _I, [_P, _I, _FP, _FP, _I, _FP, _FP, _FP, _FP,
                                ctypes.POINTER(ctypes.c_double)]),
    ("seis_batch_create", _I, [_P, _I, ctypes.POINTER(_P)]),
    ("seis_batch_destroy", _I, [_P]),
    ("seis_batch_add", _I, [_P, _P, ctypes.POINTER(_I)]),
    ("seis_batch_remove", _I, [_P, _I]),
    ("seis_batch_step", _I, [_P, _I]),
    ("seis_batch_status", _I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    ("seis_batch_slot_left", _I, [_P, _I]),
    ("seis_model_set_profiling", _I, [_P, _I]),
    ("seis_last_timing", _I, [_P, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double)]),
]

_lib = None


class SeisError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""


def load() -> ctypes.CDLL:
    """Load libseiscl_amd.so and bind the signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SeisError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or "
            "make -C seiscl_amd/csrc")
    # The library and PyTorch-ROCm link HIP runtimes of the same soname
    # (libamdhip64.so.7), so a process holds one of them: the one loaded
    # first.  With the library's loaded first, torch finds no HIP device; with
    # torch's loaded first, both run on it.  So torch (when installed) is
    # loaded first, and device tensors can pass into the C-ABI.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().seis_last_error().decode(errors="replace")
        raise SeisError(msg or f"seiscl_amd call failed ({status})")


def fptr(a: np.ndarray | None):
    """float* of a C-contiguous float32 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)
