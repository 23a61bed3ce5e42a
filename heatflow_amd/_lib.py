"""ctypes binding of libheatflow_hip.so (the C-ABI in include/heatflow_hip.h).

The library is loaded from this package directory only.  There is no CPU fallback: if the
shared library is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libheatflow_hip.so")

HF_OK = 0
HF_ERR_NAN_B, HF_ERR_NAN_Q, HF_ERR_NAN_QDENS, HF_ERR_NAN_SCAT, HF_ERR_NAN_S, HF_ERR_NAN_X = range(1, 7)
HF_ERR_ARG, HF_ERR_ALLOC, HF_ERR_HIP, HF_ERR_STATE, HF_ERR_NO_DEVICE = 10, 11, 12, 13, 14
HF_ERR_UNSTABLE = 15
HF_PREC_FP64, HF_PREC_FP32_IR = 0, 1
# kernel ids of hf_kernel_times
KERNELS = ("k_qdens", "k_rhs_res<fp64>", "k_rhs_res<fp32>", "k_cg_a<fp64>", "k_cg_b<fp64>",
           "k_cg_a<fp32>", "k_cg_b<fp32>", "k_finalize", "k_ir_res", "k_ir_step", "k_step64",
           "k_run_small", "k_ir_step2")

# every symbol include/heatflow_hip.h declares
EXPORTS = (
    "hf_params_default", "hf_create", "hf_upload_cells", "hf_assemble", "hf_export_coo",
    "hf_set_state", "hf_get_state", "hf_step", "hf_run", "hf_solve", "hf_get_vector",
    "hf_set_debug", "hf_get_stats", "hf_reset_stats", "hf_set_solver", "hf_grad",
    "hf_set_timing", "hf_kernel_times", "hf_forward_checkpointed", "hf_adjoint",
    "hf_set_fields", "hf_set_adjoint_tol", "hf_export_coo_noalloc",
    "hf_builtin_material", "hf_upload_materials", "hf_check_stability",
    "hf_last_error", "hf_destroy", "hf_version",
)


class HfParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("time_step", C.c_double),
        ("icattaneo", C.c_int32), ("isteady", C.c_int32),
        ("periodic", C.c_int32 * 3),
        ("cg", C.c_int32 * 6),
        ("kappa_bound", C.c_double * 6),
        ("t_bath", C.c_double * 6),
        ("br", C.c_double),
        ("t_bath_cg", C.c_double),
        ("power_in", C.c_double),
        ("freq", C.c_double),
        ("t_bath_rad", C.c_double),
        ("heated_steps", C.c_int64),
        ("tol", C.c_double),
        ("itmax", C.c_int64),
        ("precision", C.c_int32),
        ("inner_itmax", C.c_int32),
        ("inner_tol", C.c_double),
    ]


class HfStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("outer_iters", C.c_int64), ("inner_iters", C.c_int64),
        ("last_err", C.c_double), ("first_iter_stops", C.c_int64), ("tiny_misses", C.c_int64),
        ("residual_reruns", C.c_int64), ("alpha_reuses", C.c_int64), ("reuse_misses", C.c_int64),
        ("adjoint_solves", C.c_int64), ("adjoint_iters", C.c_int64),
    ]


class HfMaterial(C.Structure):
    _fields_ = [("index", C.c_int64), ("kappa", C.c_double), ("rho", C.c_double),
                ("heat_capacity", C.c_double), ("tau", C.c_double), ("em", C.c_double)]


class HeatFlowError(RuntimeError):
    """Raised for a non-zero HF_* return code; ``code`` holds it."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

P = C.c_void_p
D = C.POINTER(C.c_double)
I64 = C.POINTER(C.c_int64)


def load() -> C.CDLL:
    """Loads the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m heatflow_amd.build` "
            "(there is no CPU fallback)")
    # torch ships its own HIP and HSA runtimes.  Loaded first, they are the ones this library
    # binds to (same sonames), so the process has one runtime and torch tensors are device
    # memory of it; loaded after this library, a second HSA instance finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    sig = {
        "hf_params_default": (None, [C.POINTER(HfParams)]),
        "hf_create": (C.c_int, [C.POINTER(P), C.POINTER(HfParams)]),
        "hf_upload_cells": (C.c_int, [P, P, P, P, P, P, P, P, P, P, P, P]),
        "hf_assemble": (C.c_int, [P]),
        "hf_export_coo": (C.c_int, [P, I64, I64, I64, P, P, P]),
        "hf_export_coo_noalloc": (C.c_int, [P, I64, I64, I64, P, P, P]),
        "hf_builtin_material": (C.c_int, [C.c_int64, C.POINTER(HfMaterial), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32)]),
        "hf_upload_materials": (C.c_int, [P, P, P, C.c_int32, P, C.c_double, C.c_double,
                                          C.c_double]),
        "hf_check_stability": (C.c_int, [P, I64, D]),
        "hf_set_state": (C.c_int, [P, P, P]),
        "hf_get_state": (C.c_int, [P, P, P]),
        "hf_step": (C.c_int, [P, C.c_int64, I64, D]),
        "hf_run": (C.c_int, [P, C.c_int64, C.c_int64, I64]),
        "hf_solve": (C.c_int, [P, P, P, C.c_int32, C.c_double, C.c_int64, I64, D]),
        "hf_get_vector": (C.c_int, [P, C.c_int32, P]),
        "hf_set_debug": (C.c_int, [P, C.c_int32]),
        "hf_get_stats": (C.c_int, [P, C.POINTER(HfStats)]),
        "hf_reset_stats": (C.c_int, [P]),
        "hf_set_solver": (C.c_int, [P, C.c_double, C.c_int64, C.c_int32, C.c_double, C.c_int32]),
        "hf_set_timing": (C.c_int, [P, C.c_int32]),
        "hf_forward_checkpointed": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int64]),
        "hf_adjoint": (C.c_int, [P, P, P, P]),
        "hf_set_fields": (C.c_int, [P, P, P]),
        "hf_set_adjoint_tol": (C.c_int, [P, C.c_double]),
        "hf_kernel_times": (C.c_int, [P, P, P, P]),
        "hf_grad": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int64, C.c_int64, P, P, P, D, P, P]),
        "hf_last_error": (C.c_char_p, [P]),
        "hf_destroy": (None, [P]),
        "hf_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None) -> None:
    if rc != HF_OK:
        lib = load()
        msg = lib.hf_last_error(ctx)
        raise HeatFlowError(rc, msg.decode() if msg else "")
