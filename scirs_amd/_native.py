"""ctypes binding of libscirs_gpu.so (the C ABI in include/scirs_gpu.h).

The library is the only compute path: if it is missing or fails to load this
module raises, it never falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libscirs_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "scirs_gpu.h")

SCIRS_OK = 0
SCIRS_ERR_ARG = -1
SCIRS_ERR_HIP = -2
SCIRS_ERR_NOMEM = -3
SCIRS_ERR_TIMEOUT = -4
SCIRS_ERR_NOT_CONVERGED = -5
SCIRS_ERR_UNSUPPORTED = -6

PATH_FP64_LU, PATH_FP64_CHOL, PATH_MIXED_IR, PATH_MIXED_GMRES, PATH_FP64_FALLBACK, PATH_FP32_LU, PATH_FP32_CHOL = range(7)
MODE_AUTO, MODE_BF16, MODE_FP16, MODE_BF16X3, MODE_FP32, MODE_BF16X6, MODE_FP16X3 = range(7)
PRECOND_NONE, PRECOND_DIAGONAL, PRECOND_LU, PRECOND_LU_GIVEN = range(4)


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("path", ctypes.c_int32),
        ("ir_iters", ctypes.c_int32),
        ("gmres_iters", ctypes.c_int32),
        ("pivot_value", ctypes.c_double),
        ("backward_error", ctypes.c_double),
        ("time_factor_ms", ctypes.c_double),
        ("time_solve_ms", ctypes.c_double),
        ("time_total_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ProfStats(ctypes.Structure):
    _fields_ = [
        ("launches", ctypes.c_int64),
        ("total_ms", ctypes.c_double),
        ("flops", ctypes.c_double),
        ("bytes", ctypes.c_double),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("failed", ctypes.c_int32),
        ("max_iters", ctypes.c_int32),
        ("max_backward_error", ctypes.c_double),
        ("time_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("wave_max", ctypes.c_int32),
        ("graph_replayed", ctypes.c_int32),
    ]


PROF_CLASSES = ["panel", "laswp", "trsm", "update", "trsv", "residual", "convert", "other"]

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double
_INFO = ctypes.POINTER(SolveInfo)

SIGNATURES = {
    "scirs_gpu_init": ([_I], _I),
    "scirs_gpu_sgetrf_tp_dev": ([_I, _P, _L, _P, _I, _P], _I),
    "scirs_gpu_seq_plan": ([ctypes.c_uint, ctypes.c_uint, _P, _P], _I),
    "scirs_gpu_last_error": ([], ctypes.c_char_p),
    "scirs_gpu_dgetrf": ([_I, _I, _P, _L, _P, _INFO], _I),
    "scirs_gpu_sgetrf": ([_I, _I, _P, _L, _P, _INFO], _I),
    "scirs_gpu_dgesv": ([_I, _I, _P, _L, _P, _L, _P, _L, _INFO], _I),
    "scirs_gpu_sgesv": ([_I, _I, _P, _L, _P, _L, _P, _L, _INFO], _I),
    "scirs_gpu_dpotrf": ([_I, _P, _L, _INFO], _I),
    "scirs_gpu_spotrf": ([_I, _P, _L, _INFO], _I),
    "scirs_gpu_dposv": ([_I, _I, _P, _L, _P, _L, _P, _L, _INFO], _I),
    "scirs_gpu_dtrsv": ([_I, _I, _P, _L, _P, _L, _P, _L, _I, _I, _INFO], _I),
    "scirs_gpu_strsv": ([_I, _I, _P, _L, _P, _L, _P, _L, _I, _I, _INFO], _I),
    "scirs_gpu_sposv": ([_I, _I, _P, _L, _P, _L, _P, _L, _INFO], _I),
    "scirs_gpu_dgmres": ([_I, _P, _L, _P, _P, _I, _D, _I, _P, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dgmres_dev": ([_I, _P, _L, _P, _P, _I, _I, _D, _I, _P, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dgmres_precond": ([_I, _P, _L, _P, _P, _I, _D, _I, _I, _P, _P, _P, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dgmres_precond_dev": ([_I, _P, _L, _P, _P, _I, _I, _D, _I, _I, _P, _P, _P, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dbatch_matmul": ([_I, _I, _I, _I, _P, _P, _P, _INFO], _I),
    "scirs_gpu_sbatch_matmul": ([_I, _I, _I, _I, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dbatch_matvec": ([_I, _I, _I, _P, _P, _P, _INFO], _I),
    "scirs_gpu_sbatch_matvec": ([_I, _I, _I, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dbatch_add": ([_I, _I, _I, _P, _P, _I, _P, _INFO], _I),
    "scirs_gpu_sbatch_add": ([_I, _I, _I, _P, _P, _I, _P, _INFO], _I),
    "scirs_gpu_dbatch_sum": ([_I, _I, _I, _P, _P, _INFO], _I),
    "scirs_gpu_sbatch_sum": ([_I, _I, _I, _P, _P, _INFO], _I),
    "scirs_gpu_dbatch_matmul_dev": ([_I, _I, _I, _I, _P, _P, _P, _P], _I),
    "scirs_gpu_dqr": ([_I, _I, _P, _L, _P, _L, _P, _L, _INFO], _I),
    "scirs_gpu_dlstsq": ([_I, _I, _P, _L, _P, _P, _P, _INFO], _I),
    "scirs_gpu_dgesv_ir": ([_I, _I, _P, _L, _P, _L, _P, _L, _I, _I, _I, _D, _INFO], _I),
    "scirs_gpu_dgesv_dev": ([_I, _I, _P, _L, _P, _L, _P, _L, _P, _L, _P, _P, _INFO], _I),
    "scirs_gpu_dposv_dev": ([_I, _I, _P, _L, _P, _L, _P, _L, _P, _L, _P, _INFO], _I),
    "scirs_gpu_dgesv_ir_dev": ([_I, _I, _P, _L, _P, _L, _P, _L, _I, _I, _I, _D, _P, _INFO], _I),
    "scirs_gpu_dgesv_batched_dev": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _D, _P,
                                     ctypes.POINTER(BatchInfo)], _I),
    "scirs_gpu_dgesv_batched": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _D,
                                 ctypes.POINTER(BatchInfo)], _I),
    "scirs_gpu_prof_begin": ([], _I),
    "scirs_gpu_prof_end": ([ctypes.POINTER(ProfStats)], _I),
    "scirs_gpu_backward_error_dev": ([_I, _I, _P, _L, _P, _L, _P, _L, _P, ctypes.POINTER(ctypes.c_double)], _I),
    "scirs_gpu_backward_error_dev_noalloc": ([_I, _I, _P, _L, _P, _L, _P, _L, _P, ctypes.POINTER(ctypes.c_double)],
                                             _I),
}

_lib = None


def header_symbols():
    """Every function declared in include/scirs_gpu.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(scirs_gpu_\w+)\s*\(", text)))


def load():
    """Load and bind the library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m scirs_amd.build` "
            "(there is no CPU fallback)")
    # One HIP runtime per process: torch ships its own libamdhip64 (same soname),
    # which it cannot share if the system one is loaded first, while this
    # library binds to whichever copy is already resident.  So torch, when
    # installed, is loaded before the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def last_error() -> str:
    return load().scirs_gpu_last_error().decode(errors="replace")
