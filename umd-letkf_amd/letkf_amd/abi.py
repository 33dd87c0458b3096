"""ctypes mirror of include/letkf_gpu.h and the loader of the HIP library.

The library is built in-tree (``__graft_entry__.build()``) as
``umd-letkf_amd/letkf_amd/libletkf_gpu.so``.  Loading fails loudly when it
is missing: there is no CPU fallback for the analysis.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libletkf_gpu.so")

LETKF_OK = 0
LETKF_EINVAL = 1
LETKF_EUNSUPPORTED = 2
LETKF_ENOMEM = 3
LETKF_EHIP = 4
LETKF_ECAPACITY = 5
LETKF_ESTATE = 6

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2
PREC_BF16 = 3
PREC_BF16X4 = 4
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3,
              "bf16": PREC_BF16, "bf16x4": PREC_BF16X4}

PATH_PLAIN = 0
PATH_BATCHED = 1

LOC_NOVRT = 1
LOC_OCEAN = 2
LOC_VGAUSS = 3

MAX_INTERP = 32
MAX_IDLIST = 64


class Dims(C.Structure):
    _fields_ = [("ens_size", C.c_int64), ("grid_ns", C.c_int64),
                ("ij_count", C.c_int64), ("max_obs", C.c_int64),
                ("precision", C.c_int32), ("path", C.c_int32)]


class LinearInterpLat(C.Structure):
    _fields_ = [("npoints", C.c_int32), ("lat", C.c_float * MAX_INTERP),
                ("radius", C.c_float * MAX_INTERP)]


class LocDesc(C.Structure):
    _fields_ = [("cls", C.c_int32),
                ("hzloc", LinearInterpLat),
                ("hzloc_prof", LinearInterpLat),
                ("hzloc_sat", LinearInterpLat),
                ("tloc_prof", C.c_float), ("tloc_sat", C.c_float),
                ("vtloc_surf", C.c_int32), ("bkg_t_delta", C.c_float),
                ("bkg_t_var_idx", C.c_int32), ("bkg_t_var_len", C.c_int32),
                ("n_sat_obs", C.c_int32), ("sat_obs", C.c_int32 * MAX_IDLIST),
                ("n_sat_plats", C.c_int32),
                ("sat_plats", C.c_int32 * MAX_IDLIST),
                ("vt_scale", C.c_float), ("nlev", C.c_int32),
                ("lev_depth", C.POINTER(C.c_float))]


class Diag(C.Structure):
    _fields_ = [("maxhz", C.POINTER(C.c_float)),
                ("obscount", C.POINTER(C.c_float)),
                ("lg_obscount", C.POINTER(C.c_float)),
                ("lg_obsloc", C.POINTER(C.c_float)),
                ("lg_selhash", C.POINTER(C.c_uint64))]


class Timing(C.Structure):
    _fields_ = [("search_ms", C.c_double), ("product_ms", C.c_double),
                ("eigen_ms", C.c_double), ("finalize_ms", C.c_double),
                ("total_ms", C.c_double), ("n_problems", C.c_int64),
                ("n_pairs", C.c_int64), ("n_cand", C.c_int64),
                ("n_launch", C.c_int64), ("order_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("order", C.c_int32), ("bucket_bits", C.c_int32),
                ("tridiag_reg", C.c_int32), ("n_tried", C.c_int32),
                ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("overlap", C.c_int32), ("chunk_log2", C.c_int32), ("xcd", C.c_int32),
                ("tile", C.c_int32)]


EXPORTED = [
    "letkf_gpu_version", "letkf_gpu_create", "letkf_gpu_destroy",
    "letkf_gpu_last_error", "letkf_gpu_set_state", "letkf_gpu_set_obs",
    "letkf_gpu_set_statevars", "letkf_gpu_set_localizer",
    "letkf_gpu_set_solver", "letkf_gpu_run", "letkf_gpu_get_state",
    "letkf_gpu_get_state_f64", "letkf_gpu_get_diag", "letkf_gpu_maxgroups",
    "letkf_gpu_get_selection", "letkf_gpu_get_timing",
    "letkf_gpu_set_state_members", "letkf_gpu_set_obs_members",
    "letkf_gpu_get_state_members", "letkf_gpu_autotune", "letkf_gpu_get_tuning",
    "letkf_gpu_set_tuning", "letkf_gpu_get_state_members_noalloc", "letkf_gpu_run_graph",
]

_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_H = C.c_void_p

_SIGS = {
    "letkf_gpu_version": (C.c_char_p, []),
    "letkf_gpu_create": (C.c_int, [C.POINTER(Dims), C.POINTER(_H)]),
    "letkf_gpu_destroy": (C.c_int, [_H]),
    "letkf_gpu_last_error": (C.c_char_p, [_H]),
    "letkf_gpu_set_state": (C.c_int, [_H, _FP, _FP, _FP, _FP, _FP, _IP]),
    "letkf_gpu_set_obs": (C.c_int, [_H, C.c_int64, _FP, _FP, _FP, _FP, _FP,
                                    _FP, _IP, _IP, _FP, _FP]),
    "letkf_gpu_set_statevars": (C.c_int, [_H, C.c_int32, _IP, _IP, _FP, _FP]),
    "letkf_gpu_set_localizer": (C.c_int, [_H, C.POINTER(LocDesc)]),
    "letkf_gpu_set_solver": (C.c_int, [_H, C.c_float, C.c_float, C.c_float]),
    "letkf_gpu_run": (C.c_int, [_H]),
    "letkf_gpu_get_state": (C.c_int, [_H, _FP, _FP, _FP]),
    "letkf_gpu_get_state_f64": (C.c_int, [_H, _DP, _DP, _DP]),
    "letkf_gpu_get_diag": (C.c_int, [_H, C.POINTER(Diag)]),
    "letkf_gpu_maxgroups": (C.c_int, [_H, _IP]),
    "letkf_gpu_get_selection": (C.c_int, [_H, C.c_int64, C.c_int32, _IP, _FP,
                                          C.c_int64, C.POINTER(C.c_int64)]),
    "letkf_gpu_get_timing": (C.c_int, [_H, C.POINTER(Timing)]),
    "letkf_gpu_set_state_members": (C.c_int, [_H, _FP, _FP, _FP, _IP]),
    "letkf_gpu_set_obs_members": (C.c_int, [_H, C.c_int64, _FP, _FP, _FP, _FP, _FP,
                                            _FP, _IP, _IP, _IP, _FP]),
    "letkf_gpu_get_state_members": (C.c_int, [_H, _FP, _FP, _FP]),
    "letkf_gpu_get_state_members_noalloc": (C.c_int, [_H, _FP, _FP, _FP]),
    "letkf_gpu_run_graph": (C.c_int, [_H]),
    "letkf_gpu_autotune": (C.c_int, [_H, C.c_int32, C.POINTER(Tuning)]),
    "letkf_gpu_get_tuning": (C.c_int, [_H, C.POINTER(Tuning)]),
    "letkf_gpu_set_tuning": (C.c_int, [_H, C.POINTER(Tuning)]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load_library(path: str | None = None) -> C.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibraryMissing(
            f"{p} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


_NP_OF = {C.c_float: np.float32, C.c_double: np.float64, C.c_int32: np.int32,
          C.c_int64: np.int64, C.c_uint64: np.uint64}


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None -> NULL).  The array's
    dtype must be the C type's, and it must be contiguous (column-major for
    the reference's multi-dimensional arrays is the caller's contract)."""
    if a is None:
        return None
    want = _NP_OF.get(ctype)
    if want is not None and a.dtype != want:
        raise TypeError(f"array of dtype {a.dtype} passed as {ctype.__name__}")
    if not (a.flags.f_contiguous or a.flags.c_contiguous):
        raise ValueError("non-contiguous array passed to the C ABI")
    return a.ctypes.data_as(C.POINTER(ctype))
