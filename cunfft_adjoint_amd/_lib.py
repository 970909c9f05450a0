"""ctypes loader for libcunaf.so / libcunad.so (the C-ABI of include/*.h).

The libraries are built in-tree by cunfft_adjoint_amd.build; nothing here
falls back to another implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
LIBNAMES = {"f": "libcunaf.so", "d": "libcunad.so"}

c_int, c_uint, c_double, c_void_p = ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_void_p
c_char_p = ctypes.c_char_p

# status codes (include/cuna_typedefs.h)
CUNA_SUCCESS = 0
CUNA_ERROR_INVALID_VALUE = 1
CUNA_ERROR_UNSUPPORTED = 2
CUNA_ERROR_HIP = 3
CUNA_ERROR_OUT_OF_MEMORY = 4
CUNA_ERROR_NOT_FOUND = 5

CUNA_MODE_ACCURATE = 0
CUNA_MODE_COMPAT = 1
CUNA_PLAN_WINDOW2 = 0x1
CUNA_PLAN_UNFUSED = 0x2
CUNA_PLAN_GATHER = 0x4

# FLAGS (inc/cuna_typedefs.h:110-116 of the reference)
NO_FLAGS = 0x01
OUTPUT_INTERMEDIATE = 0x02
CALCULATE_WINDOW_FUNCTION = 0x04
DONT_TRANSFER_TO_CPU = 0x08
PRINT_TIMING = 0x10


def real_type(prec: str):
    return ctypes.c_float if prec == "f" else ctypes.c_double


def make_types(prec: str):
    R = real_type(prec)

    class Complex(ctypes.Structure):
        _fields_ = [("x", R), ("y", R)]

    class FilterProperties(ctypes.Structure):
        _fields_ = [("b", R), ("normfac", R), ("filter_radius", c_int),
                    ("E1", c_void_p), ("E2", c_void_p), ("E3", c_void_p)]

    class Plan(ctypes.Structure):
        _fields_ = [("x_data", c_void_p), ("f_data", c_void_p),
                    ("f_hat", c_void_p), ("f_hat_win", c_void_p),
                    ("g_f_hat", c_void_p), ("g_f_hat_win", c_void_p),
                    ("g_x_data", c_void_p), ("g_f_data", c_void_p),
                    ("flags", c_uint), ("Ndata", c_int), ("Ngrid", c_int),
                    ("filter_radius", c_int),
                    ("fprops_host", c_void_p), ("fprops_device", c_void_p)]

    return R, Complex, FilterProperties, Plan


class NfftParams(ctypes.Structure):
    _fields_ = [("mode", c_int), ("sigma", c_int), ("m", c_int), ("tile_log2", c_int),
                ("flags", c_int), ("fft_w", c_int), ("fft_h", c_int)]


class NfftInfo(ctypes.Structure):
    _fields_ = [("mode", c_int), ("sigma", c_int), ("m", c_int), ("ntaps", c_int),
                ("n", c_int), ("ncurves", c_int), ("M", c_int), ("G", c_int),
                ("tile_log2", c_int), ("ntiles", c_int), ("beta", c_double)]


def _protos(prec: str):
    R = real_type(prec)
    P = c_void_p
    RP = ctypes.POINTER(R)
    return {
        # include/cuna.h
        "cunfft_adjoint": (None, [RP, RP, P, c_int, c_int, c_uint]),
        "cunfft_adjoint_from_plan": (None, [P]),
        "cunfft_adjoint_raw": (None, [P, P, P, c_int, c_int, P]),
        "cunfft_adjoint_raw_async": (None, [P, P, P, c_int, c_int, P, P]),
        "cunfft_adjoint_raw_async_noalloc": (None, [P, P, P, c_int, c_int, P, P]),
        "cunfft_adjoint_raw_async_bootstrap": (None, [P, P, P, c_int, c_int, c_int, P, P, c_uint]),
        # include/cuna_nfft.h: the reference's commented-out batched API
        "cunfft_adjoint_raw_batched_async": (None, [P, P, P, P, P, c_int, P, P]),
        "generate_pinned_filter_properties_batch": (None, [P, P, P, c_int, P, P, P]),
        "cuna_batch_cache_clear": (None, []),
        # include/cuna_typedefs.h
        "init_plan": (None, [P, RP, RP, c_int, c_int, c_uint]),
        "free_plan": (None, [P]),
        "cuna_get_last_error": (c_int, []),
        "cuna_get_last_error_string": (c_char_p, []),
        "cuna_clear_last_error": (None, []),
        "cuna_set_abort_on_error": (None, [c_int]),
        "cuna_real_size": (c_int, []),
        # include/cuna_filter.h
        "set_filter_properties": (None, [P]),
        "generate_pinned_filter_properties": (None, [P, c_int, c_int, P, P, P]),
        "generate_filter_properties": (None, [P, c_int, c_int, P, P]),
        "free_filter_properties": (None, [P, P]),
        "cuna_filter_radius": (c_int, []),
        # include/cuna_nfft.h
        "cuna_nfft_default_params": (None, [P]),
        "cuna_nfft_plan_create": (c_int, [P, P, c_int, c_int, P, P]),
        "cuna_nfft_plan_create_times": (c_int, [P, P, c_int, c_double, c_int, P, P]),
        "cuna_nfft_plan_create_batch": (c_int, [P, P, P, c_int, c_int, P, P]),
        "cuna_nfft_plan_create_batch_times": (c_int, [P, P, P, c_int, c_double, c_int, P, P]),
        "cuna_nfft_adjoint": (c_int, [P, P, P, P]),
        "cuna_nfft_adjoint_many": (c_int, [P, P, c_int, P, P]),
        "cuna_nfft_bootstrap": (c_int, [P, P, c_int, c_uint, P, P, P]),
        "cuna_nfft_bootstrap_ls": (c_int, [P, P, c_int, c_uint, P, P, P, P]),
        "cuna_nfft_forward": (c_int, [P, P, P, P]),
        "cuna_nfft_backward": (c_int, [P, P, P, P, P, P]),
        "cuna_nfft_plan_info": (c_int, [P, P]),
        "cuna_nfft_set_profile_events": (c_int, [P, P, c_int]),
        "cuna_nfft_plan_destroy": (c_int, [P]),
        "cuna_nfft_autotune": (c_int, [P, c_int, c_int, c_double, c_char_p, P, P]),
        "cuna_stream_capture_begin": (c_int, [P]),
        "cuna_stream_capture_end": (c_int, [P, P]),
        "cuna_graph_launch": (c_int, [P, P]),
        "cuna_graph_destroy": (c_int, [P]),
        "cuna_fft": (c_int, [P, c_int, c_int, c_int, P]),
    }


_cache: dict[str, ctypes.CDLL] = {}


def lib_path(prec: str) -> str:
    return os.path.join(LIBDIR, LIBNAMES[prec])


def load(prec: str = "f") -> ctypes.CDLL:
    """Load libcunaf.so ('f') or libcunad.so ('d'); raises if not built."""
    if prec not in LIBNAMES:
        raise ValueError("precision must be 'f' or 'd'")
    if prec in _cache:
        return _cache[prec]
    path = lib_path(prec)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m cunfft_adjoint_amd.build`")
    L = ctypes.CDLL(path)
    for name, (res, args) in _protos(prec).items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _cache[prec] = L
    return L


def exported_functions(prec: str) -> list[str]:
    return list(_protos(prec).keys())


class CunaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"CUNA error {code}: {msg}")
        self.code = code


def check(L: ctypes.CDLL, status: int | None = None) -> None:
    code = L.cuna_get_last_error() if status is None else status
    if code != CUNA_SUCCESS:
        msg = L.cuna_get_last_error_string()
        raise CunaError(code, msg.decode() if msg else "")
