"""CPU oracle for the CUNA adjoint NFFT (ctypes wrapper over libcuna_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as the product.  The HIP library (cunfft_adjoint_amd) never imports it.

Parity status of each function (how it is pinned):
  wang_hash / xorshift / bootstrap_indices  -- known answers of the reference RNG
  fft                                       -- brute-force DFT and numpy ifft*n
  compat_*                                  -- independent float32 step-by-step
                                               emulation + structural invariants
  gl_adjoint                                -- direct NDFT (closed form) at 1e-8
  ndft_bands                                -- closed forms (one point, F_0 = sum y,
                                               shift, conjugate symmetry) and
                                               brute-force numpy sums
  gl_grad / ndft_grad                       -- central finite differences
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcuna_oracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_int = ctypes.c_int


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (Makefile beside this file)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_wang_hash.restype = ctypes.c_uint32
        L.orc_wang_hash.argtypes = [ctypes.c_uint32]
        L.orc_xorshift.restype = ctypes.c_uint32
        L.orc_xorshift.argtypes = [ctypes.c_uint32]
        L.orc_bootstrap_indices.argtypes = [ctypes.c_uint32, _int, _int, _i32]
        L.orc_compat_params.argtypes = [_int, _int, _int, _d, _d, ctypes.POINTER(_int)]
        L.orc_compat_point.argtypes = [ctypes.c_double, _int, _int, _int,
                                       ctypes.POINTER(_int), _d, ctypes.POINTER(_int)]
        L.orc_compat_grid.argtypes = [_d, _d, _int, _int, _int, _d]
        L.orc_compat_grid_bootstrap.argtypes = [_d, _d, _int, _int, _int, ctypes.c_uint32, _int, _d]
        L.orc_compat_adjoint.argtypes = [_d, _d, _int, _int, _int, _d]
        L.orc_compat_adjoint_bootstrap.argtypes = [_d, _d, _int, _int, _int, ctypes.c_uint32,
                                                   _int, _d]
        L.orc_fft.argtypes = [_d, _int, _int]
        L.orc_gl_adjoint.argtypes = [_d, _d, _int, _int, _int, _int, _d]
        L.orc_gl_grad.argtypes = [_d, _d, _int, _int, _int, _int, _d, _d, _d]
        L.orc_ndft_bands.argtypes = [_d, _d, _int, _i64, _int, _int, _d]
        L.orc_ndft_grad.argtypes = [_d, _d, _int, _int, _d, _d, _d]
        L.orc_num_threads.restype = _int
        L.orc_lomb_scargle.argtypes = [_d, _d, _int, ctypes.c_double, ctypes.c_double, _int, _d]
        L.orc_ndft_forward.argtypes = [_d, _int, _d, _int, _d]
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_d)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return int(lib().orc_num_threads())


def wang_hash(s: int) -> int:
    return int(lib().orc_wang_hash(s & 0xFFFFFFFF))


def xorshift(s: int) -> int:
    return int(lib().orc_xorshift(s & 0xFFFFFFFF))


def bootstrap_indices(seed: int, n: int, nbs: int) -> np.ndarray:
    d = np.empty(n * nbs, dtype=np.int32)
    lib().orc_bootstrap_indices(seed & 0xFFFFFFFF, n, nbs, d.ctypes.data_as(_i32))
    return d.reshape(nbs, n)


def compat_params(n: int, ng: int, dbl: bool):
    b, nf, r = ctypes.c_double(), ctypes.c_double(), _int()
    lib().orc_compat_params(n, ng, int(dbl), ctypes.byref(b), ctypes.byref(nf), ctypes.byref(r))
    return b.value, nf.value, r.value


def compat_point(x: float, n: int, ng: int, dbl: bool):
    u, e, j = _int(), ctypes.c_double(), _int()
    lib().orc_compat_point(float(x), n, ng, int(dbl), ctypes.byref(u), ctypes.byref(e),
                           ctypes.byref(j))
    return u.value, e.value, j.value


def compat_grid(x, y, ng: int, dbl: bool) -> np.ndarray:
    x = _f64(x)
    y = _f64(y)
    g = np.empty(ng, dtype=np.float64)
    if lib().orc_compat_grid(_dp(x), _dp(y), len(x), ng, int(dbl), _dp(g)):
        raise ValueError("bad arguments")
    return g


def compat_grid_bootstrap(x, y, ng: int, nbs: int, seed: int, dbl: bool) -> np.ndarray:
    x = _f64(x)
    y = _f64(y)
    g = np.empty(ng * nbs, dtype=np.float64)
    if lib().orc_compat_grid_bootstrap(_dp(x), _dp(y), len(x), ng, nbs, seed & 0xFFFFFFFF,
                                       int(dbl), _dp(g)):
        raise ValueError("bad arguments")
    return g.reshape(nbs, ng)


def compat_adjoint(x, y, ng: int, dbl: bool) -> np.ndarray:
    x = _f64(x)
    y = _f64(y)
    out = np.empty(2 * ng, dtype=np.float64)
    if lib().orc_compat_adjoint(_dp(x), _dp(y), len(x), ng, int(dbl), _dp(out)):
        raise ValueError("bad arguments")
    return out.view(np.complex128)


def compat_adjoint_bootstrap(x, y, ng: int, nbs: int, seed: int, dbl: bool) -> np.ndarray:
    x = _f64(x)
    y = _f64(y)
    out = np.empty(2 * ng * nbs, dtype=np.float64)
    if lib().orc_compat_adjoint_bootstrap(_dp(x), _dp(y), len(x), ng, nbs, seed & 0xFFFFFFFF,
                                          int(dbl), _dp(out)):
        raise ValueError("bad arguments")
    return out.view(np.complex128).reshape(nbs, ng)


def fft(z, sign: int = +1) -> np.ndarray:
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().orc_fft(z.view(np.float64).ctypes.data_as(_d), len(z), sign)
    return z


def gl_adjoint(x, y, M: int, sigma: int = 2, m: int = 8) -> np.ndarray:
    x = _f64(x)
    y = _f64(y)
    out = np.empty(2 * M, dtype=np.float64)
    if lib().orc_gl_adjoint(_dp(x), _dp(y), len(x), M, sigma, m, _dp(out)):
        raise ValueError("bad arguments")
    return out.view(np.complex128)


def gl_grad(x, y, M: int, g, sigma: int = 2, m: int = 8):
    x = _f64(x)
    y = _f64(y)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    n = len(x)
    dy = np.empty(n)
    dx = np.empty(n)
    if lib().orc_gl_grad(_dp(x), _dp(y), n, M, sigma, m, g.view(np.float64).ctypes.data_as(_d),
                         _dp(dy), _dp(dx)):
        raise ValueError("bad arguments")
    return dy, dx


def ndft_bands(x, y, k0, length: int) -> np.ndarray:
    """F_k = sum_j y_j exp(2 pi i k x_j) for k in k0[b] .. k0[b]+length-1."""
    x = _f64(x)
    y = _f64(y)
    k0 = np.ascontiguousarray(np.atleast_1d(k0), dtype=np.int64)
    out = np.empty(2 * len(k0) * length, dtype=np.float64)
    lib().orc_ndft_bands(_dp(x), _dp(y), len(x), k0.ctypes.data_as(_i64), len(k0), length,
                         _dp(out))
    return out.view(np.complex128).reshape(len(k0), length)


def ndft(x, y, M: int) -> np.ndarray:
    """Full spectrum k = -M/2 .. M/2-1."""
    return ndft_bands(x, y, [-(M // 2)], M)[0]


def ndft_grad(x, y, M: int, g):
    x = _f64(x)
    y = _f64(y)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    n = len(x)
    dy = np.empty(n)
    dx = np.empty(n)
    lib().orc_ndft_grad(_dp(x), _dp(y), n, M, g.view(np.float64).ctypes.data_as(_d), _dp(dy),
                        _dp(dx))
    return dy, dx


def lomb_scargle(t, y, df: float, nf: int, f0: float = 0.0) -> np.ndarray:
    """Lomb-Scargle power (centred data, standard normalisation) at
    f0 + k df, k = 0 .. nf-1, by the direct FP64 sums."""
    t = _f64(t)
    y = _f64(y)
    out = np.empty(nf, dtype=np.float64)
    lib().orc_lomb_scargle(_dp(t), _dp(y), len(t), float(f0), float(df), int(nf), _dp(out))
    return out


def ndft_forward(x, fhat) -> np.ndarray:
    """f_j = sum_k fhat_k exp(-2 pi i k x_j), k = -M/2 .. M/2-1 (fhat natural order)."""
    x = _f64(x)
    fh = np.ascontiguousarray(fhat, dtype=np.complex128)
    out = np.empty(2 * len(x), dtype=np.float64)
    lib().orc_ndft_forward(_dp(x), len(x), fh.view(np.float64).ctypes.data_as(_d), len(fh),
                           _dp(out))
    return out.view(np.complex128)
