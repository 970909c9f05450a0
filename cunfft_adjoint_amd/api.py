"""Python mirror of the CUNA interface over the C-ABI (libcunaf / libcunad).

Function names, argument order and meaning follow the reference's inc/cuna.h,
inc/cuna_filter.h and inc/cuna_typedefs.h; device buffers are torch tensors
(PyTorch provides device memory and streams), host buffers numpy arrays.
Errors raise CunaError instead of terminating the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import CunaError, check, load

_PREC = {np.dtype(np.float32): "f", np.dtype(np.float64): "d"}


def _prec_of(arr) -> str:
    import torch

    if isinstance(arr, torch.Tensor):
        dt = {torch.float32: np.dtype(np.float32), torch.float64: np.dtype(np.float64)}.get(arr.dtype)
    else:
        dt = np.asarray(arr).dtype
    if dt not in _PREC:
        raise TypeError("x must be float32 (libcunaf) or float64 (libcunad)")
    return _PREC[dt]


def _ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device buffer expected (a CUDA/HIP torch tensor)")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor expected")
    return t.data_ptr()


def _stream(stream) -> int | None:
    import torch

    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def complex_dtype(prec: str):
    import torch

    return torch.complex64 if prec == "f" else torch.complex128


def real_dtype(prec: str):
    import torch

    return torch.float32 if prec == "f" else torch.float64


# ------------------------------------------------------ reference API --
def cunfft_adjoint(x, f, ng: int, flags: int = 0) -> np.ndarray:
    """inc/cuna.h:33-34: host arrays in, host spectrum (ng complex) out."""
    x = np.ascontiguousarray(x)
    prec = _prec_of(x)
    R = np.float32 if prec == "f" else np.float64
    x = x.astype(R, copy=False)
    L = load(prec)
    RP = ctypes.POINTER(_lib.real_type(prec))
    fp = None
    if f is not None:
        f = np.ascontiguousarray(f, dtype=R)
        fp = f.ctypes.data_as(RP)
    out = np.zeros(ng, dtype=np.complex64 if prec == "f" else np.complex128)
    L.cunfft_adjoint(x.ctypes.data_as(RP), fp, out.ctypes.data, len(x), ng, flags)
    check(L)
    return out


class FilterProperties:
    """Host + device filter_properties from generate_filter_properties
    (inc/cuna_filter.h:47-49).  `device` is the pointer the raw transforms take."""

    def __init__(self, x, n: int, ng: int):
        self.prec = _prec_of(x)
        self.L = load(self.prec)
        R, _, FP, _ = _lib.make_types(self.prec)
        self._FP = FP
        h = ctypes.c_void_p()
        d = ctypes.c_void_p()
        self.L.generate_filter_properties(_ptr(x), n, ng, ctypes.byref(h), ctypes.byref(d))
        check(self.L)
        self.host, self.device = h.value, d.value

    @property
    def fields(self):
        return self._FP.from_address(self.host)

    def free(self):
        if self.device is not None:
            self.L.free_filter_properties(self.device, self.host)
            self.host = self.device = None
            check(self.L)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_filter_properties(x, n: int, ng: int) -> FilterProperties:
    return FilterProperties(x, n, ng)


def generate_pinned_filter_properties(x, n: int, ng: int, h_fprops, d_fprops, stream=None):
    """inc/cuna_filter.h:37-40; h_fprops is a ctypes filter_properties whose
    E1/E2/E3 device arrays the caller allocated, d_fprops a device pointer."""
    prec = _prec_of(x)
    L = load(prec)
    L.generate_pinned_filter_properties(_ptr(x), n, ng, ctypes.addressof(h_fprops), d_fprops,
                                        _stream(stream))
    check(L)


def cunfft_adjoint_raw(x, f_data, f_hat, n: int, ng: int, fprops: FilterProperties):
    """inc/cuna.h:48-49 (null stream)."""
    L = load(_prec_of(x))
    L.cunfft_adjoint_raw(_ptr(x), _ptr(f_data), _ptr(f_hat), n, ng,
                         fprops.device if isinstance(fprops, FilterProperties) else fprops)
    check(L)


def cunfft_adjoint_raw_async(x, f_data, f_hat, n: int, ng: int, fprops, stream=None):
    """inc/cuna.h:51-53."""
    L = load(_prec_of(x))
    L.cunfft_adjoint_raw_async(_ptr(x), _ptr(f_data), _ptr(f_hat), n, ng,
                               fprops.device if isinstance(fprops, FilterProperties) else fprops,
                               _stream(stream))
    check(L)


def cunfft_adjoint_raw_async_noalloc(x, f_data, f_hat, n: int, ng: int, fprops, stream=None):
    """cunfft_adjoint_raw_async without any device allocation (include/cuna.h;
    no counterpart in the reference): the grid lives in the library's fixed
    per-GPU scratch array, ng <= 2^22 (float) / 2^21 (double)."""
    L = load(_prec_of(x))
    L.cunfft_adjoint_raw_async_noalloc(_ptr(x), _ptr(f_data), _ptr(f_hat), n, ng,
                                       fprops.device if isinstance(fprops, FilterProperties)
                                       else fprops, _stream(stream))
    check(L)


def cunfft_adjoint_raw_async_bootstrap(x, f_data, f_hat, n: int, ng: int, nbs: int, fprops,
                                       stream=None, seed: int = 0):
    """src/cuna_kernel.cu:236-289 (the defined 9-argument form)."""
    L = load(_prec_of(x))
    L.cunfft_adjoint_raw_async_bootstrap(
        _ptr(x), _ptr(f_data), _ptr(f_hat), n, ng, nbs,
        fprops.device if isinstance(fprops, FilterProperties) else fprops, _stream(stream),
        seed & 0xFFFFFFFF)
    check(L)


def cunfft_adjoint_raw_batched_async(x, f_data, f_hat, n, ng, fprops=None, stream=None):
    """Batch of curves through the reference formula (the batched transform
    the reference keeps commented out, src/cuna_kernel.cu:302-358): curve i
    has n[i] points of the concatenated x / f_data and ng[i] cells; its
    spectrum fills the first ng[i] values of slab i of f_hat (slab stride
    max(ng)).  fprops: per-curve FilterProperties (or device pointers) from
    generate_filter_properties; when given, the multi-curve plan is cached."""
    L = load(_prec_of(x))
    nlc = len(n)
    if len(ng) != nlc:
        raise ValueError("n and ng must have the same length")
    na = (ctypes.c_int * nlc)(*[int(v) for v in n])
    ga = (ctypes.c_int * nlc)(*[int(v) for v in ng])
    fa = None
    if fprops is not None:
        if len(fprops) != nlc:
            raise ValueError("one filter_properties per curve")
        fa = (ctypes.c_void_p * nlc)(
            *[f.device if isinstance(f, FilterProperties) else f for f in fprops])
    L.cunfft_adjoint_raw_batched_async(_ptr(x), _ptr(f_data), _ptr(f_hat), na, ga, nlc, fa,
                                       _stream(stream))
    check(L)


def generate_pinned_filter_properties_batch(x, n, ng, h_fprops, d_fprops, stream=None):
    """src/gaussian_filter.cu:67-99 (commented out in the reference): h_fprops
    are ctypes filter_properties with caller-allocated E1/E2/E3, d_fprops
    device pointers, one each per curve."""
    L = load(_prec_of(x))
    nlc = len(n)
    na = (ctypes.c_int * nlc)(*[int(v) for v in n])
    ga = (ctypes.c_int * nlc)(*[int(v) for v in ng])
    ha = (ctypes.c_void_p * nlc)(*[ctypes.addressof(h) for h in h_fprops])
    da = (ctypes.c_void_p * nlc)(*d_fprops)
    L.generate_pinned_filter_properties_batch(_ptr(x), na, ga, nlc, ha, da, _stream(stream))
    check(L)


def batch_cache_clear(prec: str = "f"):
    """Frees the cached multi-curve plans of one library."""
    L = load(prec)
    L.cuna_batch_cache_clear()
    check(L)


class CunaPlan:
    """init_plan / cunfft_adjoint_from_plan / free_plan (inc/cuna_typedefs.h:151-155,
    inc/cuna.h:41) over host arrays."""

    def __init__(self, x, f, ng: int, flags: int = 0):
        x = np.ascontiguousarray(x)
        self.prec = _prec_of(x)
        R = np.float32 if self.prec == "f" else np.float64
        self.L = load(self.prec)
        _, _, _, PlanT = _lib.make_types(self.prec)
        RP = ctypes.POINTER(_lib.real_type(self.prec))
        self._x = x.astype(R, copy=False)
        self._f = None if f is None else np.ascontiguousarray(f, dtype=R)
        self._buf = ctypes.create_string_buffer(ctypes.sizeof(PlanT))
        self.L.init_plan(ctypes.addressof(self._buf), self._x.ctypes.data_as(RP),
                         None if self._f is None else self._f.ctypes.data_as(RP), len(self._x), ng,
                         flags)
        check(self.L)
        self.plan = PlanT.from_buffer(self._buf)
        self.ng, self.flags = ng, flags
        self._live = True

    def execute(self):
        self.L.cunfft_adjoint_from_plan(ctypes.addressof(self._buf))
        check(self.L)

    def f_hat(self) -> np.ndarray:
        dt = np.complex64 if self.prec == "f" else np.complex128
        return np.ctypeslib.as_array(
            ctypes.cast(self.plan.f_hat, ctypes.POINTER(_lib.real_type(self.prec))),
            shape=(2 * self.ng,)).view(dt).copy()

    def f_hat_win(self) -> np.ndarray:
        dt = np.complex64 if self.prec == "f" else np.complex128
        return np.ctypeslib.as_array(
            ctypes.cast(self.plan.f_hat_win, ctypes.POINTER(_lib.real_type(self.prec))),
            shape=(2 * self.ng,)).view(dt).copy()

    def free(self):
        # free_plan also frees the struct itself (src/utils.cu:275), so hand it
        # a malloc'd copy and keep our buffer
        if self._live:
            libc = ctypes.CDLL(None)
            libc.malloc.restype = ctypes.c_void_p
            libc.malloc.argtypes = [ctypes.c_size_t]
            p = libc.malloc(ctypes.sizeof(self._buf))
            ctypes.memmove(p, ctypes.addressof(self._buf), ctypes.sizeof(self._buf))
            self.L.free_plan(p)
            self._live = False
            check(self.L)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ----------------------------------------------------------- extensions --
class NFFTPlan:
    """Plan of include/cuna_nfft.h.

    NFFTPlan(x, M)                      accurate adjoint NFFT of points x
    NFFTPlan(x, M, mode="compat")       the reference formula (M = ng)
    NFFTPlan(t=t, df=df, M=M)           FP64 times, frequencies k*df
    NFFTPlan(x, M, offsets=[...])       batch of curves (CSR offsets, host)
    Precision follows x's dtype (float32 -> libcunaf, float64 -> libcunad);
    for times plans pass dtype.
    """

    def __init__(self, x=None, M: int = 0, *, t=None, df: float | None = None, offsets=None,
                 mode: str = "accurate", sigma: int = 2, m: int | None = None,
                 tile_log2: int = 0, window2: bool = False, unfused: bool = False,
                 gather: bool = False, dtype=None, fft_w: int = 0, fft_h: int = 0, stream=None):
        import torch

        if x is not None:
            self.prec = _prec_of(x)
        else:
            self.prec = "d" if dtype in (torch.float64, np.float64, "d") else "f"
        self.L = load(self.prec)
        p = _lib.NfftParams()
        self.L.cuna_nfft_default_params(ctypes.addressof(p))
        p.mode = _lib.CUNA_MODE_COMPAT if mode == "compat" else _lib.CUNA_MODE_ACCURATE
        p.sigma = sigma
        if m is not None:
            p.m = m
        p.tile_log2 = tile_log2
        p.fft_w, p.fft_h = fft_w, fft_h
        p.flags = ((_lib.CUNA_PLAN_WINDOW2 if window2 else 0)
                   | (_lib.CUNA_PLAN_UNFUSED if unfused else 0)
                   | (_lib.CUNA_PLAN_GATHER if gather else 0))
        h = ctypes.c_void_p()
        s = _stream(stream)
        if t is not None:
            if t.dtype != torch.float64:
                raise TypeError("times must be float64")
            if offsets is not None:
                off = np.ascontiguousarray(offsets, dtype=np.int32)
                st = self.L.cuna_nfft_plan_create_batch_times(
                    ctypes.byref(h), _ptr(t), off.ctypes.data, len(off) - 1, float(df), M,
                    ctypes.addressof(p), s)
            else:
                st = self.L.cuna_nfft_plan_create_times(ctypes.byref(h), _ptr(t), t.numel(),
                                                        float(df), M, ctypes.addressof(p), s)
            self.n = t.numel()
        elif offsets is not None:
            off = np.ascontiguousarray(offsets, dtype=np.int32)
            st = self.L.cuna_nfft_plan_create_batch(ctypes.byref(h), _ptr(x), off.ctypes.data,
                                                    len(off) - 1, M, ctypes.addressof(p), s)
            self.n = int(off[-1])
        else:
            st = self.L.cuna_nfft_plan_create(ctypes.byref(h), _ptr(x), x.numel(), M,
                                              ctypes.addressof(p), s)
            self.n = x.numel()
        check(self.L, st)
        self.h = h.value
        self.params = p
        info = self.info()
        self.M, self.G, self.ncurves = info.M, info.G, info.ncurves
        self.mode = mode

    def info(self) -> _lib.NfftInfo:
        inf = _lib.NfftInfo()
        check(self.L, self.L.cuna_nfft_plan_info(self.h, ctypes.addressof(inf)))
        return inf

    def _out(self, rows: int):
        import torch

        return torch.empty((rows, self.M), dtype=complex_dtype(self.prec), device="cuda")

    def adjoint(self, y=None, out=None, stream=None):
        if out is None:
            out = self._out(self.ncurves)
        check(self.L, self.L.cuna_nfft_adjoint(self.h, _ptr(y), _ptr(out), _stream(stream)))
        return out

    def adjoint_many(self, y, out=None, stream=None):
        """y: (nrhs, n) values; returns (ncurves * nrhs, M), row c * nrhs + r
        the spectrum of right-hand side r for curve c."""
        nrhs = int(y.shape[0])
        if out is None:
            out = self._out(self.ncurves * nrhs)
        check(self.L, self.L.cuna_nfft_adjoint_many(self.h, _ptr(y), nrhs, _ptr(out),
                                                    _stream(stream)))
        return out

    def bootstrap(self, y, nbs: int, seed: int = 0, window: bool = False, stream=None,
                  out_y=None, out_w=None):
        if out_y is None:
            out_y = self._out(self.ncurves * nbs)
        if window and out_w is None:
            out_w = self._out(self.ncurves * nbs)
        if not window:
            out_w = None
        check(self.L, self.L.cuna_nfft_bootstrap(self.h, _ptr(y), nbs, seed & 0xFFFFFFFF,
                                                  _ptr(out_y), _ptr(out_w), _stream(stream)))
        return (out_y, out_w) if window else out_y

    def bootstrap_ls(self, y, nbs: int, seed: int = 0, power: bool = True, peak: bool = True,
                     stream=None, out_power=None, out_peak=None, out_kpeak=None):
        """Lomb-Scargle periodograms of the bootstrap replicates (plan made
        with window2=True): power (ncurves*nbs, M/2) for k = 0..M/2-1 and/or
        peak (ncurves*nbs,) = max over k >= 1 with kpeak its k."""
        import torch

        rows = self.ncurves * nbs
        rd = real_dtype(self.prec)
        if power and out_power is None:
            out_power = torch.empty((rows, self.M // 2), dtype=rd, device="cuda")
        if peak and out_peak is None:
            out_peak = torch.empty(rows, dtype=rd, device="cuda")
        if peak and out_kpeak is None:
            out_kpeak = torch.empty(rows, dtype=torch.int32, device="cuda")
        check(self.L, self.L.cuna_nfft_bootstrap_ls(
            self.h, _ptr(y), nbs, seed & 0xFFFFFFFF, _ptr(out_power if power else None),
            _ptr(out_peak if peak else None), _ptr(out_kpeak if peak else None),
            _stream(stream)))
        return {"power": out_power if power else None, "peak": out_peak if peak else None,
                "kpeak": out_kpeak if peak else None}

    def forward(self, fhat, out=None, stream=None):
        """Forward (type-2) NFFT: f_j = sum_k fhat_k e^{-2 pi i k x_j} with fhat
        of shape (ncurves, M) in the adjoint's layout; returns (n,) complex."""
        import torch

        if out is None:
            out = torch.empty(self.n, dtype=complex_dtype(self.prec), device="cuda")
        fh = fhat.contiguous()
        check(self.L, self.L.cuna_nfft_forward(self.h, _ptr(fh), _ptr(out), _stream(stream)))
        return out

    def backward(self, y, grad_out, want_y: bool = True, want_x: bool = True, stream=None):
        import torch

        rd = real_dtype(self.prec)
        dy = torch.empty(self.n, dtype=rd, device="cuda") if want_y else None
        dx = torch.empty(self.n, dtype=rd, device="cuda") if want_x else None
        g = grad_out.contiguous()
        check(self.L, self.L.cuna_nfft_backward(self.h, _ptr(y), _ptr(g), _ptr(dy), _ptr(dx),
                                                 _stream(stream)))
        return dy, dx

    def close(self):
        if getattr(self, "h", None):
            self.L.cuna_nfft_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(x, M: int, tol: float, cache_path: str | None = None, stream=None) -> dict:
    """cuna_nfft_autotune: plan parameters (sigma, m, tile_log2, fft_w, fft_h)
    for points x meeting `tol` against the direct NDFT, fastest on this GPU.
    Pass the result as keyword arguments to NFFTPlan."""
    prec = _prec_of(x)
    L = load(prec)
    p = _lib.NfftParams()
    cp = None if cache_path is None else cache_path.encode()
    check(L, L.cuna_nfft_autotune(_ptr(x), x.numel(), M, float(tol), cp, ctypes.addressof(p),
                                  _stream(stream)))
    return {"sigma": p.sigma, "m": p.m, "tile_log2": p.tile_log2, "fft_w": p.fft_w,
            "fft_h": p.fft_h}


class Graph:
    """A captured sequence of library calls (cuna_stream_capture_*)."""

    def __init__(self, L, handle):
        self.L, self.h = L, handle

    def launch(self, stream=None):
        check(self.L, self.L.cuna_graph_launch(self.h, _stream(stream)))

    def close(self):
        if self.h:
            self.L.cuna_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def capture(fn, stream, prec: str = "f") -> Graph:
    """Record the library calls made by fn() on `stream` (a torch.cuda.Stream,
    not the null stream) into a graph.  Run fn() once before capturing so that
    its workspaces exist, and pass preallocated outputs."""
    L = load(prec)
    s = _stream(stream)
    check(L, L.cuna_stream_capture_begin(s))
    g = ctypes.c_void_p()
    try:
        fn()
    except BaseException:
        L.cuna_stream_capture_end(s, ctypes.byref(g))
        if g.value:
            L.cuna_graph_destroy(g.value)
        raise
    check(L, L.cuna_stream_capture_end(s, ctypes.byref(g)))
    return Graph(L, g.value)


def fft(data, sign: int = +1, stream=None):
    """In-place unnormalised DFT along the last axis (any length)."""
    import torch

    prec = "f" if data.dtype == torch.complex64 else "d"
    L = load(prec)
    n = data.shape[-1]
    check(L, L.cuna_fft(_ptr(data), n, data.numel() // n, sign, _stream(stream)))
    return data


__all__ = [
    "CunaError", "cunfft_adjoint", "FilterProperties", "generate_filter_properties",
    "generate_pinned_filter_properties", "cunfft_adjoint_raw", "cunfft_adjoint_raw_async",
    "cunfft_adjoint_raw_async_bootstrap", "cunfft_adjoint_raw_async_noalloc", "CunaPlan",
    "NFFTPlan", "fft", "Graph", "capture",
    "autotune",
]
