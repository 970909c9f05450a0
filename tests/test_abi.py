"""C-ABI checks that need no GPU: exported symbols, struct layouts, headers
usable from C and C++, and the host-side argument validation (which runs
before any HIP call)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from cunfft_adjoint_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(INCLUDE, "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"#[^\n]*", "", src)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}()]*(\([^()]*\)[^;{}()]*)*)\)\s*;", src):
            name = m.group(1)
            if name in {"if", "while", "for", "return", "sizeof"}:
                continue
            names.add(name)
    return names


@pytest.fixture(scope="module", params=["f", "d"])
def lib(request):
    return request.param, _lib.load(request.param)


def test_every_declared_function_is_exported(lib):
    prec, L = lib
    declared = _declared_functions()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.lib_path(prec)],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = declared - exported
    assert not missing, missing
    # and the Python mirror binds exactly those entry points
    assert declared == set(_lib.exported_functions(prec))
    for name in declared:
        getattr(L, name)


def test_precision_of_each_library(lib):
    prec, L = lib
    assert L.cuna_real_size() == (4 if prec == "f" else 8)
    assert L.cuna_filter_radius() == (6 if prec == "f" else 12)


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "cuna.h"
#include "cuna_filter.h"
#include "cuna_nfft.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu\n", sizeof(filter_properties),
           offsetof(filter_properties, normfac), offsetof(filter_properties, filter_radius),
           offsetof(filter_properties, E1), offsetof(filter_properties, E3), sizeof(Complex));
    printf("%zu %zu %zu %zu %zu\n", sizeof(plan), offsetof(plan, flags), offsetof(plan, Ngrid),
           offsetof(plan, fprops_host), offsetof(plan, fprops_device));
    printf("%zu %zu %zu\n", sizeof(cuna_nfft_params), sizeof(cuna_nfft_info),
           offsetof(cuna_nfft_info, beta));
    printf("%d %d %d %d %d\n", NO_FLAGS, OUTPUT_INTERMEDIATE, CALCULATE_WINDOW_FUNCTION,
           DONT_TRANSFER_TO_CPU, PRINT_TIMING);
    return 0;
}
"""


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("lang", ["c", "c++"])
def test_headers_compile_and_layouts(tmp_path, prec, lang):
    src = tmp_path / ("layout.c" if lang == "c" else "layout.cpp")
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    cc = "gcc" if lang == "c" else "g++"
    cmd = [cc, "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", "-I/opt/rocm/include", str(src), "-o",
           str(exe)]
    if lang == "c":
        cmd.insert(1, "-std=c99")
    if prec == "d":
        cmd.insert(1, "-DDOUBLE_PRECISION")
    subprocess.run(cmd, check=True, capture_output=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    R, Cx, FP, PlanT = _lib.make_types(prec)
    fp = [int(v) for v in lines[0].split()]
    assert fp == [ctypes.sizeof(FP), FP.normfac.offset, FP.filter_radius.offset, FP.E1.offset,
                  FP.E3.offset, ctypes.sizeof(Cx)]
    # the reference's field order (inc/cuna_typedefs.h:92-108): b, normfac,
    # filter_radius, E1, E2, E3
    rs = ctypes.sizeof(R)
    assert fp[1] == rs and fp[2] == 2 * rs and fp[5] == 2 * rs
    pl = [int(v) for v in lines[1].split()]
    assert pl == [ctypes.sizeof(PlanT), PlanT.flags.offset, PlanT.Ngrid.offset,
                  PlanT.fprops_host.offset, PlanT.fprops_device.offset]
    nf = [int(v) for v in lines[2].split()]
    assert nf == [ctypes.sizeof(_lib.NfftParams), ctypes.sizeof(_lib.NfftInfo),
                  _lib.NfftInfo.beta.offset]
    assert [int(v) for v in lines[3].split()] == [1, 2, 4, 8, 16]


def test_validation_errors_without_gpu(lib):
    prec, L = lib
    h = ctypes.c_void_p()
    p = _lib.NfftParams()
    L.cuna_nfft_default_params(ctypes.addressof(p))
    assert (p.mode, p.sigma, p.m) == (0, 2, 8 if prec == "f" else 13)
    # M not a power of two
    st = L.cuna_nfft_plan_create(ctypes.byref(h), None, 0, 3, ctypes.addressof(p), None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE
    assert b"power of two" in L.cuna_get_last_error_string()
    assert L.cuna_get_last_error() == st
    # m out of range
    p.m = 40
    st = L.cuna_nfft_plan_create(ctypes.byref(h), None, 0, 64, ctypes.addressof(p), None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE
    p.m = 8
    # unknown mode, bad tile
    p.mode = 7
    assert L.cuna_nfft_plan_create(ctypes.byref(h), None, 0, 64, ctypes.addressof(p),
                                   None) == _lib.CUNA_ERROR_INVALID_VALUE
    p.mode = 0
    p.tile_log2 = 30
    assert L.cuna_nfft_plan_create(ctypes.byref(h), None, 0, 64, ctypes.addressof(p),
                                   None) == _lib.CUNA_ERROR_INVALID_VALUE
    p.tile_log2 = 0
    # compat mode: ng beyond the FFT, n >= 2 ng (b would be negative), NULL x
    # (any ng >= 1 of a supported size is accepted, as cuFFT takes any length)
    p.mode = 1
    st = L.cuna_nfft_plan_create(ctypes.byref(h), None, 10, 1 << 27, ctypes.addressof(p), None)
    assert st == _lib.CUNA_ERROR_UNSUPPORTED
    st = L.cuna_nfft_plan_create(ctypes.byref(h), None, 10, 100, ctypes.addressof(p), None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE and b"NULL" in L.cuna_get_last_error_string()
    st = L.cuna_nfft_plan_create(ctypes.byref(h), None, 64, 32, ctypes.addressof(p), None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE and b"2*ng" in L.cuna_get_last_error_string()
    st = L.cuna_nfft_plan_create(ctypes.byref(h), None, 10, 64, ctypes.addressof(p), None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE
    # times plans: df must be positive
    p.mode = 0
    st = L.cuna_nfft_plan_create_times(ctypes.byref(h), None, 0, -1.0, 64, ctypes.addressof(p),
                                       None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE
    # batches: decreasing offsets
    off = (ctypes.c_int * 3)(0, 5, 3)
    st = L.cuna_nfft_plan_create_batch(ctypes.byref(h), None, off, 2, 64, ctypes.addressof(p),
                                       None)
    assert st == _lib.CUNA_ERROR_INVALID_VALUE
    # FFT stage: a length whose Bluestein convolution exceeds the FFT; bad
    # sign; NULL data
    st = L.cuna_fft(ctypes.c_void_p(16), (1 << 24) + 1, 1, 1, None)
    assert st == _lib.CUNA_ERROR_UNSUPPORTED
    assert L.cuna_fft(ctypes.c_void_p(16), 16, 1, 2, None) == _lib.CUNA_ERROR_INVALID_VALUE
    assert L.cuna_fft(None, 16, 1, 1, None) == _lib.CUNA_ERROR_INVALID_VALUE
    # reference entry points record errors instead of exiting
    L.cunfft_adjoint_raw(None, None, None, 10, 64, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE
    L.cuna_clear_last_error()
    assert L.cuna_get_last_error() == 0
    R = _lib.real_type(prec)
    L.generate_filter_properties(None, 0, 64, None, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE
    x = (R * 4)(0.0, 0.1, 0.2, 0.3)
    out = (R * 8)()
    L.cunfft_adjoint(x, None, out, 4, 2, 0)  # n = 2 ng: b would not be positive
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE
    # batched compat API: sizes are checked before any device work
    fake = ctypes.c_void_p(256)
    n2, ng2 = (ctypes.c_int * 2)(5, 5), (ctypes.c_int * 2)(16, 16)
    L.cunfft_adjoint_raw_batched_async(fake, None, None, n2, ng2, 2, None, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE  # f_hat NULL
    L.cunfft_adjoint_raw_batched_async(fake, None, fake, n2, ng2, 0, None, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE  # no curves
    L.cunfft_adjoint_raw_batched_async(fake, None, fake, n2, (ctypes.c_int * 2)(16, 1 << 27),
                                       2, None, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_UNSUPPORTED  # ng = 2^27
    L.cunfft_adjoint_raw_batched_async(fake, None, fake, (ctypes.c_int * 2)(5, 40), ng2, 2,
                                       None, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE  # n >= 2 ng
    assert b"2*ng" in L.cuna_get_last_error_string()
    L.generate_pinned_filter_properties_batch(fake, n2, ng2, 2, None, None, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE
    L.cuna_clear_last_error()


def test_abort_on_error_is_opt_in(tmp_path):
    # with abort-on-error the process exits like the reference's checkCudaErrors
    code = ("import ctypes,sys; sys.path.insert(0, %r); from cunfft_adjoint_amd import _lib; "
            "L=_lib.load('f'); L.cuna_set_abort_on_error(1); "
            "h=ctypes.c_void_p(); L.cuna_nfft_plan_create(ctypes.byref(h), None, 0, 3, None, None); "
            "print('not reached')") % ROOT
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True)
    assert r.returncode == _lib.CUNA_ERROR_INVALID_VALUE
    assert "not reached" not in r.stdout
    assert "CUNA ERROR" in r.stderr
