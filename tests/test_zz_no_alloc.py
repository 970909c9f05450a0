"""The allocation-free compat transform cunfft_adjoint_raw_async_noalloc
(include/cuna.h), through the C-ABI: the same output as
cunfft_adjoint_raw_async on two calls in a row (the second smaller), in both
libraries; no change in device memory use; oversize and unregistered inputs
rejected before any device work; a graph capture of the call replays."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from cunfft_adjoint_amd import _lib  # noqa: E402
from tests._data import rel_l2, uniform_points  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"f": np.float32, "d": np.float64}
TOL = {"f": 1e-5, "d": 1e-12}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return h


def _free_bytes(h):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("prec", ["f", "d"])
def test_noalloc_matches_raw_async_two_calls(prec):
    L = cu.load(prec)
    cdt = torch.complex64 if prec == "f" else torch.complex128
    view = np.uint32 if prec == "f" else np.uint64
    s = torch.cuda.Stream()
    cases = []
    for n, ng, seed in ((100_000, 1 << 18, 1), (3000, 1 << 12, 2)):  # second call smaller
        x, y = uniform_points(n, seed, NP[prec])
        xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
        fp = cu.generate_filter_properties(xd, n, ng)
        ref = torch.zeros(ng, dtype=cdt, device="cuda")
        cu.cunfft_adjoint_raw_async(xd, yd, ref, n, ng, fp, stream=s)
        s.synchronize()
        cases.append((x, y, xd, yd, fp, ref, n, ng))
    h = _hip()
    outs = []
    for x, y, xd, yd, fp, ref, n, ng in cases:
        out = torch.full((ng,), 7.0, dtype=cdt, device="cuda")
        outs.append(out)
    torch.cuda.synchronize()
    free0 = _free_bytes(h)
    for (x, y, xd, yd, fp, ref, n, ng), out in zip(cases, outs):
        L.cunfft_adjoint_raw_async_noalloc(xd.data_ptr(), yd.data_ptr(), out.data_ptr(), n, ng,
                                           fp.device, s.cuda_stream)
        assert L.cuna_get_last_error() == 0, L.cuna_get_last_error_string()
    s.synchronize()
    assert _free_bytes(h) == free0  # nothing allocated on the device
    for (x, y, xd, yd, fp, ref, n, ng), out in zip(cases, outs):
        got = out.cpu().numpy()
        assert np.array_equal(got.view(view), ref.cpu().numpy().view(view)), (n, ng)
        assert rel_l2(got, oracle.compat_adjoint(x, y, ng, prec == "d")) < TOL[prec]
    # unit weights (f_data NULL) as in the reference
    x, y, xd, yd, fp, ref, n, ng = cases[1]
    out = torch.zeros(ng, dtype=cdt, device="cuda")
    L.cunfft_adjoint_raw_async_noalloc(xd.data_ptr(), None, out.data_ptr(), n, ng, fp.device,
                                       s.cuda_stream)
    s.synchronize()
    assert rel_l2(out.cpu().numpy(), oracle.compat_adjoint(x, None, ng, prec == "d")) < TOL[prec]
    for c in cases:
        c[4].free()


@pytest.mark.parametrize("prec", ["f", "d"])
def test_noalloc_rejects_before_device_work(prec):
    L = cu.load(prec)
    n, ng = 2000, 1 << 12
    x, y = uniform_points(n, 3, NP[prec])
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    cdt = torch.complex64 if prec == "f" else torch.complex128
    out = torch.zeros(ng, dtype=cdt, device="cuda")
    # not registered (no generate_* call for this x): no plan may be built here
    L.cunfft_adjoint_raw_async_noalloc(xd.data_ptr(), yd.data_ptr(), out.data_ptr(), n, ng, None,
                                       None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE
    # beyond the scratch: 2^23 cells (float) / 2^22 (double)
    big = 1 << (23 if prec == "f" else 22)
    L.cunfft_adjoint_raw_async_noalloc(xd.data_ptr(), yd.data_ptr(), out.data_ptr(), n, big, None,
                                       None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_UNSUPPORTED
    assert b"scratch" in L.cuna_get_last_error_string()
    # filter properties of another size
    fp = cu.generate_filter_properties(xd, n, ng)
    L.cunfft_adjoint_raw_async_noalloc(xd.data_ptr(), yd.data_ptr(), out.data_ptr(), n, ng // 2,
                                       fp.device, None)
    assert L.cuna_get_last_error() == _lib.CUNA_ERROR_INVALID_VALUE
    torch.cuda.synchronize()
    assert np.all(out.cpu().numpy() == 0)
    fp.free()


def test_noalloc_graph_replays():
    """Captured into a graph (no allocation, no memset node), replays equal
    the eager call bitwise, with other library work in between."""
    n, ng = 50_000, 1 << 16
    x, y = uniform_points(n, 4)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    fp = cu.generate_filter_properties(xd, n, ng)
    s = torch.cuda.Stream()
    out = torch.zeros(ng, dtype=torch.complex64, device="cuda")
    run = lambda: cu.cunfft_adjoint_raw_async_noalloc(xd, yd, out, n, ng, fp, stream=s)
    run()
    s.synchronize()
    eager = out.cpu().numpy().copy()
    g = cu.capture(run, s)
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        assert np.array_equal(out.cpu().numpy().view(np.uint32), eager.view(np.uint32)), rep
        other = torch.zeros_like(out)
        cu.cunfft_adjoint_raw_async(xd, yd * 2, other, n, ng, fp, stream=s)
        s.synchronize()
    g.close()
    fp.free()
