"""GPU parity of the HIP library (through its C-ABI) against the CPU oracle.

Tolerances: compat mode vs the oracle's restatement of the reference formula
rel L2 <= 1e-5 (libcunaf) / 1e-12 (libcunad); accurate mode vs the FP64 direct
NDFT rel L2 <= 1e-5 (libcunaf, m = 8) / 1e-12 (libcunad, m = 13).
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from cunfft_adjoint_amd.autograd import adjoint_nfft  # noqa: E402
from tests._data import clustered_points, rel_l2, uniform_points  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = {"f": 1e-5, "d": 1e-12}
NP = {"f": np.float32, "d": np.float64}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    # the product must be the HIP library: both .so must load
    cu.load("f")
    cu.load("d")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# ------------------------------------------------------------ FFT stage --
@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("lg", [1, 2, 3, 4, 5, 8, 10, 12, 13, 14, 15, 16, 18, 20])
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_stage_vs_numpy(prec, lg, sign):
    n = 1 << lg
    batch = 3 if lg <= 16 else 1
    rng = np.random.default_rng(lg)
    z = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n)))
    z = z.astype(np.complex64 if prec == "f" else np.complex128)
    t = dev(z)
    cu.fft(t, sign=sign)
    got = host(t)
    ref = np.fft.ifft(z.astype(np.complex128), axis=-1) * n if sign > 0 else np.fft.fft(
        z.astype(np.complex128), axis=-1)
    assert rel_l2(got, ref) < (2e-6 if prec == "f" else 1e-13)


# --------------------------------------------------------- compat mode --
@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("n,ng", [(10000, 1 << 15), (1000, 1 << 10), (7, 16), (1, 4), (300, 256)])
def test_compat_raw_vs_oracle(prec, n, ng):
    x, y = uniform_points(n, 11 + n, NP[prec])
    xd, yd = dev(x), dev(y)
    fp = cu.generate_filter_properties(xd, n, ng)
    f_hat = torch.zeros(ng, dtype=torch.complex64 if prec == "f" else torch.complex128,
                        device="cuda")
    cu.cunfft_adjoint_raw(xd, yd, f_hat, n, ng, fp)
    got = host(f_hat)
    ref = oracle.compat_adjoint(x, y, ng, prec == "d")
    assert rel_l2(got, ref) < TOL[prec]
    # filter constants and E1/E2/E3 as src/gaussian_filter.cu:54-57,204-228
    b, nf, rf = oracle.compat_params(n, ng, prec == "d")
    fl = fp.fields
    assert fl.filter_radius == rf
    assert fl.b == pytest.approx(b, rel=1e-7 if prec == "f" else 1e-15)
    assert fl.normfac == pytest.approx(nf, rel=1e-7 if prec == "f" else 1e-15)
    R = NP[prec]
    e1 = np.empty(n, R)
    e2 = np.empty(n, R)
    e3 = np.empty(rf, R)
    torch.cuda.synchronize()
    _copy_d2h(e1, fl.E1)
    _copy_d2h(e2, fl.E2)
    _copy_d2h(e3, fl.E3)
    eps = np.array([oracle.compat_point(xi, n, ng, prec == "d")[1] for xi in x[:200]])
    tol = 2e-6 if prec == "f" else 1e-13
    np.testing.assert_allclose(e1[:200], np.exp(-eps * eps / b) / np.sqrt(np.pi * b), rtol=tol)
    np.testing.assert_allclose(e2[:200], np.exp(2 * eps / b), rtol=tol)
    np.testing.assert_allclose(e3, np.exp(-np.arange(rf) ** 2 / b), rtol=tol)
    fp.free()


def _copy_d2h(dst, ptr):
    import ctypes
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert lib.hipMemcpy(dst.ctypes.data, ptr, dst.nbytes, 2) == 0


@pytest.mark.parametrize("prec", ["f", "d"])
def test_compat_async_and_bootstrap_vs_oracle(prec):
    n, ng, nbs, seed = 3000, 1 << 13, 5, 12345
    x, y = uniform_points(n, 3, NP[prec])
    xd, yd = dev(x), dev(y)
    fp = cu.generate_filter_properties(xd, n, ng)
    s = torch.cuda.Stream()
    cdt = torch.complex64 if prec == "f" else torch.complex128
    f_hat = torch.zeros(ng, dtype=cdt, device="cuda")
    cu.cunfft_adjoint_raw_async(xd, yd, f_hat, n, ng, fp, stream=s)
    s.synchronize()
    assert rel_l2(host(f_hat), oracle.compat_adjoint(x, y, ng, prec == "d")) < TOL[prec]
    fb = torch.zeros(nbs * ng, dtype=cdt, device="cuda")
    cu.cunfft_adjoint_raw_async_bootstrap(xd, yd, fb, n, ng, nbs, fp, stream=s, seed=seed)
    s.synchronize()
    ref = oracle.compat_adjoint_bootstrap(x, y, ng, nbs, seed, prec == "d")
    assert rel_l2(host(fb).reshape(nbs, ng), ref) < TOL[prec]
    # f_data == NULL: unit weights, no draw (src/gridding.cu:98-99)
    fb.zero_()
    cu.cunfft_adjoint_raw_async_bootstrap(xd, None, fb, n, ng, nbs, fp, stream=s, seed=seed)
    s.synchronize()
    ref = oracle.compat_adjoint_bootstrap(x, None, ng, nbs, seed, prec == "d")
    assert rel_l2(host(fb).reshape(nbs, ng), ref) < TOL[prec]
    fp.free()


@pytest.mark.parametrize("prec", ["f", "d"])
def test_compat_host_api_and_plan_path(prec):
    n, ng = 2000, 1 << 12
    x, y = uniform_points(n, 5, NP[prec])
    got = cu.cunfft_adjoint(x, y, ng)
    ref = oracle.compat_adjoint(x, y, ng, prec == "d")
    assert rel_l2(got, ref) < TOL[prec]
    plan = cu.CunaPlan(x, y, ng, cu._lib.CALCULATE_WINDOW_FUNCTION)
    plan.execute()
    assert rel_l2(plan.f_hat(), ref) < TOL[prec]
    assert rel_l2(plan.f_hat_win(), oracle.compat_adjoint(x, None, ng, prec == "d")) < TOL[prec]
    plan.free()


def test_compat_unregistered_fprops_rebuilds():
    n, ng = 500, 1 << 10
    x, y = uniform_points(n, 9)
    xd, yd = dev(x), dev(y)
    f_hat = torch.zeros(ng, dtype=torch.complex64, device="cuda")
    cu.cunfft_adjoint_raw(xd, yd, f_hat, n, ng, 0)
    assert rel_l2(host(f_hat), oracle.compat_adjoint(x, y, ng, False)) < 1e-5


# ---------------------------------------------- batches of compat curves --
BATCH_N = [1, 50, 1000, 3000, 7, 200, 5000]
BATCH_NG = [16, 1024, 4096, 4096, 16, 256, 8192]


def _hip_malloc(nbytes):
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p.value


@pytest.mark.parametrize("prec", ["f", "d"])
def test_compat_batched_vs_single_curves(prec):
    """cunfft_adjoint_raw_batched_async (the reference's commented-out batched
    API): each curve's slab equals the single-curve transform and the
    oracle, the rest of each slab is untouched, unit weights work, and the
    cached multi-curve plan gives the same bits on a second call."""
    rng = np.random.default_rng(31)
    n, ng = BATCH_N, BATCH_NG
    nlc, idist = len(n), max(ng)
    off = np.concatenate([[0], np.cumsum(n)])
    x = rng.uniform(-0.5, 0.5, off[-1]).astype(NP[prec])
    y = rng.standard_normal(off[-1]).astype(NP[prec])
    xd, yd = dev(x), dev(y)
    cdt = torch.complex64 if prec == "f" else torch.complex128
    view = np.uint32 if prec == "f" else np.uint64
    fps = [cu.generate_filter_properties(xd[off[i]:off[i + 1]], n[i], ng[i]) for i in range(nlc)]
    s = torch.cuda.Stream()
    sentinel = 3.0 + 4.0j
    for f_data, yh in ((yd, y), (None, None)):
        fh = torch.full((nlc * idist,), sentinel, dtype=cdt, device="cuda")
        cu.cunfft_adjoint_raw_batched_async(xd, f_data, fh, n, ng, fps, stream=s)
        s.synchronize()
        got = host(fh).reshape(nlc, idist)
        for i in range(nlc):
            sl = slice(off[i], off[i + 1])
            single = torch.zeros(ng[i], dtype=cdt, device="cuda")
            cu.cunfft_adjoint_raw_async(xd[sl], None if f_data is None else yd[sl], single, n[i],
                                        ng[i], fps[i], stream=s)
            s.synchronize()
            assert rel_l2(got[i, :ng[i]], host(single)) < (1e-6 if prec == "f" else 1e-14), i
            ref = oracle.compat_adjoint(x[sl], None if yh is None else yh[sl], ng[i], prec == "d")
            assert rel_l2(got[i, :ng[i]], ref) < TOL[prec], i
            assert np.all(got[i, ng[i]:] == sentinel), i
    # second call: the cached plan, same bits
    a = torch.zeros(nlc * idist, dtype=cdt, device="cuda")
    b = torch.zeros_like(a)
    cu.cunfft_adjoint_raw_batched_async(xd, yd, a, n, ng, fps, stream=s)
    cu.cunfft_adjoint_raw_batched_async(xd, yd, b, n, ng, fps, stream=s)
    s.synchronize()
    assert np.array_equal(host(a).view(view), host(b).view(view))
    # freeing one curve's filter properties drops the cached batch; without
    # filter properties the plan is rebuilt, with the same result
    fps[0].free()
    c = torch.zeros_like(a)
    cu.cunfft_adjoint_raw_batched_async(xd, yd, c, n, ng, None, stream=s)
    s.synchronize()
    assert np.array_equal(host(a).view(view), host(c).view(view))
    for f in fps[1:]:
        f.free()
    cu.batch_cache_clear(prec)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_pinned_filter_properties_batch(prec):
    """generate_pinned_filter_properties_batch fills each curve's b, normfac,
    R_f and E1/E2/E3 as src/gaussian_filter.cu:54-57,204-228 define them, and
    its structs drive the batched transform."""
    rng = np.random.default_rng(32)
    n, ng = [300, 20, 4000], [1024, 64, 4096]
    nlc = len(n)
    off = np.concatenate([[0], np.cumsum(n)])
    dt = NP[prec]
    x = rng.uniform(-0.5, 0.5, off[-1]).astype(dt)
    y = rng.standard_normal(off[-1]).astype(dt)
    xd, yd = dev(x), dev(y)
    L = cu.load(prec)
    _, _, FP, _ = cu._lib.make_types(prec)
    rf = L.cuna_filter_radius()
    tdt = torch.float32 if prec == "f" else torch.float64
    E1 = [torch.empty(k, dtype=tdt, device="cuda") for k in n]
    E2 = [torch.empty(k, dtype=tdt, device="cuda") for k in n]
    E3 = [torch.empty(rf, dtype=tdt, device="cuda") for _ in n]
    h = [FP(0, 0, 0, E1[i].data_ptr(), E2[i].data_ptr(), E3[i].data_ptr()) for i in range(nlc)]
    d = [_hip_malloc(ctypes.sizeof(FP)) for _ in range(nlc)]
    s = torch.cuda.Stream()
    cu.generate_pinned_filter_properties_batch(xd, n, ng, h, d, stream=s)
    s.synchronize()
    rtol = 2e-6 if prec == "f" else 1e-13
    for i in range(nlc):
        b, nf, r = oracle.compat_params(n[i], ng[i], prec == "d")
        assert h[i].filter_radius == r == rf
        assert h[i].b == pytest.approx(b, rel=1e-7 if prec == "f" else 1e-15)
        assert h[i].normfac == pytest.approx(nf, rel=1e-7 if prec == "f" else 1e-15)
        eps = np.array([oracle.compat_point(v, n[i], ng[i], prec == "d")[1]
                        for v in x[off[i]:off[i + 1]]])
        np.testing.assert_allclose(host(E1[i]), np.exp(-eps * eps / b) / np.sqrt(np.pi * b),
                                   rtol=rtol)
        np.testing.assert_allclose(host(E2[i]), np.exp(2 * eps / b), rtol=rtol)
        np.testing.assert_allclose(host(E3[i]), np.exp(-np.arange(rf) ** 2 / b), rtol=rtol)
    idist = max(ng)
    cdt = torch.complex64 if prec == "f" else torch.complex128
    fh = torch.zeros(nlc * idist, dtype=cdt, device="cuda")
    cu.cunfft_adjoint_raw_batched_async(xd, yd, fh, n, ng, d, stream=s)
    s.synchronize()
    got = host(fh).reshape(nlc, idist)
    for i in range(nlc):
        sl = slice(off[i], off[i + 1])
        ref = oracle.compat_adjoint(x[sl], y[sl], ng[i], prec == "d")
        assert rel_l2(got[i, :ng[i]], ref) < TOL[prec], i
    for p in d:
        L.free_filter_properties(p, None)
    cu.batch_cache_clear(prec)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_pinned_filter_properties_single(prec):
    """generate_pinned_filter_properties (inc/cuna_filter.h:37-40): the caller
    owns the host struct, its device E1/E2/E3 arrays and the device struct;
    the call fills them on a stream (src/gaussian_filter.cu:43-66), and the
    raw transform then matches the oracle."""
    rng = np.random.default_rng(33)
    n, ng = 2500, 1 << 12
    x = rng.uniform(-0.5, 0.5, n).astype(NP[prec])
    y = rng.standard_normal(n).astype(NP[prec])
    xd, yd = dev(x), dev(y)
    L = cu.load(prec)
    _, _, FP, _ = cu._lib.make_types(prec)
    rf = L.cuna_filter_radius()
    tdt = torch.float32 if prec == "f" else torch.float64
    E1, E2, E3 = (torch.empty(k, dtype=tdt, device="cuda") for k in (n, n, rf))
    h = FP(0, 0, 0, E1.data_ptr(), E2.data_ptr(), E3.data_ptr())
    d = _hip_malloc(ctypes.sizeof(FP))
    s = torch.cuda.Stream()
    cu.generate_pinned_filter_properties(xd, n, ng, h, d, stream=s)
    s.synchronize()
    b, nf, r = oracle.compat_params(n, ng, prec == "d")
    assert h.filter_radius == r == rf
    assert h.b == pytest.approx(b, rel=1e-7 if prec == "f" else 1e-15)
    eps = np.array([oracle.compat_point(v, n, ng, prec == "d")[1] for v in x[:300]])
    rtol = 2e-6 if prec == "f" else 1e-13
    np.testing.assert_allclose(host(E1)[:300], np.exp(-eps * eps / b) / np.sqrt(np.pi * b),
                               rtol=rtol)
    np.testing.assert_allclose(host(E3), np.exp(-np.arange(rf) ** 2 / b), rtol=rtol)
    fh = torch.zeros(ng, dtype=torch.complex64 if prec == "f" else torch.complex128,
                     device="cuda")
    cu.cunfft_adjoint_raw_async(xd, yd, fh, n, ng, d, stream=s)
    s.synchronize()
    assert rel_l2(host(fh), oracle.compat_adjoint(x, y, ng, prec == "d")) < TOL[prec]
    L.free_filter_properties(d, None)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_pinned_filter_properties_does_not_wait(prec):
    """generate_pinned_filter_properties is ordered on its stream and does not
    wait for it, as src/gaussian_filter.cu:43-66 is: enqueued behind a bounded
    spin kernel on the same stream, it returns while that stream is still busy
    (hipStreamQuery: not ready), and the transform ordered after it matches
    the oracle."""
    import time
    rng = np.random.default_rng(34)
    n, ng = 3000, 1 << 12
    x = rng.uniform(-0.5, 0.5, n).astype(NP[prec])
    y = rng.standard_normal(n).astype(NP[prec])
    xd, yd = dev(x), dev(y)
    L = cu.load(prec)
    _, _, FP, _ = cu._lib.make_types(prec)
    rf = L.cuna_filter_radius()
    tdt = torch.float32 if prec == "f" else torch.float64
    warm = cu.generate_filter_properties(xd, n, ng)  # this size's FFT tables exist now
    warm.free()
    E1, E2, E3 = (torch.empty(k, dtype=tdt, device="cuda") for k in (n, n, rf))
    h = FP(0, 0, 0, E1.data_ptr(), E2.data_ptr(), E3.data_ptr())
    d = _hip_malloc(ctypes.sizeof(FP))
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamQuery.argtypes = [ctypes.c_void_p]
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        torch.cuda._sleep(200_000_000)  # bounded spin, ~0.1 s
    t0 = time.perf_counter()
    cu.generate_pinned_filter_properties(xd, n, ng, h, d, stream=s)
    elapsed = time.perf_counter() - t0
    busy = hip.hipStreamQuery(ctypes.c_void_p(s.cuda_stream))
    assert busy == 600, (busy, elapsed)  # hipErrorNotReady: the spin still runs
    fh = torch.zeros(ng, dtype=torch.complex64 if prec == "f" else torch.complex128,
                     device="cuda")
    cu.cunfft_adjoint_raw_async(xd, yd, fh, n, ng, d, stream=s)
    s.synchronize()
    assert h.b == pytest.approx(oracle.compat_params(n, ng, prec == "d")[0],
                                rel=1e-7 if prec == "f" else 1e-15)
    assert rel_l2(host(fh), oracle.compat_adjoint(x, y, ng, prec == "d")) < TOL[prec]
    L.free_filter_properties(d, None)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_compat_batched_rejects_bad_sizes(prec):
    x = dev(np.zeros(10, dtype=NP[prec]))
    fh = torch.zeros(64, dtype=torch.complex64 if prec == "f" else torch.complex128,
                     device="cuda")
    for n, ng in (([5, 5], [16, 1 << 27]), ([5, 5], [16, 2]), ([5, 40], [16, 16]), ([], [])):
        with pytest.raises(cu.CunaError):
            cu.cunfft_adjoint_raw_batched_async(x, None, fh, n, ng)


# ------------------------------------------------------- accurate mode --
@pytest.mark.parametrize("prec,m", [("f", 8), ("d", 13)])
def test_accurate_c1_vs_ndft(prec, m):
    n, M = 10000, 1 << 14
    x, y = uniform_points(n, 21, NP[prec])
    plan = cu.NFFTPlan(dev(x), M, m=m)
    got = host(plan.adjoint(dev(y)))[0]
    ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
    err = rel_l2(got, ref)
    assert err < TOL[prec], err
    # same algorithm in FP64 on the CPU
    gl = oracle.gl_adjoint(x, y, M, 2, m)
    assert rel_l2(got, gl) < (1e-5 if prec == "f" else 1e-12)


def test_accurate_c2_bands_vs_ndft():
    n, M = 1_000_000, 1 << 20
    x, y = clustered_points(n, 2024)
    plan = cu.NFFTPlan(dev(x), M)
    got = host(plan.adjoint(dev(y)))[0]
    rng = np.random.default_rng(7)
    k0 = np.sort(rng.integers(-M // 2, M // 2 - 64, 6))
    k0 = np.concatenate([k0, [-M // 2, -32, M // 2 - 64]])
    ref = oracle.ndft_bands(x.astype(np.float64), y.astype(np.float64), k0, 64)
    sel = np.stack([got[k + M // 2: k + M // 2 + 64] for k in k0])
    assert rel_l2(sel, ref) < 1e-5
    # property at full size: F_0 = sum y, Hermitian symmetry for real y
    assert abs(got[M // 2] - y.astype(np.float64).sum()) < 1e-5 * np.abs(got).max()
    assert rel_l2(got[1:M // 2][::-1], np.conj(got[M // 2 + 1:])) < 1e-6


@pytest.mark.parametrize("prec", ["f", "d"])
def test_accurate_edge_cases(prec):
    R = NP[prec]
    M = 64
    # one point at 0 with y = 1: F == 1
    plan = cu.NFFTPlan(dev(np.zeros(1, R)), M)
    got = host(plan.adjoint(dev(np.ones(1, R))))[0]
    np.testing.assert_allclose(got, np.ones(M), atol=TOL[prec] * 10)
    # all points in one cell (one hot tile), x = -1/2 and x just below 1/2
    x = np.array([0.123] * 500 + [-0.5, np.nextafter(R(0.5), R(0))] * 10, dtype=R)
    y = np.random.default_rng(0).standard_normal(len(x)).astype(R)
    for MM in (64, 1 << 12):
        plan = cu.NFFTPlan(dev(x), MM)
        got = host(plan.adjoint(dev(y)))[0]
        ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), MM)
        assert rel_l2(got, ref) < TOL[prec]
    # unit weights (y = NULL)
    got = host(plan.adjoint(None))[0]
    assert rel_l2(got, oracle.ndft(x.astype(np.float64), None, 1 << 12)) < TOL[prec]


def test_accurate_small_grids_and_tiles():
    for M, tl in ((2, 0), (4, 0), (16, 0), (256, 6), (1 << 12, 7), (1 << 15, 9)):
        x, y = uniform_points(777, M)
        plan = cu.NFFTPlan(dev(x), M, tile_log2=tl)
        got = host(plan.adjoint(dev(y)))[0]
        ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
        assert rel_l2(got, ref) < 1e-5, (M, tl)


def test_determinism_bitwise():
    x, y = clustered_points(200_000, 5)
    plan = cu.NFFTPlan(dev(x), 1 << 18)
    a = host(plan.adjoint(dev(y)))
    b = host(plan.adjoint(dev(y)))
    plan2 = cu.NFFTPlan(dev(x), 1 << 18)
    c = host(plan2.adjoint(dev(y)))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("pts,M,tile", [("clustered", 1 << 16, 0), ("dense", 1 << 12, 8),
                                         ("uniform", 1 << 14, 6), ("tiny", 8, 0)])
def test_scatter_and_gather_forms(pts, M, tile, prec):
    """The spread has two forms: fixed-point scatter (default) and gather.
    Both match the NDFT and each other; "dense" puts ~20k points in one tile
    (many scatter rounds), "tiny" is a grid smaller than two filter widths.
    FP64 (m = 13): both within 1e-11 of the NDFT and 1e-12 of each other."""
    if pts == "clustered":
        x, y = clustered_points(300_000, 12)
    elif pts == "dense":
        rng = np.random.default_rng(13)
        x = rng.uniform(0.1, 0.1 + 256 / (2 * M), 20_000).astype(np.float32)
        y = rng.standard_normal(20_000).astype(np.float32)
    else:
        x, y = uniform_points(5000 if pts == "tiny" else 100_000, 14)
    if prec == "d":
        x, y = x.astype(np.float64), y.astype(np.float64)
    tol_ndft, tol_pair = (1e-5, 2e-6) if prec == "f" else (1e-11, 1e-12)
    view = np.uint32 if prec == "f" else np.uint64
    ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
    # FP64: the same method in FP64 on the CPU pins the spread more tightly
    gl = oracle.gl_adjoint(x, y, M, 2, 13) if prec == "d" else None
    outs = []
    for gather in (False, True):
        plan = cu.NFFTPlan(dev(x), M, gather=gather, tile_log2=tile)
        a = host(plan.adjoint(dev(y)))[0]
        b = host(plan.adjoint(dev(y)))[0]
        assert np.array_equal(a.view(view), b.view(view))
        assert rel_l2(a, ref) < tol_ndft, (gather, rel_l2(a, ref))
        if gl is not None:
            assert rel_l2(a, gl) < 1e-12, (gather, rel_l2(a, gl))
        outs.append(a)
    assert rel_l2(outs[0], outs[1]) < tol_pair


@pytest.mark.parametrize("prec", ["f", "d"])
def test_nonfinite_value_propagates(prec):
    x, y = uniform_points(5000, 15)
    y = y.copy()
    y[1234] = np.nan
    if prec == "d":
        x, y = x.astype(np.float64), y.astype(np.float64)
    for gather in (False, True):
        out = host(cu.NFFTPlan(dev(x), 1 << 10, gather=gather).adjoint(dev(y)))[0]
        assert np.isnan(out).all()


def test_profile_events_ordered_and_harmless():
    """cuna_nfft_set_profile_events: the spread's start and stop (taken by its
    own dispatch), the mark between the FFT passes and the end come back in
    order with a positive spread time, and profiling does not change the
    result (bitwise)."""
    h = ctypes.CDLL("libamdhip64.so")
    h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                      ctypes.c_void_p]
    h.hipEventDestroy.argtypes = [ctypes.c_void_p]
    x, y = clustered_points(200_000, 41)
    M = 1 << 18  # two FFT passes
    plan = cu.NFFTPlan(dev(x), M)
    yd = dev(y)
    ref = host(plan.adjoint(yd))
    ev = []
    for _ in range(4):
        e = ctypes.c_void_p()
        assert h.hipEventCreate(ctypes.byref(e)) == 0
        ev.append(e.value)
    s = torch.cuda.Stream()
    arr = (ctypes.c_void_p * 4)(*ev)
    assert plan.L.cuna_nfft_set_profile_events(plan.h, arr, 4) == 0
    out = plan.adjoint(yd, stream=s)
    s.synchronize()
    assert plan.L.cuna_nfft_set_profile_events(plan.h, None, 0) == 0
    got = host(out)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    t = ctypes.c_float()
    ms = []
    for a, b in ((0, 1), (1, 2), (2, 3)):
        assert h.hipEventElapsedTime(ctypes.byref(t), ev[a], ev[b]) == 0
        ms.append(t.value)
    assert ms[0] > 0 and ms[1] >= 0 and ms[2] >= 0, ms
    for e in ev:
        h.hipEventDestroy(e)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_mostly_empty_tiles_repeat_bitwise(prec):
    """Points in 3% of the grid: most spread tiles are empty.  An empty tile
    must still write zeros, whatever earlier launches left in LDS: twenty
    transforms in a row are bitwise equal and match the NDFT."""
    rng = np.random.default_rng(21)
    dt = np.float32 if prec == "f" else np.float64
    x = rng.uniform(0.0, 0.03, 3000).astype(dt)
    y = rng.standard_normal(3000).astype(dt)
    M = 1 << 16
    plan = cu.NFFTPlan(dev(x), M)
    yd = dev(y)
    outs = [plan.adjoint(yd) for _ in range(20)]
    got = [host(o)[0] for o in outs]
    view = np.uint32 if prec == "f" else np.uint64
    for g in got[1:]:
        assert np.array_equal(g.view(view), got[0].view(view))
    ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
    assert rel_l2(got[0], ref) < (1e-5 if prec == "f" else 1e-11)


def test_batch_equals_single_curves():
    rng = np.random.default_rng(3)
    lens = [0, 1, 100, 5000, 20000, 333]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, offs[-1]).astype(np.float32)
    y = rng.standard_normal(offs[-1]).astype(np.float32)
    M = 1 << 13
    bp = cu.NFFTPlan(dev(x), M, offsets=offs)
    got = host(bp.adjoint(dev(y)))
    for c in range(len(lens)):
        xs, ys = x[offs[c]:offs[c + 1]], y[offs[c]:offs[c + 1]]
        if len(xs) == 0:
            assert np.all(got[c] == 0)
            continue
        single = host(cu.NFFTPlan(dev(xs), M).adjoint(dev(ys)))[0]
        assert rel_l2(got[c], single) < 1e-6
        ref = oracle.ndft(xs.astype(np.float64), ys.astype(np.float64), M)
        assert rel_l2(got[c], ref) < 1e-5


def test_times_plan_fp64_phase_reduction():
    # long baseline: t in [0, 36500) days, df = 1/(5*baseline)
    rng = np.random.default_rng(4)
    n = 20000
    t = np.sort(rng.uniform(0, 36500, n))
    y = rng.standard_normal(n).astype(np.float32)
    df = 1.0 / (5 * 36500.0)
    M = 1 << 16
    plan = cu.NFFTPlan(t=dev(t), df=df, M=M, dtype=torch.float32)
    got = host(plan.adjoint(dev(y)))[0]
    xs = (df * t) % 1.0
    ref = oracle.ndft(xs, y.astype(np.float64), M)
    assert rel_l2(got, ref) < 1e-5
    # an FP32 x loses the phase at this baseline (why the times path exists)
    x32 = ((df * t + 0.5) % 1.0 - 0.5).astype(np.float32)
    bad = host(cu.NFFTPlan(dev(x32), M).adjoint(dev(y)))[0]
    assert rel_l2(bad, ref) > rel_l2(got, ref)


# ------------------------------------------------------------ bootstrap --
def test_accurate_bootstrap_vs_oracle():
    n, M, nbs, seed = 2000, 1 << 11, 4, 777
    x, y = uniform_points(n, 8)
    plan = cu.NFFTPlan(dev(x), M, window2=True)
    oy, ow = plan.bootstrap(dev(y), nbs, seed, window=True)
    oy, ow = host(oy), host(ow)
    d = oracle.bootstrap_indices(seed, n, nbs)
    xd = x.astype(np.float64)
    for s in range(nbs):
        ref = oracle.ndft(xd[d[s]], y.astype(np.float64)[d[s]], M)
        assert rel_l2(oy[s], ref) < 1e-5
        refw = oracle.ndft(2 * xd[d[s]], None, M)
        assert rel_l2(ow[s], refw) < 1e-5


# ------------------------------------------------------------- backward --
@pytest.mark.parametrize("prec,m", [("f", 8), ("d", 13)])
def test_backward_vs_oracle(prec, m):
    n, M = 3000, 1 << 10
    x, y = uniform_points(n, 31, NP[prec])
    rng = np.random.default_rng(32)
    g = (rng.standard_normal(M) + 1j * rng.standard_normal(M))
    plan = cu.NFFTPlan(dev(x), M, m=m)
    gd = dev(g.astype(np.complex64 if prec == "f" else np.complex128)).reshape(1, M)
    dy, dx = plan.backward(dev(y), gd)
    dy, dx = host(dy), host(dx)
    rdy, rdx = oracle.gl_grad(x, y, M, g, 2, m)
    tol = 1e-5 if prec == "f" else 1e-11
    assert rel_l2(dy, rdy) < tol
    assert rel_l2(dx, rdx) < tol
    ady, adx = oracle.ndft_grad(x, y, M, g)
    tol = 1e-5 if prec == "f" else 1e-11
    assert rel_l2(dy, ady) < tol
    assert rel_l2(dx, adx) < tol


@pytest.mark.parametrize("prec,m", [("f", 8), ("d", 13)])
def test_forward_nfft_vs_oracle_and_adjointness(prec, m):
    n, M = 3000, 1 << 10
    x, y = uniform_points(n, 33, NP[prec])
    rng = np.random.default_rng(34)
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    plan = cu.NFFTPlan(dev(x), M, m=m)
    gd = dev(g.astype(np.complex64 if prec == "f" else np.complex128)).reshape(1, M)
    f = host(plan.forward(gd))
    ref = oracle.ndft_forward(x.astype(np.float64), g)
    tol = 1e-5 if prec == "f" else 1e-11
    assert rel_l2(f, ref) < tol
    # adjoint pair on the device: <forward g, y> = <g, adjoint y>
    F = host(plan.adjoint(dev(y)))[0]
    lhs = np.vdot(f.astype(np.complex128), y.astype(np.float64))
    rhs = np.vdot(g, F.astype(np.complex128))
    assert abs(lhs - rhs) < (1e-4 if prec == "f" else 1e-10) * abs(rhs)


def test_forward_nfft_batch_times():
    from tests._data import light_curves
    t, _, offs = light_curves(5, 8, nmin=20, nmax=1500)
    df, M = 1.0 / 3650.0, 1 << 9
    rng = np.random.default_rng(35)
    g = (rng.standard_normal((5, M)) + 1j * rng.standard_normal((5, M))).astype(np.complex64)
    plan = cu.NFFTPlan(t=dev(t), df=df, offsets=offs, M=M)
    f = host(plan.forward(dev(g)))
    xs = (df * t) % 1.0
    for c in range(5):
        sl = slice(offs[c], offs[c + 1])
        ref = oracle.ndft_forward(xs[sl], g[c].astype(np.complex128))
        assert rel_l2(f[sl], ref) < 1e-5


def test_autograd_batched_loss():
    B, n, M = 4, 2000, 1 << 12
    rng = np.random.default_rng(40)
    x = torch.tensor(rng.uniform(-0.5, 0.5, (B, n)), dtype=torch.float64, device="cuda",
                     requires_grad=True)
    y = torch.tensor(rng.standard_normal((B, n)), dtype=torch.float64, device="cuda",
                     requires_grad=True)
    w = torch.tensor(rng.uniform(0, 1, M), dtype=torch.float64, device="cuda")
    F = adjoint_nfft(x, y, M, m=13)
    L = (w * (F.real ** 2 + F.imag ** 2)).sum()
    L.backward()
    xh, yh = x.detach().cpu().numpy(), y.detach().cpu().numpy()
    for b in range(B):
        Fb = oracle.ndft(xh[b], yh[b], M)
        g = 2 * w.cpu().numpy() * Fb
        ady, adx = oracle.ndft_grad(xh[b], yh[b], M, g)
        assert rel_l2(y.grad[b].cpu().numpy(), ady) < 1e-10
        assert rel_l2(x.grad[b].cpu().numpy(), adx) < 1e-10


# ------------------------------------------- batched bootstrap + graphs --
def test_fused_batched_bootstrap_vs_unfused_and_oracle():
    from tests._data import light_curves
    t, y, offs = light_curves(24, 5, nmin=50, nmax=3000)
    df, M, nbs, seed = 1.0 / 3650.0, 1 << 12, 6, 31
    td, yd = dev(t), dev(y)
    fused = cu.NFFTPlan(t=td, df=df, offsets=offs, M=M, window2=True)
    unf = cu.NFFTPlan(t=td, df=df, offsets=offs, M=M, window2=True, unfused=True)
    fy, fw = fused.bootstrap(yd, nbs, seed, window=True)
    uy, uw = unf.bootstrap(yd, nbs, seed, window=True)
    fy, fw, uy, uw = host(fy), host(fw), host(uy), host(uw)
    assert rel_l2(fy, uy) < 1e-6 and rel_l2(fw, uw) < 1e-6
    xs = (df * t) % 1.0
    for c in (0, 7, 23):
        n = offs[c + 1] - offs[c]
        d = oracle.bootstrap_indices(seed, n, nbs) + offs[c]
        for s in (0, nbs - 1):
            ref = oracle.ndft(xs[d[s]], y.astype(np.float64)[d[s]], M)
            assert rel_l2(fy[c * nbs + s], ref) < 1e-5
            refw = oracle.ndft(2 * xs[d[s]], None, M)
            assert rel_l2(fw[c * nbs + s], refw) < 1e-5
    # per-curve plans give the same replicates as the batch
    c = 7
    sl = slice(offs[c], offs[c + 1])
    one = cu.NFFTPlan(t=dev(t[sl]), df=df, M=M, window2=True)
    oy = host(one.bootstrap(dev(y[sl]), nbs, seed))
    assert rel_l2(oy, fy[c * nbs:(c + 1) * nbs]) < 1e-6


@pytest.mark.parametrize("prec", ["f", "d"])
def test_bootstrap_lomb_scargle_vs_oracle(prec):
    """Fused Lomb-Scargle of bootstrap replicates vs the direct FP64
    periodogram of the same draws; the peak is the row's arg-max over k >= 1."""
    from tests._data import light_curves
    t, y, offs = light_curves(6, 21, nmin=50, nmax=2000)
    df, M, nbs, seed = 1.0 / 3650.0, 1 << 12, 3, 17
    yv = y.astype(NP[prec])
    plan = cu.NFFTPlan(t=dev(t), df=df, offsets=offs, M=M, window2=True,
                       dtype=torch.float32 if prec == "f" else torch.float64)
    r = plan.bootstrap_ls(dev(yv), nbs, seed)
    pw, pk, kp = host(r["power"]), host(r["peak"]), host(r["kpeak"])
    assert pw.shape == (6 * nbs, M // 2)
    tol = 2e-5 if prec == "f" else 1e-10
    worst = 0.0
    for c in range(6):
        n = offs[c + 1] - offs[c]
        d = oracle.bootstrap_indices(seed, n, nbs) + offs[c]
        for s in range(nbs):
            ref = oracle.lomb_scargle(t[d[s]], yv.astype(np.float64)[d[s]], df, M // 2)
            got = pw[c * nbs + s]
            worst = max(worst, float(np.max(np.abs(got - ref))))
            k = 1 + int(np.argmax(got[1:]))
            assert kp[c * nbs + s] == k and pk[c * nbs + s] == got[k]
    assert worst < tol, worst
    # peak-only call: same draws, same peaks
    r2 = plan.bootstrap_ls(dev(yv), nbs, seed, power=False)
    assert r2["power"] is None
    assert np.array_equal(host(r2["peak"]), pk) and np.array_equal(host(r2["kpeak"]), kp)


def test_bootstrap_lomb_scargle_edge_curves_and_graph_replay():
    """Empty, one-point and constant curves give zero power; replays of a
    captured call are bitwise equal to the eager call."""
    rng = np.random.default_rng(61)
    lens = [0, 1, 300, 50, 700]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    t = np.concatenate([np.sort(rng.uniform(0, 3650, n)) for n in lens])
    y = rng.standard_normal(offs[-1]).astype(np.float32)
    y[offs[3]:offs[4]] = 2.5  # constant curve
    M, nbs, seed = 1 << 10, 4, 5
    plan = cu.NFFTPlan(t=dev(t), df=1 / 3650.0, offsets=offs, M=M, window2=True)
    yd = dev(y)
    r = plan.bootstrap_ls(yd, nbs, seed)
    pw, pk = host(r["power"]), host(r["peak"])
    for c in (0, 1, 3):
        assert np.all(pw[c * nbs:(c + 1) * nbs] == 0) and np.all(pk[c * nbs:(c + 1) * nbs] == 0)
    assert np.all(np.isfinite(pw)) and np.all(pw >= 0) and np.all(pw <= 1 + 1e-4)
    d = oracle.bootstrap_indices(seed, 300, nbs) + offs[2]
    ref = oracle.lomb_scargle(t[d[1]], y.astype(np.float64)[d[1]], 1 / 3650.0, M // 2)
    assert np.max(np.abs(pw[2 * nbs + 1] - ref)) < 2e-5
    stream = torch.cuda.Stream()
    out = torch.empty_like(r["power"])
    pkd = torch.empty_like(r["peak"])
    kpd = torch.empty_like(r["kpeak"])
    run = lambda: plan.bootstrap_ls(yd, nbs, seed, stream=stream, out_power=out, out_peak=pkd,
                                    out_kpeak=kpd)
    run()
    stream.synchronize()
    eager = host(out).copy()
    g = cu.capture(run, stream)
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert np.array_equal(host(out).view(np.uint32), eager.view(np.uint32)), rep
        plan.bootstrap_ls(yd, nbs, seed + 1)
        torch.cuda.synchronize()
    g.close()


def test_forward_nfft_single_point_and_empty_curve():
    x = np.array([0.125], dtype=np.float32)
    M = 1 << 6
    g = np.zeros(M, np.complex64)
    g[M // 2 + 3] = 1.0  # k = 3
    f = host(cu.NFFTPlan(dev(x), M).forward(dev(g).reshape(1, M)))
    assert abs(f[0] - np.exp(-2j * np.pi * 3 * 0.125)) < 1e-5
    # batch with an empty curve
    offs = np.array([0, 0, 5], dtype=np.int32)
    x2 = np.linspace(-0.4, 0.4, 5).astype(np.float32)
    G2 = np.stack([np.zeros(M, np.complex64), g])
    f2 = host(cu.NFFTPlan(dev(x2), M, offsets=offs).forward(dev(G2)))
    ref = oracle.ndft_forward(x2.astype(np.float64), g.astype(np.complex128))
    assert rel_l2(f2, ref) < 1e-5


def test_bootstrap_lomb_scargle_rejects_unsupported():
    x, y = uniform_points(1000, 3)
    plan = cu.NFFTPlan(dev(x), 1 << 10)  # no window2
    with pytest.raises(cu.CunaError):
        plan.bootstrap_ls(dev(y), 2, 1)
    big = cu.NFFTPlan(dev(x), 1 << 16, window2=True)  # grid 2^17: no fused kernel
    with pytest.raises(cu.CunaError):
        big.bootstrap_ls(dev(y), 2, 1)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_fft_stage_matches_golden_fixtures(prec):
    """The FFT stage (CUFFT_INVERSE convention, unnormalised) against the
    committed numpy fixtures of tests/golden/."""
    import os
    gz = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_inverse_unnormalised.npz"))
    cdt = torch.complex64 if prec == "f" else torch.complex128
    for n in gz["sizes"]:
        d = torch.tensor(gz[f"x{n}"], dtype=cdt, device="cuda")
        got = host(cu.fft(d, +1)).astype(np.complex128)
        ref = gz[f"X{n}"]
        tol = (2e-6 if prec == "f" else 1e-13) * np.log2(max(int(n), 2))
        assert rel_l2(got, ref) < tol, int(n)


def test_graph_replays_match_eager():
    stream = torch.cuda.Stream()
    x, y = clustered_points(100_000, 9)
    M = 1 << 16
    plan = cu.NFFTPlan(dev(x), M)
    yd = dev(y)
    out = torch.empty((1, M), dtype=torch.complex64, device="cuda")
    with torch.cuda.stream(stream):
        plan.adjoint(yd, out, stream=stream)
    stream.synchronize()
    eager = host(out).copy()
    g = cu.capture(lambda: plan.adjoint(yd, out, stream=stream), stream)
    for rep in range(4):
        out.zero_()
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert np.array_equal(host(out).view(np.uint32), eager.view(np.uint32)), rep
        # eager library work between replays
        plan.adjoint(yd * 2, out)
        torch.cuda.synchronize()
    g.close()
    # the fused bootstrap (work-queue counter reset by a kernel node) replays too
    from tests._data import light_curves
    t, yy, offs = light_curves(16, 6, nmin=50, nmax=2000)
    bp = cu.NFFTPlan(t=dev(t), df=1 / 3650.0, offsets=offs, M=1 << 12, window2=True)
    ty = dev(yy)
    oy = torch.empty((16 * 4, 1 << 12), dtype=torch.complex64, device="cuda")
    ow = torch.empty_like(oy)
    bp.bootstrap(ty, 4, 5, window=True, out_y=oy, out_w=ow, stream=stream)
    stream.synchronize()
    ey, ew = host(oy).copy(), host(ow).copy()
    g = cu.capture(lambda: bp.bootstrap(ty, 4, 5, window=True, out_y=oy, out_w=ow,
                                        stream=stream), stream)
    for rep in range(3):
        oy.zero_()
        ow.zero_()
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        # fixed-point accumulation: replays are bit-identical to the eager call
        assert np.array_equal(host(oy).view(np.uint32), ey.view(np.uint32)), rep
        assert np.array_equal(host(ow).view(np.uint32), ew.view(np.uint32)), rep
        bp.bootstrap(ty, 4, 9, window=True)
        torch.cuda.synchronize()
    g.close()


# ------------------------------------------------------------ autotuner --
def test_autotune_meets_tolerance_and_caches(tmp_path):
    import time
    n, M = 50_000, 1 << 16
    x, y = uniform_points(n, 17)
    xd, yd = dev(x), dev(y)
    cache = str(tmp_path / "tune.txt")
    ref = oracle.ndft_bands(x.astype(np.float64), y.astype(np.float64),
                            np.array([-M // 2, -100, 5000], dtype=np.int64), 64)
    for tol in (1e-3, 1e-5):
        p = cu.autotune(xd, M, tol, cache_path=cache)
        assert p["sigma"] in (2, 4) and 2 <= p["m"] <= 16
        plan = cu.NFFTPlan(xd, M, **p)
        got = host(plan.adjoint(yd))[0]
        sel = np.stack([got[k + M // 2:k + M // 2 + 64] for k in (-M // 2, -100, 5000)])
        assert rel_l2(sel, ref) < tol
    t0 = time.perf_counter()
    p2 = cu.autotune(xd, M, 1e-5, cache_path=cache)
    assert time.perf_counter() - t0 < 0.5  # served from the cache file
    assert p2 == p
    assert len(open(cache).read().splitlines()) == 2


@pytest.mark.parametrize("prec", ["f", "d"])
def test_sigma4_and_explicit_fft_shapes(prec):
    n, M = 20_000, 1 << 15
    x, y = uniform_points(n, 41, NP[prec])
    ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
    for kw in ({"sigma": 4, "m": 6 if prec == "f" else 10}, {"fft_w": 1, "fft_h": 4},
               {"fft_w": 8, "fft_h": 1}, {"fft_w": 4, "fft_h": 2}):
        plan = cu.NFFTPlan(dev(x), M, **kw)
        got = host(plan.adjoint(dev(y)))[0]
        assert rel_l2(got, ref) < (1e-5 if prec == "f" else 1e-10), kw


def test_many_tiles_one_wave_per_sort_block():
    # 8192 tiles: the sort's per-wave key counters (3 x 8192 ints) allow one
    # wave per block; every chunk must still be binned
    n, M = 300_000, 1 << 20
    x, y = clustered_points(n, 77)
    for tl in (8, 9):
        plan = cu.NFFTPlan(dev(x), M, tile_log2=tl)
        got = host(plan.adjoint(dev(y)))[0]
        k0 = np.array([-M // 2, -20, 12345], dtype=np.int64)
        ref = oracle.ndft_bands(x.astype(np.float64), y.astype(np.float64), k0, 64)
        sel = np.stack([got[k + M // 2:k + M // 2 + 64] for k in k0])
        assert rel_l2(sel, ref) < 1e-5, tl
