"""GPU parity at the full sizes of the benchmark workloads, through the C-ABI.

Each workload of the benchmark runs here at its own size against the CPU
oracle (FP64 direct sums), on subsets where a full direct sum would take too
long, plus size-independent identities on the whole output:
  - N=1e6 clustered points, M=2^20, in the FP64 library (m = 13);
  - the batched bootstrap: 1024 uneven light curves (8e6 points), 100
    replicates x 2 transforms at M=2^14 in one fused launch, graph-replayed,
    and its Lomb-Scargle form;
  - N=2e5 long-baseline times with seasonal gaps, M=2^18, FP64 phase
    reduction with FP32 gridding, and all-FP64;
  - 64 curves x 2e4 points, M=2^16: backward (d/dy, d/dx);
  - forward (type-2) NFFT and backward at the sizes where the inverse FFT
    changes form (one workgroup up to N = 2^13, two passes from 2^14 on).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from tests._data import (clustered_points, light_curves, long_baseline_times,  # noqa: E402
                         rel_l2, uniform_points)

pytestmark = pytest.mark.gpu

NP = {"f": np.float32, "d": np.float64}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    cu.load("f")
    cu.load("d")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _bands(got, x, y, M, k0, length):
    ref = oracle.ndft_bands(x.astype(np.float64), None if y is None else y.astype(np.float64),
                            np.asarray(k0, dtype=np.int64), length)
    sel = np.stack([got[k + M // 2:k + M // 2 + length] for k in k0])
    return sel, ref


# ------------------------------------------- N = 1e6, M = 2^20, FP64 --
def test_c2_fp64_bands_vs_ndft():
    """FP64 library, accurate mode (sigma = 2, m = 13) on the clustered
    N = 1e6 points: rel L2 <= 1e-12 against the FP64 direct NDFT on nine
    bands of 64 modes (both ends of the spectrum included); F_0 = sum y and
    conjugate symmetry on the whole output."""
    n, M = 1_000_000, 1 << 20
    x, y = clustered_points(n, 2024, np.float64)
    plan = cu.NFFTPlan(dev(x), M)
    assert plan.info().m == 13
    got = host(plan.adjoint(dev(y)))[0]
    rng = np.random.default_rng(8)
    k0 = np.concatenate([np.sort(rng.integers(-M // 2, M // 2 - 64, 6)),
                         [-M // 2, -32, M // 2 - 64]])
    sel, ref = _bands(got, x, y, M, k0, 64)
    err = rel_l2(sel, ref)
    assert err < 1e-12, err
    assert abs(got[M // 2] - y.sum()) < 1e-12 * np.abs(got).max()
    assert rel_l2(got[1:M // 2][::-1], np.conj(got[M // 2 + 1:])) < 1e-13
    plan.close()


# ------------------------------------------------ batched bootstrap --
def test_bootstrap_1024_curves_m16384_graph():
    """1024 light curves of log-uniform length in [100, 5e4] (times in days,
    FP64 phase reduction, df = 1/3650), 100 replicates, M = 2^14, spectra of
    y at f and of unit weights at 2f from ONE fused launch captured in a
    graph: replays equal the eager call bitwise, and spot checks (longest,
    middle and shortest curve; first and last replicate) match the FP64
    direct NDFT of the drawn points within 1e-5.  The Lomb-Scargle form of
    the same replicates matches the FP64 periodogram within 2e-5."""
    ncurves, M, nbs, seed, df = 1024, 1 << 14, 100, 2024, 1.0 / 3650.0
    t, y, offs = light_curves(ncurves, 3650)
    lens = np.diff(offs)
    assert lens.min() >= 100 and lens.max() <= 50_000
    td, yd = dev(t), dev(y)
    plan = cu.NFFTPlan(t=td, df=df, offsets=offs, M=M, window2=True)
    stream = torch.cuda.Stream()
    out_y = torch.empty((ncurves * nbs, M), dtype=torch.complex64, device="cuda")
    out_w = torch.empty_like(out_y)
    run = lambda: plan.bootstrap(yd, nbs, seed, window=True, out_y=out_y, out_w=out_w,
                                 stream=stream)
    run()
    stream.synchronize()
    cs = [int(np.argmax(lens)), ncurves // 2, int(np.argmin(lens))]
    rows = [c * nbs + s for c in cs for s in (0, nbs - 1)]
    eager_y = host(out_y[rows]).copy()
    eager_w = host(out_w[rows]).copy()
    g = cu.capture(run, stream)
    for rep in range(2):
        out_y[rows] = 0
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert np.array_equal(host(out_y[rows]).view(np.uint32), eager_y.view(np.uint32)), rep
        assert np.array_equal(host(out_w[rows]).view(np.uint32), eager_w.view(np.uint32)), rep
    g.close()
    xs = (df * t) % 1.0
    yv = y.astype(np.float64)
    i = 0
    for c in cs:
        n = int(lens[c])
        d = oracle.bootstrap_indices(seed, n, nbs) + offs[c]
        for s in (0, nbs - 1):
            ref = oracle.ndft(xs[d[s]], yv[d[s]], M)
            assert rel_l2(eager_y[i], ref) < 1e-5, (c, s)
            refw = oracle.ndft(2 * xs[d[s]], None, M)
            assert rel_l2(eager_w[i], refw) < 1e-5, (c, s)
            i += 1
    del out_y, out_w
    torch.cuda.empty_cache()
    r = plan.bootstrap_ls(yd, nbs, seed, stream=stream)
    stream.synchronize()
    pw, pk, kp = r["power"], r["peak"], r["kpeak"]
    for c in cs:
        n = int(lens[c])
        d = oracle.bootstrap_indices(seed, n, nbs) + offs[c]
        for s in (0, nbs - 1):
            ref = oracle.lomb_scargle(t[d[s]], yv[d[s]], df, M // 2)
            got = host(pw[c * nbs + s])
            assert np.max(np.abs(got - ref)) < 2e-5, (c, s)
            k = 1 + int(np.argmax(got[1:]))
            assert int(host(kp[c * nbs + s])) == k
            assert float(host(pk[c * nbs + s])) == got[k]
    plan.close()


# -------------------------------------- long baseline, M = 2^18 --
@pytest.mark.parametrize("prec,m,tol", [("f", 8, 1e-5), ("d", 13, 1e-11), ("d", 16, 1e-12)])
def test_long_baseline_bands_vs_ndft(prec, m, tol):
    """N = 2e5 times over 100 years (nightly cadence, +-4 h jitter, 120-day
    seasonal gaps), df = 1/(5 baseline), M = 2^18: FP64 phase reduction with
    FP32 gridding (libcunaf) and all-FP64 (libcunad), against the FP64 direct
    NDFT per band of 64 modes across the spectrum.  The FP64 bound at m = 13
    is 1e-11: the Gaussian truncation (~1e-13) is amplified by the
    deconvolution, up to ~30x at the band edges k = +-M/2; m = 16 meets 1e-12."""
    n, M = 200_000, 1 << 18
    t, y = long_baseline_times(n, 36500)
    df = 1.0 / (5.0 * (t[-1] - t[0]))
    plan = cu.NFFTPlan(t=dev(t), df=df, M=M, m=m,
                       dtype=torch.float32 if prec == "f" else torch.float64)
    got = host(plan.adjoint(dev(y.astype(NP[prec]))))[0]
    xs = (df * t) % 1.0
    k0 = np.linspace(-M // 2, M // 2 - 64, 8).astype(np.int64)
    ref = oracle.ndft_bands(xs, y.astype(np.float64), k0, 64)
    for b, k in enumerate(k0):
        sl = slice(int(k) + M // 2, int(k) + M // 2 + 64)
        assert rel_l2(got[sl], ref[b]) < tol, (int(k), rel_l2(got[sl], ref[b]))
    plan.close()


# -------------------------------- differentiation, 64 x 2e4, M = 2^16 --
def test_backward_64_curves_m65536_vs_ndft_grad():
    """64 curves x N = 2e4, M = 2^16, L = sum_k w_k |F_k|^2 (FP32): the
    backward's d/dy and d/dx of two curves against the analytic gradient of
    the FP64 direct NDFT within 1e-5; the grid transform is the two-pass
    inverse (N = 2^16)."""
    B, n, M = 64, 20_000, 1 << 16
    rng = np.random.default_rng(5)
    x = rng.uniform(-0.5, 0.5, (B, n)).astype(np.float32)
    y = rng.standard_normal((B, n)).astype(np.float32)
    w = rng.uniform(0, 1, M)
    offs = np.arange(B + 1, dtype=np.int32) * n
    plan = cu.NFFTPlan(dev(x.reshape(-1)), M, offsets=offs)
    yd = dev(y.reshape(-1))
    F = plan.adjoint(yd)
    g = (2 * torch.from_numpy(w).to(F.device, torch.float32)) * F
    dy, dx = plan.backward(yd, g)
    dy, dx = host(dy).reshape(B, n), host(dx).reshape(B, n)
    Fh = host(F)
    for b in (0, B - 1):
        xb, yb = x[b].astype(np.float64), y[b].astype(np.float64)
        Fb = oracle.ndft(xb, yb, M)
        assert rel_l2(Fh[b], Fb) < 1e-5
        ady, adx = oracle.ndft_grad(xb, yb, M, 2 * w * Fb)
        assert rel_l2(dy[b], ady) < 1e-5, b
        assert rel_l2(dx[b], adx) < 1e-5, b
    plan.close()


# ------------------------- forward / backward across the inverse FFT forms --
@pytest.mark.parametrize("prec,lgM", [("f", 13), ("f", 14), ("f", 16), ("d", 12), ("d", 13),
                                      ("d", 16)])
def test_forward_and_backward_sizes(prec, lgM):
    """Forward NFFT against the FP64 forward NDFT and backward against the
    oracle's exact gradient, at M where the inverse transform of the grid
    (N = M with sigma = 2) is one workgroup (N <= 2^13) or two passes
    (N >= 2^14), in both libraries."""
    M = 1 << lgM
    m = 8 if prec == "f" else 13
    n = 4000
    x, y = uniform_points(n, 50 + lgM, NP[prec])
    rng = np.random.default_rng(lgM)
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    gd = dev(g.astype(np.complex64 if prec == "f" else np.complex128)).reshape(1, M)
    plan = cu.NFFTPlan(dev(x), M, m=m)
    f = host(plan.forward(gd))
    tol = 1e-5 if prec == "f" else 1e-11
    ref = oracle.ndft_forward(x.astype(np.float64), g)
    assert rel_l2(f, ref) < tol, rel_l2(f, ref)
    dy, dx = plan.backward(dev(y), gd)
    ady, adx = oracle.ndft_grad(x.astype(np.float64), y.astype(np.float64), M, g)
    assert rel_l2(host(dy), ady) < tol
    assert rel_l2(host(dx), adx) < tol
    plan.close()
