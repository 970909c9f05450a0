"""The compat entry points at grid lengths that are not powers of two, through
the C-ABI, against the oracle's restatement of the reference formula.

The reference hands any ng to cuFFT (src/cuna_kernel.cu:82,169,208,261) and
its own test program builds ng = R*N (test.cpp:114), e.g. 2 * 1000 = 2000.  Here
such lengths take the Bluestein transform (csrc/cuna_chirp.hip) over the
power-of-two Stockham FFT.  Tolerances as for the power-of-two grids: rel L2
1e-5 (libcunaf) / 1e-12 (libcunad).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from tests._data import rel_l2, uniform_points  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = {"f": 1e-5, "d": 1e-12}
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


def cdt(prec):
    return torch.complex64 if prec == "f" else torch.complex128


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("ng", [2000, 3000, 6144, 10000, 100000])
def test_compat_entries_any_ng(prec, ng):
    """cunfft_adjoint (host arrays), cunfft_adjoint_raw_async and
    cunfft_adjoint_raw_async_bootstrap at ng = 2 n (the reference test program's
    R = 2), against the oracle.  ng = 100000: 391 spread tiles, a partial last
    tile, and two sort passes over the tile index."""
    n = ng // 2
    x, y = uniform_points(n, ng, NP[prec])
    dbl = prec == "d"
    ref = oracle.compat_adjoint(x, y, ng, dbl)
    got = cu.cunfft_adjoint(x, y, ng)
    assert rel_l2(got, ref) < TOL[prec]
    xd, yd = dev(x), dev(y)
    fp = cu.generate_filter_properties(xd, n, ng)
    s = torch.cuda.Stream()
    fh = torch.zeros(ng, dtype=cdt(prec), device="cuda")
    cu.cunfft_adjoint_raw_async(xd, yd, fh, n, ng, fp, stream=s)
    s.synchronize()
    assert rel_l2(host(fh), ref) < TOL[prec]
    nbs, seed = 3, 99
    fb = torch.zeros(nbs * ng, dtype=cdt(prec), device="cuda")
    cu.cunfft_adjoint_raw_async_bootstrap(xd, yd, fb, n, ng, nbs, fp, stream=s, seed=seed)
    s.synchronize()
    refb = oracle.compat_adjoint_bootstrap(x, y, ng, nbs, seed, dbl)
    assert rel_l2(host(fb).reshape(nbs, ng), refb) < TOL[prec]
    fp.free()


@pytest.mark.parametrize("prec", ["f", "d"])
def test_compat_small_and_odd_ng(prec):
    """Lengths 1, 2, 3 and odd, prime and near-power-of-two lengths, with the
    plan path's window transform."""
    for ng in (1, 2, 3, 5, 7, 13, 97, 1023, 1025, 4099):
        n = max(1, (2 * ng - 1) // 3)
        x, y = uniform_points(n, 7 * ng + 1, NP[prec])
        ref = oracle.compat_adjoint(x, y, ng, prec == "d")
        plan = cu.CunaPlan(x, y, ng, cu._lib.CALCULATE_WINDOW_FUNCTION)
        plan.execute()
        assert rel_l2(plan.f_hat(), ref) < TOL[prec], ng
        refw = oracle.compat_adjoint(x, None, ng, prec == "d")
        assert rel_l2(plan.f_hat_win(), refw) < TOL[prec], ng
        plan.free()


@pytest.mark.parametrize("prec", ["f", "d"])
def test_compat_batched_mixed_lengths(prec):
    """The batched compat API with lengths of both kinds in one call: each
    slab equals the oracle and the rest of each slab is untouched."""
    rng = np.random.default_rng(71)
    n = [900, 30, 2500, 700, 1]
    ng = [2000, 64, 3000, 2000, 3]
    nlc, idist = len(n), max(ng)
    off = np.concatenate([[0], np.cumsum(n)])
    x = rng.uniform(-0.5, 0.5, off[-1]).astype(NP[prec])
    y = rng.standard_normal(off[-1]).astype(NP[prec])
    xd, yd = dev(x), dev(y)
    sentinel = 3.0 - 1.0j
    fh = torch.full((nlc * idist,), sentinel, dtype=cdt(prec), device="cuda")
    cu.cunfft_adjoint_raw_batched_async(xd, yd, fh, n, ng)
    got = host(fh).reshape(nlc, idist)
    for i in range(nlc):
        sl = slice(off[i], off[i + 1])
        ref = oracle.compat_adjoint(x[sl], y[sl], ng[i], prec == "d")
        assert rel_l2(got[i, :ng[i]], ref) < TOL[prec], i
        assert np.all(got[i, ng[i]:] == sentinel), i


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_stage_any_length(prec, sign):
    """cuna_fft at lengths that are not powers of two (both signs, batches)
    against numpy; n = 1 is the identity."""
    for n in (1, 3, 5, 12, 100, 1000, 4097, 10007, 100003):
        batch = 2
        rng = np.random.default_rng(n)
        z = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
        t = dev(z.astype(np.complex64 if prec == "f" else np.complex128))
        cu.fft(t, sign=sign)
        got = host(t)
        ref = np.fft.ifft(z, axis=-1) * n if sign > 0 else np.fft.fft(z, axis=-1)
        tol = (2e-6 if prec == "f" else 1e-13) * max(1.0, np.log2(n))
        assert rel_l2(got, ref) < tol, (n, rel_l2(got, ref))
