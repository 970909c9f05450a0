"""Pins the CPU oracle (CPU only).

The reference ships no golden vectors and cannot run here (CUDA + cuFFT), so
each oracle function is pinned by something independent of its own code:
known answers of the bootstrap generator, brute-force sums, closed forms of
the NDFT, finite differences, and -- for the compat formula -- a separate
float32 step-by-step emulation of the reference's arithmetic (E1 * E2^m * E3
products, src/gaussian_filter.cu:204-273) written with numpy, against the
oracle's double-precision closed form exp(-(eps-m)^2/b)/sqrt(pi b).
"""
import os

import numpy as np
import pytest

import oracle
from tests._data import rel_l2, uniform_points


# ----------------------------------------------------------------- RNG --
def test_rng_known_answers():
    # known answers of wang_hash / xorshift (src/gridding.cu:67-82)
    assert oracle.wang_hash(0) == 3232319850
    assert oracle.wang_hash(1) == 663891101
    d = oracle.bootstrap_indices(0, 1000, 1)[0, :8]
    assert list(d) == [136, 933, 190, 253, 477, 128, 637, 479]
    d = oracle.bootstrap_indices(12345, 10000, 1)[0, :8]
    assert list(d) == [8717, 5010, 4898, 436, 7010, 4961, 9979, 3531]


def _np_wang(s):
    s = s.astype(np.uint64) & 0xFFFFFFFF
    s = (s ^ 61) ^ (s >> 16)
    s = (s * 9) & 0xFFFFFFFF
    s = s ^ (s >> 4)
    s = (s * 0x27D4EB2D) & 0xFFFFFFFF
    s = s ^ (s >> 15)
    return s


def _np_xorshift(s):
    s = s ^ ((s << 13) & 0xFFFFFFFF)
    s = s ^ (s >> 17)
    s = s ^ ((s << 5) & 0xFFFFFFFF)
    return s


def test_rng_matches_vectorised_restatement():
    seed, n, nbs = 4242, 777, 5
    t = np.arange(n * nbs, dtype=np.uint64)
    d = (_np_xorshift(_np_wang((seed + t) & 0xFFFFFFFF)) % n).astype(np.int32)
    assert np.array_equal(oracle.bootstrap_indices(seed, n, nbs).ravel(), d)


# ----------------------------------------------------------------- FFT --
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 16, 31, 64, 100])
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_brute_force(n, sign):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    g = np.arange(n)
    ref = np.exp(sign * 2j * np.pi * np.outer(g, g) / n) @ z
    assert np.abs(oracle.fft(z, sign) - ref).max() < 1e-12 * max(1, n)


@pytest.mark.parametrize("n", [1 << 10, 1 << 15, 3000])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(1)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(oracle.fft(z, +1), np.fft.ifft(z) * n) < 1e-13
    assert rel_l2(oracle.fft(z, -1), np.fft.fft(z)) < 1e-13


# ---------------------------------------------------------------- NDFT --
def test_ndft_brute_force_and_closed_forms():
    rng = np.random.default_rng(2)
    n, M = 300, 256
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.standard_normal(n)
    k = np.arange(-M // 2, M // 2)
    ref = np.exp(2j * np.pi * np.outer(k, x)) @ y
    F = oracle.ndft(x, y, M)
    assert rel_l2(F, ref) < 1e-13
    # one point at 0 with unit weight: F == 1
    assert np.allclose(oracle.ndft(np.zeros(1), np.ones(1), 64), 1.0, atol=1e-15)
    # F_0 = sum y; conjugate symmetry for real y
    assert abs(F[M // 2] - y.sum()) < 1e-12
    assert np.allclose(F[1:M // 2][::-1], np.conj(F[M // 2 + 1:]), atol=1e-11)
    # shift x -> x + s/G multiplies F_k by exp(2 pi i k s/G)
    G, s = 512, 3
    Fs = oracle.ndft(x + s / G, y, M)
    assert rel_l2(Fs, F * np.exp(2j * np.pi * k * s / G)) < 1e-12
    # linearity and unit weights (NULL y)
    y2 = rng.standard_normal(n)
    assert rel_l2(oracle.ndft(x, 2 * y - 3 * y2, M),
                  2 * F - 3 * oracle.ndft(x, y2, M)) < 1e-13
    assert rel_l2(oracle.ndft(x, None, M), oracle.ndft(x, np.ones(n), M)) < 1e-15


def test_ndft_large_k_phase_reduction():
    # |k| up to 2^21: the phase must be reduced exactly (fma split)
    x = np.array([0.123456789012345678, -0.3141592653589793])
    y = np.array([1.0, 0.5])
    k0 = np.array([(1 << 21) - 8], dtype=np.int64)
    F = oracle.ndft_bands(x, y, k0, 8)[0]
    from fractions import Fraction
    ref = []
    for k in range(int(k0[0]), int(k0[0]) + 8):
        acc = 0
        for xi, yi in zip(x, y):
            ph = Fraction(k) * Fraction(xi)
            ph -= int(ph)
            acc += yi * np.exp(2j * np.pi * float(ph))
        ref.append(acc)
    assert np.abs(F - np.array(ref)).max() < 1e-13


# ----------------------------------------------------- Greengard-Lee --
@pytest.mark.parametrize("m,tol", [(6, 2e-6), (8, 3e-8), (13, 1e-12)])
def test_gl_vs_ndft(m, tol):
    x, y = uniform_points(3000, 5, np.float64)
    M = 1024
    err = rel_l2(oracle.gl_adjoint(x, y, M, 2, m), oracle.ndft(x, y, M))
    assert err < tol
    # the error falls with m (a dropped term or wrong width would not)
    assert err > 1e-16


def test_gl_keeps_every_bit_of_negative_fp64_x():
    """At M = 2^20 the phase of the top modes needs every bit of x: shifting a
    negative x by 1 (or G x by G) before the grid reduction drops its last
    bit, an error of ~1e-10 at |k| ~ 2^19.  m = 16 keeps the method's own
    error below 1e-13, so 1e-12 separates the two."""
    rng = np.random.default_rng(11)
    n, M = 4000, 1 << 20
    x = rng.uniform(-0.5, 0.5, n) * (2.0 / 3.0)  # every mantissa bit in use
    y = rng.standard_normal(n)
    neg = x[x < 0]
    assert np.mean((neg + 1.0) - 1.0 != neg) > 0.3  # x + 1 would round these
    got = oracle.gl_adjoint(x, y, M, 2, 16)
    k0 = np.array([-M // 2, M // 2 - 64, 1000], dtype=np.int64)
    ref = oracle.ndft_bands(x, y, k0, 64)
    sel = np.stack([got[k + M // 2:k + M // 2 + 64] for k in k0])
    assert rel_l2(sel, ref) < 1e-12


def test_gl_sigma4_and_unit_weights():
    x, y = uniform_points(1000, 6, np.float64)
    M = 256
    assert rel_l2(oracle.gl_adjoint(x, y, M, 4, 6), oracle.ndft(x, y, M)) < 1e-7
    assert rel_l2(oracle.gl_adjoint(x, None, M, 2, 8), oracle.ndft(x, None, M)) < 1e-7


# ---------------------------------------------------------- gradients --
def _loss(x, y, w, M):
    F = oracle.ndft(x, y, M)
    return float(np.sum(w * np.abs(F) ** 2)), F


def test_ndft_grad_vs_finite_differences():
    rng = np.random.default_rng(7)
    n, M, h = 12, 32, 1e-6
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.standard_normal(n)
    w = rng.uniform(0, 1, M)
    L0, F = _loss(x, y, w, M)
    g = 2 * w * F
    dy, dx = oracle.ndft_grad(x, y, M, g)
    fdy = np.array([(_loss(x, y + h * e, w, M)[0] - _loss(x, y - h * e, w, M)[0]) / (2 * h)
                    for e in np.eye(n)])
    fdx = np.array([(_loss(x + h * e, y, w, M)[0] - _loss(x - h * e, y, w, M)[0]) / (2 * h)
                    for e in np.eye(n)])
    assert rel_l2(dy, fdy) < 1e-7
    assert rel_l2(dx, fdx) < 1e-7


def test_gl_grad_vs_ndft_grad():
    x, y = uniform_points(500, 9, np.float64)
    M = 256
    rng = np.random.default_rng(10)
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    a = oracle.ndft_grad(x, y, M, g)
    b = oracle.gl_grad(x, y, M, g, 2, 13)
    assert rel_l2(b[0], a[0]) < 1e-11
    assert rel_l2(b[1], a[1]) < 1e-11


# ------------------------------------------------------- compat formula --
def _compat_float32_steps(x, y, ng, bootstrap=None):
    """The reference's single-precision arithmetic, step by step in numpy
    float32: src/gaussian_filter.cu:54-57 (b), :213-226 (u, eps, E1, E2, E3),
    :264-273 (filter = E1 * E2^m * E3[|m|]), src/gridding.cu:43-63 (j, taps),
    src/gaussian_filter.cu:284-287 (normalize)."""
    f32 = np.float32
    x = x.astype(f32)
    n = len(x)
    R = f32(ng) / f32(n)
    b = f32(float((f32(2) * R) / (f32(2) * R - f32(1))) * (6 / np.pi))
    prod = (f32(ng) * x).astype(f32)
    u = np.trunc((prod - f32(6)).astype(f32)).astype(np.int64)
    eps = (prod - u.astype(f32)).astype(f32)
    eps2 = (eps / b).astype(f32)
    E1 = (np.exp(-eps * eps2).astype(f32) / f32(np.sqrt(f32(np.pi * float(b))))).astype(f32)
    E2 = np.exp(f32(2) * eps2).astype(f32)
    E3 = np.exp((-(np.arange(6) ** 2)).astype(f32) / b).astype(f32)
    j = np.trunc((x.astype(np.float64) + 0.5) * (ng - 1)).astype(np.int64)
    nslab = 1 if bootstrap is None else bootstrap[1]
    grid = np.zeros((nslab, ng), np.float64)
    for s in range(nslab):
        if bootstrap is None:
            jj = np.clip(j, 0, ng - 1)
            d = np.arange(n)
        else:
            jj = j
            seed = bootstrap[0]
            t = np.arange(n, dtype=np.uint64) + s * n
            d = (_np_xorshift(_np_wang((seed + t) & 0xFFFFFFFF)) % n).astype(np.int64)
            if y is None:
                d = np.arange(n)
        val = np.ones(n, f32) if y is None else y.astype(f32)[d]
        for m in range(-5, 6):
            w = (val * ((E1[d] * np.power(E2[d], f32(m))).astype(f32) * E3[abs(m)])).astype(f32)
            ok = (jj + m >= 0) & (jj + m < ng)
            np.add.at(grid[s], (jj + m)[ok], w[ok].astype(np.float64))
    F = np.fft.ifft(grid, axis=-1) * ng
    K = (np.arange(ng, dtype=f32) / f32(ng)).astype(f32)
    fac = f32(np.sqrt(2 * np.pi)) * np.exp((K * K * b).astype(np.float64) * 0.25).astype(f32)
    F = F * fac
    return F[0] if bootstrap is None else F


@pytest.mark.parametrize("n,ng", [(10000, 1 << 15), (1000, 1 << 10), (300, 256)])
def test_compat_oracle_vs_float32_steps(n, ng):
    x, y = uniform_points(n, n, np.float32)
    ref = _compat_float32_steps(x, y, ng)
    got = oracle.compat_adjoint(x, y, ng, False)
    assert rel_l2(got, ref) < 1e-5
    got_w = oracle.compat_adjoint(x, None, ng, False)
    assert rel_l2(got_w, _compat_float32_steps(x, None, ng)) < 1e-5


def test_compat_bootstrap_oracle_vs_float32_steps():
    n, ng, nbs, seed = 800, 1 << 11, 3, 99
    x, y = uniform_points(n, 1, np.float32)
    ref = _compat_float32_steps(x, y, ng, bootstrap=(seed, nbs))
    got = oracle.compat_adjoint_bootstrap(x, y, ng, nbs, seed, False)
    assert rel_l2(got, ref) < 1e-5
    ref = _compat_float32_steps(x, None, ng, bootstrap=(seed, nbs))
    got = oracle.compat_adjoint_bootstrap(x, None, ng, nbs, seed, False)
    assert rel_l2(got, ref) < 1e-5


def test_compat_structure():
    ng = 1 << 12
    # one point: support of exactly 2 R_f - 1 cells around j = (x+0.5)(ng-1)
    for dbl, R in ((False, 6), (True, 12)):
        for x0 in (0.0, 0.1234, -0.4999, 0.49):
            g = oracle.compat_grid(np.array([x0]), np.array([1.0]), ng, dbl)
            j = int((x0 + 0.5) * (ng - 1))
            nz = np.nonzero(g)[0]
            lo, hi = max(0, j - R + 1), min(ng - 1, j + R - 1)
            assert nz.min() >= lo and nz.max() <= hi
            # the one-sided window: the largest tap sits at the top end
            # (eps in [R_f - 1, R_f + 1): u truncates ng*x - R_f toward zero,
            # src/gaussian_filter.cu:213-214)
            u, eps, jj = oracle.compat_point(x0, 2, ng, dbl)
            assert jj == j and R - 1 <= eps < R + 1
    # linearity in y, and the reference formula's distance from the NDFT:
    # rel L2 between 1.2 and 1.6 at n=1e4, ng=2^15 (its window is one-sided)
    x, y = uniform_points(10000, 11, np.float32)
    y2 = np.random.default_rng(3).standard_normal(10000).astype(np.float32)
    a = oracle.compat_adjoint(x, y, 1 << 15, False)
    b2 = oracle.compat_adjoint(x, y2, 1 << 15, False)
    c = oracle.compat_adjoint(x, (y + y2).astype(np.float64), 1 << 15, False)
    assert rel_l2(c, a + b2) < 1e-6
    F = oracle.ndft(x.astype(np.float64), y.astype(np.float64), 1 << 14)
    bias = rel_l2(a[: 1 << 13], F[1 << 13:])
    assert 1.2 < bias < 1.6


def test_compat_params_closed_form():
    for n, ng in ((10000, 1 << 15), (1000, 1 << 10), (7, 16)):
        b, nf, r = oracle.compat_params(n, ng, True)
        R = ng / n
        assert r == 12 and b == pytest.approx(2 * R / (2 * R - 1) * 12 / np.pi, rel=1e-15)
        assert nf == pytest.approx(np.sqrt(2 * np.pi), rel=1e-15)
        bf, nff, rf = oracle.compat_params(n, ng, False)
        assert rf == 6 and bf == pytest.approx(2 * R / (2 * R - 1) * 6 / np.pi, rel=1e-6)
        assert np.float32(bf) == bf


# ------------------------------------------------------- Lomb-Scargle --
def test_lomb_scargle_equals_least_squares_sinusoid_fit():
    """The classical periodogram of centred data is the fraction of variance
    explained by the least-squares fit of a cos wt + b sin wt (the tau shift
    only rotates the basis); numpy lstsq gives that independently."""
    rng = np.random.default_rng(70)
    t = np.sort(rng.uniform(0, 100.0, 300))
    y = np.sin(2 * np.pi * 0.137 * t + 0.4) + 0.5 * rng.standard_normal(300)
    df, nf = 0.01, 60
    P = oracle.lomb_scargle(t, y, df, nf)
    yc = y - y.mean()
    for k in range(1, nf):
        w = 2 * np.pi * k * df
        X = np.stack([np.cos(w * t), np.sin(w * t)], axis=1)
        coef, *_ = np.linalg.lstsq(X, yc, rcond=None)
        rss = np.sum((yc - X @ coef) ** 2)
        assert abs(P[k] - (1 - rss / np.sum(yc ** 2))) < 1e-10, k
    assert np.all(P >= 0) and np.all(P <= 1 + 1e-12)


def test_lomb_scargle_peak_and_edge_cases():
    t = np.linspace(0, 50, 400, endpoint=False)
    y = np.cos(2 * np.pi * 0.5 * t)
    P = oracle.lomb_scargle(t, y, 0.02, 50)
    assert int(np.argmax(P)) == 25 and abs(P[25] - 1) < 1e-9
    # constant data: no variance, power 0
    assert np.all(oracle.lomb_scargle(t, np.full(400, 3.0), 0.02, 10) == 0)


# ------------------------------------------------------- forward NDFT --
def test_ndft_forward_brute_force_and_adjointness():
    rng = np.random.default_rng(80)
    x = rng.uniform(-0.5, 0.5, 200)
    M = 64
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    k = np.arange(-M // 2, M // 2)
    brute = np.exp(-2j * np.pi * np.outer(x, k)) @ g
    f = oracle.ndft_forward(x, g)
    assert np.max(np.abs(f - brute)) < 1e-11
    # <F2 g, y> = <g, F y> with F the (independently pinned) adjoint NDFT
    y = rng.standard_normal(200)
    lhs = np.vdot(f, y)                      # sum conj(f_j) y_j
    rhs = np.vdot(g, oracle.ndft(x, y, M))   # sum conj(g_k) F_k
    assert abs(lhs - rhs) < 1e-9 * abs(rhs)


# ------------------------------------------------- FFT golden fixtures --
def _fft_golden():
    path = os.path.join(os.path.dirname(__file__), "golden", "fft_inverse_unnormalised.npz")
    return np.load(path)


def test_fft_matches_golden_fixtures():
    gz = _fft_golden()
    for n in gz["sizes"]:
        got = oracle.fft(gz[f"x{n}"], +1)
        ref = gz[f"X{n}"]
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)) * np.log2(max(n, 2)), n
