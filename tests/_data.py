"""Seeded synthetic inputs of the workloads (shared by tests and bench.py)."""
import numpy as np


def uniform_points(n, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, n).astype(dtype)
    y = rng.standard_normal(n).astype(dtype)
    return x, y


def clustered_points(n, seed, dtype=np.float32, nclump=50, width=0.005, frac=0.8):
    """80 % of the points in 50 Gaussian clumps of width 0.005, 20 % uniform,
    wrapped into [-1/2, 1/2); y = 3 sinusoids (10..1e5 cycles, amplitude
    0.1..1) + N(0, 0.3)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-0.5, 0.5, nclump)
    nc = int(round(frac * n))
    which = rng.integers(0, nclump, nc)
    xc = centres[which] + width * rng.standard_normal(nc)
    xu = rng.uniform(-0.5, 0.5, n - nc)
    x = np.concatenate([xc, xu])
    x = x[rng.permutation(n)]
    x = (x + 0.5) % 1.0 - 0.5
    f = rng.uniform(10, 1e5, 3)
    a = rng.uniform(0.1, 1.0, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    y = sum(a[i] * np.sin(2 * np.pi * f[i] * x + ph[i]) for i in range(3))
    y = y + 0.3 * rng.standard_normal(n)
    return x.astype(dtype), y.astype(dtype)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def light_curves(ncurves, seed, nmin=100, nmax=50000, baseline=3650.0):
    """Uneven light curves: lengths log-uniform in [nmin, nmax], times sorted
    Uniform[0, baseline) days (float64), y = sinusoid of period log-uniform in
    [0.1, 100] d + Normal(0, 1) (float32).  Returns t, y, CSR offsets."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(nmin), np.log(nmax), ncurves)).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
    t = np.empty(offs[-1], np.float64)
    y = np.empty(offs[-1], np.float32)
    for c in range(ncurves):
        tc = np.sort(rng.uniform(0.0, baseline, n[c]))
        P = np.exp(rng.uniform(np.log(0.1), np.log(100.0)))
        ph = rng.uniform(0, 2 * np.pi)
        t[offs[c]:offs[c + 1]] = tc
        y[offs[c]:offs[c + 1]] = np.sin(2 * np.pi * tc / P + ph) + rng.standard_normal(n[c])
    return t, y, offs


def long_baseline_times(n, seed, years=100, gap_days=120, jitter_days=1.0 / 3.0):
    """Times over `years` years (days): nightly cadence, several visits per
    night with +-4 h jitter (uniform over `jitter_days`), no visits during a
    `gap_days` seasonal gap each year; y ~ Normal(0, 1) float32."""
    rng = np.random.default_rng(seed)
    nights = np.arange(int(365 * years))
    nights = nights[(nights % 365) >= gap_days]
    night = rng.choice(nights, n)
    t = np.sort(night + rng.uniform(-jitter_days / 2, jitter_days / 2, n))
    t -= t[0]
    y = rng.standard_normal(n).astype(np.float32)
    return t, y
