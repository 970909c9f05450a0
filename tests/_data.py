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
