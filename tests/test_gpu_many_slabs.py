"""Launches whose slab or batch count exceeds one grid's y dimension (65535):
the spread, the plan's cover kernel and the two-pass FFT split them into
several launches with a slab offset.  Each case checks slabs on both sides
of the 65535 boundary against the oracle (numpy's FFT for the FFT stage).
Tolerances: rel L2 1e-5 (FP32), 1e-13 (FP64 FFT)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from tests._data import rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    cu.load("f")
    cu.load("d")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_plan_with_more_curves_than_one_launch():
    """70,000 curves of 0-3 points (accurate mode, M = 64): the spread and
    the plan's cover kernel run as two launches each."""
    rng = np.random.default_rng(101)
    nc, M = 70_000, 64
    lens = rng.integers(1, 4, nc)
    lens[[7, 65_535]] = 0  # empty curves, one on each side of the split
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, offs[-1]).astype(np.float32)
    y = rng.standard_normal(offs[-1]).astype(np.float32)
    plan = cu.NFFTPlan(dev(x), M, offsets=offs)
    got = plan.adjoint(dev(y))
    torch.cuda.synchronize()
    assert got.shape == (nc, M)
    for c in (0, 7, 1, 65_534, 65_535, 65_536, nc - 1):
        row = got[c].cpu().numpy()
        sl = slice(offs[c], offs[c + 1])
        if offs[c + 1] == offs[c]:
            assert not np.any(row), c
            continue
        ref = oracle.ndft(x[sl].astype(np.float64), y[sl].astype(np.float64), M)
        assert rel_l2(row, ref) < 1e-5, c
    plan.close()


def test_compat_bootstrap_more_replicates_than_one_launch():
    """cunfft_adjoint_raw_async_bootstrap with nbs = 66,000 replicates: the
    spread runs as two launches over the replicates."""
    n, ng, nbs, seed = 40, 128, 66_000, 777
    rng = np.random.default_rng(102)
    x = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    xd, yd = dev(x), dev(y)
    fp = cu.generate_filter_properties(xd, n, ng)
    s = torch.cuda.Stream()
    fb = torch.zeros(nbs * ng, dtype=torch.complex64, device="cuda")
    cu.cunfft_adjoint_raw_async_bootstrap(xd, yd, fb, n, ng, nbs, fp, stream=s, seed=seed)
    s.synchronize()
    got = fb.reshape(nbs, ng)
    ref = oracle.compat_adjoint_bootstrap(x, y, ng, nbs, seed, False)
    for r in (0, 65_534, 65_535, 65_536, nbs - 1):
        assert rel_l2(got[r].cpu().numpy(), ref[r]) < 1e-5, r
    fp.free()


def test_two_pass_fft_batch_more_than_one_launch():
    """cuna_fft (libcunad) of 65,537 transforms of 2^14 points (two-pass in
    FP64): both passes run as two launches over the batch.  About 35 GB of
    device memory (the data and the library's work buffer)."""
    n, batch = 1 << 14, 65_537
    data = torch.zeros((batch, n), dtype=torch.complex128, device="cuda")
    rng = np.random.default_rng(103)
    rows = (0, 65_534, 65_535, 65_536)
    src = {}
    for r in rows:
        src[r] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        data[r] = dev(src[r])
    cu.fft(data, +1)
    torch.cuda.synchronize()
    for r in rows:
        ref = np.fft.ifft(src[r]) * n
        assert rel_l2(data[r].cpu().numpy(), ref) < 1e-13, r
    assert not torch.any(data[1]).item()
    del data
    torch.cuda.empty_cache()
