"""cuna_nfft_adjoint_many: several right-hand sides of one accurate plan in one
spread launch and one batched FFT.  Every row must equal the single-vector
transform of the same values: bitwise where the FFT runs as one workgroup per
transform (M <= 2^14 in FP32), within rounding (rel L2 1e-6 FP32, 1e-13 FP64)
where the two-pass FFT picks its workgroup shape from the batch size.  One
row is also checked against the direct NDFT.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from tests._data import clustered_points, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"f": np.float32, "d": np.float64}
VIEW = {"f": np.uint32, "d": np.uint64}


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


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("M,nrhs", [(1 << 12, 3), (1 << 16, 8), (1 << 16, 1)])
def test_many_rhs_equal_single_transforms(prec, M, nrhs):
    x, y0 = clustered_points(20_000, 91, NP[prec])
    rng = np.random.default_rng(92)
    y = np.stack([y0 * (r + 1) + rng.standard_normal(x.size) for r in range(nrhs)]).astype(NP[prec])
    plan = cu.NFFTPlan(dev(x), M)
    got = host(plan.adjoint_many(dev(y)))
    assert got.shape == (nrhs, M)
    small = M <= (1 << 14 if prec == "f" else 1 << 13)
    for r in range(nrhs):
        one = host(plan.adjoint(dev(y[r])))[0]
        if small:
            assert np.array_equal(got[r].view(VIEW[prec]), one.view(VIEW[prec])), r
        else:
            assert rel_l2(got[r], one) < (1e-6 if prec == "f" else 1e-13), r
    k0 = np.array([-M // 2, -7, M // 2 - 32])
    ref = oracle.ndft_bands(x.astype(np.float64), y[-1].astype(np.float64), k0, 32)
    sel = np.stack([got[-1][k + M // 2: k + M // 2 + 32] for k in k0])
    assert rel_l2(sel, ref) < (1e-5 if prec == "f" else 1e-11)
    plan.close()


def test_many_rhs_batch_plan_layout():
    """A plan over several curves: row c * nrhs + r is curve c, vector r."""
    rng = np.random.default_rng(93)
    lens = [500, 0, 3000, 1]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, offs[-1]).astype(np.float32)
    nrhs, M = 4, 1 << 13
    y = rng.standard_normal((nrhs, offs[-1])).astype(np.float32)
    plan = cu.NFFTPlan(dev(x), M, offsets=offs)
    got = host(plan.adjoint_many(dev(y)))
    assert got.shape == (len(lens) * nrhs, M)
    for r in range(nrhs):
        one = host(plan.adjoint(dev(y[r])))
        for c in range(len(lens)):
            assert np.array_equal(got[c * nrhs + r].view(np.uint32), one[c].view(np.uint32)), (c, r)
    plan.close()


def test_many_rhs_rejects():
    x = np.random.default_rng(94).uniform(-0.5, 0.5, 100).astype(np.float32)
    plan = cu.NFFTPlan(dev(x), 64)
    y = torch.zeros((0, 100), dtype=torch.float32, device="cuda")
    with pytest.raises(cu.CunaError):
        plan.adjoint_many(y, out=torch.zeros((1, 64), dtype=torch.complex64, device="cuda"))
    plan.close()
    cp = cu.NFFTPlan(dev(x), 256, mode="compat")
    with pytest.raises(cu.CunaError):
        cp.adjoint_many(dev(np.ones((2, 100), np.float32)))
    cp.close()
