"""The spread's fixed-point scale is bounded by each tile's cover (the most
points whose taps reach one cell, counted at plan time over the tile's own
bucket and its neighbours' edge classes).  A cover that undercounted would
let a dense cell wrap its integer sum.  These cases put hundreds of
same-sign values into one window of taps: across a tile boundary, across
the periodic wrap of the grid, and in a grid that is a single periodic
tile; each is compared with the same method in FP64 on the CPU
(oracle.gl_adjoint) and with the direct NDFT.  Tolerances: rel L2 1e-5
(libcunaf), 1e-11 (libcunad, m = 13).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import cunfft_adjoint_amd as cu  # noqa: E402
from tests._data import rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"f": np.float32, "d": np.float64}
TOL = {"f": 1e-5, "d": 1e-11}
M_OF = {"f": 8, "d": 13}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    cu.load("f")
    cu.load("d")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(x, y, M, prec):
    plan = cu.NFFTPlan(dev(x), M)
    torch.cuda.synchronize()
    got = plan.adjoint(dev(y))[0].cpu().numpy()
    info = plan.info()
    plan.close()
    return got, info


def dense_window(centre_cell, G, n, rng, dt):
    """n points within +-3 cells of a grid cell, y = 1 (all the same sign)."""
    pos = centre_cell + rng.uniform(-3.0, 3.0, n)
    x = ((pos / G + 0.5) % 1.0) - 0.5
    return x.astype(dt), np.ones(n, dt)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("where", ["tile_edge", "grid_wrap"])
def test_dense_window_across_tiles(prec, where):
    rng = np.random.default_rng(81)
    M = 1 << 14
    G = 2 * M
    plan = cu.NFFTPlan(dev(np.zeros(1, NP[prec])), M)
    T = 1 << plan.info().tile_log2
    plan.close()
    centre = 11 * T if where == "tile_edge" else 0
    x, y = dense_window(centre, G, 400, rng, NP[prec])
    # a sparse background so that other tiles are not empty
    xb = rng.uniform(-0.5, 0.5, 3000).astype(NP[prec])
    yb = rng.standard_normal(3000).astype(NP[prec])
    x, y = np.concatenate([x, xb]), np.concatenate([y, yb])
    got, info = run(x, y, M, prec)
    assert info.ntiles > 1
    gl = oracle.gl_adjoint(x, y, M, 2, M_OF[prec])
    assert rel_l2(got, gl) < TOL[prec], rel_l2(got, gl)
    ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
    assert rel_l2(got, ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["f", "d"])
def test_dense_window_single_periodic_tile(prec):
    """M = 64: the grid (128 cells) is one tile whose taps wrap onto itself."""
    rng = np.random.default_rng(82)
    M = 64
    G = 2 * M
    x, y = dense_window(0, G, 400, rng, NP[prec])
    got, info = run(x, y, M, prec)
    assert info.ntiles == 1
    gl = oracle.gl_adjoint(x, y, M, 2, M_OF[prec])
    assert rel_l2(got, gl) < TOL[prec], rel_l2(got, gl)
    ref = oracle.ndft(x.astype(np.float64), y.astype(np.float64), M)
    assert rel_l2(got, ref) < TOL[prec]
