"""GPU parity of Game of Life (u8, B3/S23) against the CPU oracle: bit-exact."""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K

pytestmark = pytest.mark.gpu


def _run(arr, iters, out=None, **kw):
    with api.Grid(arr.shape, "u8") as g:
        g.upload(0, arr)
        g.upload(1, arr if out is None else out)
        g.run(K.life(), iters, **kw)
        return g.download()


@pytest.mark.parametrize("n", [1, 3, 14, 62, 100, 1022, 1023, 1030])
@pytest.mark.parametrize("variant", ["plain", "tuned"])
def test_life_sizes(gpu, n, variant):
    b = inputs.life_field(n, seed=n)
    for iters in (0, 1, 2, 9):
        assert np.array_equal(_run(b, iters, variant=variant), oracle.run_2d(K.life(), b, iters)), (n, iters)


@pytest.mark.parametrize("shape", [(5, 2000), (2000, 5), (333, 2049), (77, 4100)])
def test_life_ragged(gpu, shape):
    b = (inputs.uniform(21, 0, shape[0] * shape[1]).reshape(shape) < 0.35).astype(np.uint8)
    for variant in ("plain", "tuned"):
        assert np.array_equal(_run(b, 7, variant=variant), oracle.run_2d(K.life(), b, 7)), variant


@pytest.mark.parametrize("ry,nwy", [(1, 1), (2, 4), (4, 4), (8, 2), (16, 1)])
def test_life_tuned_shapes(gpu, ry, nwy):
    b = inputs.life_field(700, seed=3)
    assert np.array_equal(_run(b, 5, variant="tuned", tile=(0, ry, nwy)), oracle.run_2d(K.life(), b, 5))


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("cy", [16, 32, 48, 64])
def test_life_temporal(gpu, k, cy):
    """k generations per launch on bit-packed tiles of cy rows per lane, rounded
    up to 24, 32, 40, 48 or 64 and to at least 2k + 8."""
    b = inputs.life_field(500, seed=4)
    for iters in (k, 2 * k + 1, 11):
        got = _run(b, iters, variant="tuned", time_tile=k, tile=(0, cy, 0))
        assert np.array_equal(got, oracle.run_2d(K.life(), b, iters)), (k, cy, iters)


def test_life_distinct_rings(gpu):
    b = inputs.life_field(200, seed=5)
    o = b.copy()
    o[0, ::3] = 1
    o[-1, ::2] = 1
    o[:, 0] = 1
    o[:, -1] = 1
    o[5:9, -1] = 0
    for kw in ({"variant": "plain"}, {"variant": "tuned"}, {"variant": "tuned", "time_tile": 2},
               {"variant": "tuned", "time_tile": 3}, {"variant": "tuned", "time_tile": 7},
               {"variant": "tuned", "time_tile": 8}):
        for iters in (1, 2, 5, 6, 15, 16):
            assert np.array_equal(_run(b, iters, out=o, **kw), oracle.run_2d(K.life(), b, iters, out=o)), (kw, iters)


@pytest.mark.parametrize("shape", [(1, 1), (2, 40), (3, 3), (40, 2), (33, 33), (70, 1985), (41, 1986),
                                   (130, 4000), (2, 4100), (4100, 3)])
@pytest.mark.parametrize("k", [2, 5, 8])
def test_life_temporal_ragged(gpu, shape, k):
    """Tile, lane and word boundaries against the grid's edges, distinct rings."""
    b = (inputs.uniform(31, 0, shape[0] * shape[1]).reshape(shape) < 0.4).astype(np.uint8)
    o = (inputs.uniform(32, 0, shape[0] * shape[1]).reshape(shape) < 0.5).astype(np.uint8)
    for iters in (k, k + 3):
        for cy in (0, 24):  # default 32-row tiles, and 24-row tiles (8 core rows at k = 8)
            got = _run(b, iters, out=o, variant="tuned", time_tile=k, tile=(0, cy, 0))
            assert np.array_equal(got, oracle.run_2d(K.life(), b, iters, out=o)), (shape, k, iters, cy)


def test_life_rejects_non_binary_cells(gpu):
    from pskel_amd._lib import PSkelError

    b = inputs.life_field(16, seed=6)
    b[3, 3] = 2
    with api.Grid(b.shape, "u8") as g:
        with pytest.raises(PSkelError):
            g.upload(0, b)


def test_life_full_size(gpu):
    """8192^2 board (8194^2 with the dead ring) x 40 generations, all variants."""
    b = inputs.life_field(8192, seed=2)
    want = oracle.run_2d(K.life(), b, 40)
    for kw in ({"variant": "plain", "block": (64, 4)}, {"variant": "tuned"}, {"variant": "tuned", "time_tile": 4},
               {"variant": "tuned", "time_tile": 8, "tile": (0, 48, 0)}):
        assert np.array_equal(_run(b, 40, **kw), want), kw
