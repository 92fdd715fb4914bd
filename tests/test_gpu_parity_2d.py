"""GPU parity of the 2D sweeps against the CPU oracle (bit-exact).

Float kernels are compared with == on the f32 bit patterns: both sides round
every product and sum on its own in the same order, so any difference is a bug.
"""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (2, 9), (3, 3), (5, 7), (64, 64), (67, 131), (129, 257), (300, 1000), (513, 260)]


def _run(stencil, arr, iters, out=None, storage="f32", **kw):
    with api.Grid(arr.shape, storage) as g:
        g.upload(0, arr)
        g.upload(1, arr if out is None else out)
        st = g.run(stencil, iters, **kw)
        return g.download(), st


def _bits_equal(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("variant", ["plain", "tuned"])
def test_jacobi2d(gpu, shape, variant):
    a = inputs.jacobi2d_field(*shape, seed=sum(shape))
    for iters in (0, 1, 2, 7):
        got, st = _run(K.jacobi2d(), a, iters, variant=variant)
        want = oracle.run_2d(K.jacobi2d(), a, iters)
        assert _bits_equal(got, want), (shape, variant, iters)
        assert st["sweeps"] == iters


@pytest.mark.parametrize("shape", [(64, 64), (67, 131), (300, 1000), (513, 260)])
@pytest.mark.parametrize("variant", ["plain", "tuned"])
def test_diffusion9(gpu, shape, variant):
    a = inputs.blobs_field(max(shape), seed=3, nblobs=4, sigma=8.0)[: shape[0], : shape[1]].copy()
    s = K.diffusion9()
    for iters in (1, 6):
        got, _ = _run(s, a, iters, variant=variant)
        want = oracle.run_2d(s, a, iters)
        assert _bits_equal(got, want), (shape, variant, iters)


@pytest.mark.parametrize("block", [(32, 4), (64, 4), (16, 16), (256, 1), (1, 64), (7, 3)])
def test_plain_block_shapes(gpu, block):
    a = inputs.jacobi2d_field(131, 197, seed=5)
    got, _ = _run(K.jacobi2d(), a, 5, variant="plain", block=block)
    assert _bits_equal(got, oracle.run_2d(K.jacobi2d(), a, 5))


@pytest.mark.parametrize("tile", [(0, 1, 1), (0, 2, 2), (0, 4, 4), (0, 8, 2), (0, 16, 1), (0, 8, 4)])
def test_tuned_tile_shapes(gpu, tile):
    a = inputs.jacobi2d_field(300, 700, seed=6)
    got, _ = _run(K.jacobi2d(), a, 3, variant="tuned", tile=tile)
    assert _bits_equal(got, oracle.run_2d(K.jacobi2d(), a, 3))


@pytest.mark.parametrize("stencil_name", ["jacobi2d", "diffusion9"])
@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("cy", [8, 16, 32, 64])
def test_temporal_tiles(gpu, stencil_name, k, cy):
    s = getattr(K, stencil_name)()
    a = inputs.blobs_field(400, seed=7, nblobs=6, sigma=10.0)[:333, :397].copy()
    for iters in (k, 2 * k + 1, 10):
        got, st = _run(s, a, iters, variant="tuned", time_tile=k, tile=(0, cy, 0))
        want = oracle.run_2d(s, a, iters)
        assert _bits_equal(got, want), (stencil_name, k, cy, iters)


def test_temporal_tiles_distinct_rings(gpu):
    """Input and output rings differ: every fused level must read the ring of the
    buffer that level lives in, as the single-sweep ping-pong does."""
    s = K.jacobi2d()
    a = inputs.jacobi2d_field(150, 170, seed=8)
    b = a.copy()
    b[0, :] = 3.0
    b[-1, :] = -2.0
    b[:, 0] = 0.5
    b[:, -1] = 7.0
    for k in (2, 3, 4):
        for iters in (1, 2, 3, 4, 5, 8, 9):
            got, _ = _run(s, a, iters, out=b, variant="tuned", time_tile=k)
            want = oracle.run_2d(s, a, iters, out=b)
            assert _bits_equal(got, want), (k, iters)


def test_radius_larger_than_reach(gpu):
    """An explicit mask range shrinks the interior (Mask range argument)."""
    s = K.jacobi2d()
    s.radius = 3
    a = inputs.jacobi2d_field(40, 50, seed=9)
    for variant in ("plain", "tuned"):
        got, _ = _run(s, a, 4, variant=variant)
        assert _bits_equal(got, oracle.run_2d(s, a, 4))


def test_generic_mask(gpu):
    """A mask with no specialised kernel (radius 2, unusual order) runs the plain kernel."""
    offs = [(0, 2), (-2, 0), (1, 1), (0, -1), (2, -2), (0, 0)]
    s = K.Stencil(K.WEIGHTED_SUM, 2, offs, weights=[0.1, 0.2, 0.15, 0.25, 0.05, 0.25])
    a = inputs.jacobi2d_field(90, 77, seed=10)
    for variant in ("plain", "tuned"):
        got, _ = _run(s, a, 5, variant=variant)
        assert _bits_equal(got, oracle.run_2d(s, a, 5))
    s2 = K.Stencil(K.SCALED_SUM, 2, offs, coeff=1.0 / 6.0)
    got, _ = _run(s2, a, 3)
    assert _bits_equal(got, oracle.run_2d(s2, a, 3))


def test_full_size_jacobi2d(gpu):
    """2048^2 x 100 sweeps: the plain and tuned kernels, bit-exact to the oracle."""
    a = inputs.jacobi2d_field(2048, 2048, seed=1)
    want = oracle.run_2d(K.jacobi2d(), a, 100)
    for kw in ({"variant": "plain", "block": (32, 4)}, {"variant": "tuned"}, {"variant": "tuned", "time_tile": 4}):
        got, _ = _run(K.jacobi2d(), a, 100, **kw)
        assert _bits_equal(got, want), kw
