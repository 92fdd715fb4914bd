"""GPU parity of the 2D register march (k2d_march: 2..4 sweeps per launch
pipelined in registers, cfg.tile_x = 4) against the CPU oracle, bit-exact."""
import numpy as np
import pytest

import oracle
from pskel_amd import _lib as L
from pskel_amd import api, inputs
from pskel_amd import kernels as K

pytestmark = pytest.mark.gpu


def _run(stencil, arr, iters, out=None, **kw):
    with api.Grid(arr.shape, "f32") as g:
        g.upload(0, arr)
        g.upload(1, arr if out is None else out)
        st = g.run(stencil, iters, variant="tuned", **kw)
        return g.download(), st


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("stencil_name", ["jacobi2d", "diffusion9"])
@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("seg", [0, 5, 16])
def test_register_march(gpu, stencil_name, k, seg):
    """tile = (4 cells per lane, 0, rows per segment; 0: automatic)."""
    s = getattr(K, stencil_name)()
    a = inputs.blobs_field(400, seed=7, nblobs=6, sigma=10.0)[:333, :397].copy()
    for iters in (k, 2 * k + 1, 10):
        got, st = _run(s, a, iters, time_tile=k, tile=(4, 0, seg))
        assert _bits_equal(got, oracle.run_2d(s, a, iters)), (stencil_name, k, seg, iters)
        assert st["sweeps"] == iters


# narrow and short grids, tile edges (248 cells per tile), no interior rows
SHAPES = [(3, 9), (4, 20), (2, 50), (9, 300), (300, 9), (70, 248), (70, 249), (41, 250), (17, 497), (66, 744),
          (130, 745), (1000, 17)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("stencil_name", ["jacobi2d", "diffusion9"])
def test_register_march_shapes(gpu, shape, stencil_name):
    s = getattr(K, stencil_name)()
    a = inputs.jacobi2d_field(*shape, seed=sum(shape))
    for k, seg in ((4, 0), (3, 7), (2, 1)):
        for iters in (k, 9):
            got, _ = _run(s, a, iters, time_tile=k, tile=(4, 0, seg))
            assert _bits_equal(got, oracle.run_2d(s, a, iters)), (shape, k, seg, iters)


@pytest.mark.parametrize("stencil_name", ["jacobi2d", "diffusion9"])
def test_register_march_distinct_rings(gpu, stencil_name):
    """Input and output rings differ: each fused sweep reads the ring of the
    buffer it lives in (ring rows from the buffers, ring columns carried)."""
    s = getattr(K, stencil_name)()
    for shape in ((150, 170), (37, 520), (12, 13)):
        a = inputs.jacobi2d_field(*shape, seed=8)
        b = a.copy()
        b[0, :] = 3.0
        b[-1, :] = -2.0
        b[:, 0] = 0.5
        b[:, -1] = 7.0
        for k in (2, 3, 4):
            for seg in (0, 3):
                for iters in (1, 2, 3, 4, 5, 8, 9):
                    got, _ = _run(s, a, iters, out=b, time_tile=k, tile=(4, 0, seg))
                    want = oracle.run_2d(s, a, iters, out=b)
                    assert _bits_equal(got, want), (shape, k, seg, iters)


def test_register_march_residuals_and_graph(gpu):
    """Residual steps split the fused launches; a captured graph replays the
    same launches (three replays, each checked)."""
    s = K.diffusion9()
    a = inputs.blobs_field(300, seed=3, nblobs=4, sigma=8.0)[:257, :300].copy()
    got, st = _run(s, a, 24, time_tile=4, tile=(4, 0, 0), residual_every=8)
    assert _bits_equal(got, oracle.run_2d(s, a, 24))
    assert len(st["residuals"]) == 3
    want = [oracle.run_2d(s, a, 12 * r) for r in (1, 2, 3)]
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        for r in range(3):
            g.run(s, 12, variant="tuned", time_tile=4, tile=(4, 0, 0), use_graph=True)
            api.synchronize()
            assert _bits_equal(g.download(), want[r]), r


def test_register_march_refused(gpu):
    """Shapes the march does not build are refused, not run."""
    s = K.jacobi2d()
    a = inputs.jacobi2d_field(40, 40, seed=1)
    for tile in ((2, 0, 0), (8, 0, 0), (4, 0, -1), (4, 3, 0)):
        with pytest.raises(L.PSkelError) as e:
            _run(s, a, 4, time_tile=4, tile=tile)
        assert e.value.status == L.ERR_UNSUPPORTED
    with pytest.raises(L.PSkelError):
        _run(s, inputs.jacobi2d_field(40, 8, seed=1), 4, time_tile=4, tile=(4, 0, 0))
    # a mask range above 1 widens the ring the march does not keep
    s.radius = 3
    with pytest.raises(L.PSkelError) as e:
        _run(s, a, 4, time_tile=4, tile=(4, 0, 0))
    assert e.value.status == L.ERR_UNSUPPORTED
    got, _ = _run(s, a, 4, time_tile=4)  # the LDS tiles keep any range
    assert _bits_equal(got, oracle.run_2d(s, a, 4))


def test_register_march_full_size(gpu):
    """Jacobi-2D 2048^2 x100 and the 9-point stencil on 4096^2 x12 through the
    march, bit-exact to the oracle."""
    a = inputs.jacobi2d_field(2048, 2048, seed=1)
    want = oracle.run_2d(K.jacobi2d(), a, 100)
    for k in (2, 3, 4):
        got, _ = _run(K.jacobi2d(), a, 100, time_tile=k, tile=(4, 0, 0))
        assert _bits_equal(got, want), k
    b = inputs.blobs_field(4096, seed=2, nblobs=32, sigma=40.0)
    want = oracle.run_2d(K.diffusion9(), b, 12)
    got, _ = _run(K.diffusion9(), b, 12, time_tile=4, tile=(4, 0, 32))
    assert _bits_equal(got, want)
