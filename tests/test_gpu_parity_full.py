"""Parity at the full sizes of the benchmark workloads, against the oracle.

* 9-point diffusion 4096^2 (the workload's grid): plain, tuned, and 4-sweep
  temporal tiles of 16 rows, 12 sweeps, bit-exact; and the workload itself,
  1000 sweeps through the register march and the LDS tiles, bit-exact.
* Jacobi-3D 512^3 with bf16 and f16 storage, stochastic rounding, the fused
  pair (two sweeps per launch), 6 sweeps, bit-exact against the oracle's
  16-bit runner from the same initial words, FP64 residuals to 1e-12.
* The 256-grid batch: every grid after 200 sweeps in one graph, bit-exact.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K

pytestmark = pytest.mark.gpu


def test_diffusion9_4096(gpu):
    s = K.diffusion9()
    a = inputs.blobs_field(4096, seed=3)
    want = oracle.run_2d(s, a, 12)
    for kw in ({"variant": "plain", "block": (32, 4)}, {"variant": "tuned"},
               {"variant": "tuned", "time_tile": 4, "tile": (0, 16, 0)}):
        with api.Grid(a.shape, "f32") as g:
            g.upload_pair(a)
            g.run(s, 12, **kw)
            got = g.download()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), kw


def test_diffusion9_4096_x1000(gpu):
    """The 9-point workload as configured (4096^2, 32 blobs, 1000 sweeps):
    4 sweeps per launch in registers (the tuner's launch) and in LDS tiles."""
    s = K.diffusion9()
    a = inputs.blobs_field(4096, seed=3)
    want = oracle.run_2d(s, a, 1000)
    for kw in ({"time_tile": 4, "tile": (4, 0, 0)}, {"time_tile": 4, "tile": (0, 16, 0)}):
        with api.Grid(a.shape, "f32") as g:
            g.upload_pair(a)
            g.run(s, 1000, variant="tuned", **kw)
            got = g.download()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), kw


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_jacobi3d_512_16bit_fused_stochastic(gpu, store):
    code = oracle.STORE_BF16 if store == "bf16" else oracle.STORE_F16
    a = inputs.jacobi3d_field(512, seed=4)
    off = 0.5
    with api.Grid(a.shape, store, store_offset=off, rounding="stochastic", seed=17) as g:
        g.upload_pair(a)
        words = g.download(raw=True)  # the initial state as stored (u - off, to nearest)
        st = g.run(K.jacobi3d(), 6, variant="tuned", tile=(4, 3, 0), block=(768, 2), time_tile=2,
                   residual_every=2)
        got = g.download(raw=True)
    assert st["launches"] == 3
    want, want_res = oracle.run_3d_r16(K.jacobi3d(), code, words, 6, store_offset=off,
                                       rounding=oracle.ROUND_STOCHASTIC, seed=17, residual_every=2)
    assert np.array_equal(got, want)
    np.testing.assert_allclose(np.array(st["residuals"]), want_res, rtol=1e-12, atol=0)


def test_batch_256_every_grid(gpu):
    """All 256 grids of the batched workload (sides log-uniform in [64, 2048],
    half Jacobi f32, half life u8), 200 sweeps in one hipGraph: at depth 1,
    at life 7 generations per launch on bit tiles with Jacobi 3 sweeps on LDS
    tiles, and with Jacobi 3 sweeps through the register march (32-row
    segments), the families the benchmark measures best."""
    shapes = inputs.batch_shapes(256, seed=5)
    grids, stencils, fields = [], [], []
    try:
        for i, (h, w, kind) in enumerate(shapes):
            if kind == "life":
                f = (inputs.uniform(1000 + i, 0, h * w).reshape(h, w) < np.float32(0.3)).astype(np.uint8)
                g, s = api.Grid((h, w), "u8"), K.life()
            else:
                f = inputs.jacobi2d_field(h, w, seed=1000 + i)
                g, s = api.Grid((h, w), "f32"), K.jacobi2d()
            g.upload_pair(f)
            grids.append(g)
            stencils.append(s)
            fields.append(f)
        with api.Batch(grids, stencils) as b:
            st = b.run(200, use_graph=True)
            assert st["launches"] == 200
        got = [g.download() for g in grids]
        # four oracle calls at a time, four OpenMP threads each (ctypes drops the GIL)
        with cf.ThreadPoolExecutor(4) as ex:
            want = list(ex.map(lambda i: oracle.run_2d(stencils[i], fields[i], 200, threads=4), range(256)))
        bad = [i for i in range(256) if not np.array_equal(got[i].view(np.uint8), want[i].view(np.uint8))]
        assert not bad, [shapes[i] for i in bad]
        for g, f in zip(grids, fields):
            g.upload_pair(f)
        rows = [32 if kind == "life" else 16 for _, _, kind in shapes]
        depths = [7 if kind == "life" else 3 for _, _, kind in shapes]
        with api.Batch(grids, stencils, rows, depths) as b:
            b.run(200, use_graph=True)
        got = [g.download() for g in grids]
        bad = [i for i in range(256) if not np.array_equal(got[i].view(np.uint8), want[i].view(np.uint8))]
        assert not bad, [shapes[i] for i in bad]
        for g, f in zip(grids, fields):
            g.upload_pair(f)
        march = [0 if kind == "life" else 32 for _, _, kind in shapes]
        with api.Batch(grids, stencils, rows, depths, march) as b:
            b.run(200, use_graph=True)
        got = [g.download() for g in grids]
        bad = [i for i in range(256) if not np.array_equal(got[i].view(np.uint8), want[i].view(np.uint8))]
        assert not bad, [shapes[i] for i in bad]
    finally:
        for g in grids:
            g.close()
