"""Batched persistent sweeps of uneven grids: every grid bit-exact to the oracle."""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K
from pskel_amd._lib import PSkelError

pytestmark = pytest.mark.gpu


def _make(shapes, seed=0):
    grids, stencils, fields = [], [], []
    for i, (h, w, kind) in enumerate(shapes):
        if kind == "life":
            f = (inputs.uniform(seed + i, 0, h * w).reshape(h, w) < 0.3).astype(np.uint8)
            g = api.Grid((h, w), "u8")
            st = K.life()
        else:
            f = inputs.jacobi2d_field(h, w, seed=seed + i)
            g = api.Grid((h, w), "f32")
            st = K.jacobi2d() if kind == "jacobi" else K.diffusion9()
        g.upload_pair(f)
        grids.append(g)
        stencils.append(st)
        fields.append(f)
    return grids, stencils, fields


SHAPES = [(64, 64, "jacobi"), (70, 1030, "life"), (257, 66, "jacobi"), (1000, 300, "life"),
          (3, 9, "jacobi"), (129, 513, "diff9"), (2048, 64, "life"), (333, 777, "jacobi")]


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("rows", [None, [1, 2, 4, 8, 1, 2, 4, 8]])
def test_batch_matches_oracle(gpu, use_graph, rows):
    grids, stencils, fields = _make(SHAPES)
    try:
        with api.Batch(grids, stencils, rows) as b:
            info = b.info()
            assert info["items"] >= len(grids) and 1 <= info["blocks"] <= info["items"]
            b.run(7, use_graph=use_graph)
            b.run(6, use_graph=use_graph)  # resume: 13 sweeps in total
            for g, s, f in zip(grids, stencils, fields):
                assert np.array_equal(g.download(), oracle.run_2d(s, f, 13)), g.shape
    finally:
        for g in grids:
            g.close()


DEPTH_SHAPES = [(64, 64, "jacobi"), (70, 1030, "life"), (257, 66, "jacobi"), (1000, 300, "life"), (3, 9, "jacobi"),
                (129, 513, "diff9"), (2048, 64, "life"), (333, 777, "jacobi"), (40, 2100, "life"), (500, 130, "diff9"),
                (2, 2, "life"), (5, 3000, "diff9")]
DEPTHS = [3, 7, 1, 5, 3, 3, 3, 3, 7, 1, 5, 3]
ROWS = [16, 0, 4, 0, 32, 16, 0, 32, 0, 2, 0, 16]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_temporal_depths(gpu, use_graph):
    """Per-grid depths (3 for f32 LDS tiles, 3/5/7 for bit-packed life, 1 for
    register tiles) and runs of 13, 8 and 1 sweeps, whose remainders take no,
    one or two tail launches: every grid equals the oracle after each run."""
    grids, stencils, fields = _make(DEPTH_SHAPES, seed=7)
    try:
        with api.Batch(grids, stencils, ROWS, DEPTHS) as b:
            done = 0
            for n in (13, 8, 1, 14):
                st = b.run(n, use_graph=use_graph)
                done += n
                assert st["launches"] == max(-(-n // k) + (1 if n % k and (n % k) % 2 == 0 else 0) for k in DEPTHS)
                for g, s, f, k in zip(grids, stencils, fields, DEPTHS):
                    assert np.array_equal(g.download(), oracle.run_2d(s, f, done)), (g.shape, k, done)
    finally:
        for g in grids:
            g.close()


MARCH_SHAPES = [(300, 700, "jacobi"), (129, 513, "diff9"), (9, 20, "jacobi"), (1000, 260, "diff9"),
                (64, 64, "jacobi"), (70, 1030, "life"), (257, 66, "jacobi"), (40, 2100, "diff9")]
MARCH_DEPTHS = [3, 3, 3, 3, 3, 7, 1, 3]
MARCH_ROWS = [0, 0, 0, 0, 16, 0, 4, 0]
MARCH_SEGS = [32, 8, 32, 100, 0, 0, 0, 5]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_register_march(gpu, use_graph):
    """f32 grids through the register march (segments of 5 to 100 rows, tiny
    and wide grids), beside LDS tiles, bit tiles and register tiles; runs of
    13, 8, 1 and 14 sweeps take no, one or two single-sweep tail launches."""
    grids, stencils, fields = _make(MARCH_SHAPES, seed=11)
    try:
        with api.Batch(grids, stencils, MARCH_ROWS, MARCH_DEPTHS, MARCH_SEGS) as b:
            done = 0
            for n in (13, 8, 1, 14):
                b.run(n, use_graph=use_graph)
                done += n
                for g, s, f, m in zip(grids, stencils, fields, MARCH_SEGS):
                    assert np.array_equal(g.download(), oracle.run_2d(s, f, done)), (g.shape, m, done)
    finally:
        for g in grids:
            g.close()


def test_batch_search(gpu):
    """The batch-level family search (pskel_amd.tuner.batch.search) on scratch
    grids: every family choice is timed, the fastest plan is returned, and a
    batch run with it is bit-exact."""
    from pskel_amd.tuner import batch as TB

    shapes = [(h, w, "life" if kind == "life" else "jacobi") for h, w, kind in MARCH_SHAPES]
    res = TB.search(shapes, 12, reps=1)
    assert res is not None and len(res["table"]) == len(TB.FAMILY_CHOICES)
    assert abs(res["ms_per_run"] - min(r["ms_per_run"] for r in res["table"])) < 1e-3  # the table rounds to 1 us
    assert len(res["rows"]) == len(res["depths"]) == len(res["march"]) == len(shapes)
    grids, stencils, fields = _make(shapes, seed=13)
    try:
        with api.Batch(grids, stencils, res["rows"], res["depths"], res["march"]) as b:
            b.run(12, use_graph=True)
        for g, s, f in zip(grids, stencils, fields):
            assert np.array_equal(g.download(), oracle.run_2d(s, f, 12)), g.shape
    finally:
        for g in grids:
            g.close()


def test_batch_march_validation(gpu):
    grids, stencils, _ = _make([(20, 8, "jacobi"), (64, 64, "jacobi")], seed=12)
    try:
        for depths, segs in (([3, 3], [32, 0]), ([1, 5], [0, 32])):
            with pytest.raises(PSkelError):
                api.Batch(grids, stencils, None, depths, segs)
    finally:
        for g in grids:
            g.close()


def test_batch_depth_validation(gpu):
    grids, stencils, _ = _make([(64, 64, "jacobi"), (64, 64, "life")], seed=9)
    try:
        for depths, rows in (([2, 1], None), ([5, 1], None), ([1, 4], None), ([3, 1], [8, 4]), ([1, 9], None)):
            with pytest.raises(PSkelError):
                api.Batch(grids, stencils, rows, depths)
    finally:
        for g in grids:
            g.close()


def test_batch_graph_replays(gpu):
    """The same captured graph replayed several times, with eager work between."""
    grids, stencils, fields = _make(SHAPES[:5], seed=3)
    try:
        with api.Batch(grids, stencils) as b:
            for rep in range(1, 4):
                b.run(10, use_graph=True)
                grids[0].run(stencils[0], 1)  # eager sweep of one grid between replays
                api.synchronize()
                for j, (g, s, f) in enumerate(zip(grids, stencils, fields)):
                    n = 10 * rep + (rep if j == 0 else 0)
                    assert np.array_equal(g.download(), oracle.run_2d(s, f, n)), (rep, j)
    finally:
        for g in grids:
            g.close()


def test_batch_rejects_3d_and_generic(gpu):
    g3 = api.Grid((8, 8, 8), "f32")
    g2 = api.Grid((32, 32), "f32")
    try:
        with pytest.raises(PSkelError):
            api.Batch([g3], [K.jacobi3d()])
        odd = K.Stencil(K.SCALED_SUM, 2, [(0, 2), (0, -2)], coeff=0.5)
        with pytest.raises(PSkelError):
            api.Batch([g2], [odd])
    finally:
        g3.close()
        g2.close()


def test_batch_256_uneven_grids(gpu):
    """The full workload shape: 256 grids, sides log-uniform in [64, 2048],
    half Jacobi f32 half life u8, 200 sweeps in one graph; all bit-exact."""
    shapes = inputs.batch_shapes(256, seed=5)
    grids, stencils, fields = _make(shapes, seed=100)
    try:
        with api.Batch(grids, stencils) as b:
            st = b.run(200, use_graph=True)
            assert st["launches"] == 200
        for i in range(0, 256, 17):  # spot-check a spread of grids (the oracle is the slow part)
            assert np.array_equal(grids[i].download(), oracle.run_2d(stencils[i], fields[i], 200)), shapes[i]
    finally:
        for g in grids:
            g.close()
