"""Tuner host logic that needs no device: candidate spaces, features, heuristic."""
import ctypes as C

import numpy as np
import pytest

from pskel_amd import _lib as L
from pskel_amd import tuner as T


def _q(shape, storage, name, it=100):
    return T.Query(tuple(shape), storage, name, it)


def test_2d_march_candidates():
    """The register march joins the 2D Jacobi and 9-point candidates (3 depths
    x 4 segment lengths) on grids wider than 8 cells, not life's."""
    march = lambda q: [c for c in T.candidates(q) if c.tile_x == 4 and c.time_tile >= 2]  # noqa: E731
    for name in ("jacobi2d", "diffusion9"):
        m = march(_q((300, 700), "f32", name))
        assert sorted((c.time_tile, c.tile_z) for c in m) == [(k, s) for k in (2, 3, 4) for s in (0, 8, 24, 64)]
        assert march(_q((300, 8), "f32", name)) == []
    assert march(_q((300, 700), "u8", "life")) == []
    # a mask range above 1: no march (it keeps a ring of one cell)
    q = _q((3000, 3000), "f32", "jacobi2d", 100)
    s = q.stencil()
    s.radius = 3
    cq, d = q.c_query(), s.desc()
    buf = (L.LaunchCfg * L.TUNER_MAX_CANDIDATES)()
    n = C.c_size_t()
    L.check(L.lib().pskel_tuner_candidates(C.byref(cq), C.byref(d), buf, L.TUNER_MAX_CANDIDATES, C.byref(n)))
    assert not any(buf[i].tile_x == 4 and buf[i].time_tile >= 2 for i in range(n.value))
    h = L.LaunchCfg()
    L.check(L.lib().pskel_tuner_heuristic(C.byref(cq), C.byref(d), C.byref(h)))
    assert h.tile_x == 0 and h.time_tile == 4


@pytest.mark.parametrize("q,expect_tuned", [
    (_q((2048, 2048), "f32", "jacobi2d"), True),
    (_q((8194, 8194), "u8", "life"), True),
    (_q((4096, 4096), "f32", "diffusion9"), True),
    (_q((512, 512, 512), "f32", "jacobi3d"), True),
    (_q((100, 200, 300), "bf16", "jacobi3d"), True),
])
def test_candidates_cover_tuned_and_plain(q, expect_tuned):
    c = T.candidates(q)
    assert len(c) >= 10
    assert any(x.variant == L.VARIANT_PLAIN for x in c)
    assert any(x.variant == L.VARIANT_TUNED for x in c) == expect_tuned
    keys = [T.cfg_key(x) for x in c]
    assert len(set(keys)) == len(keys)


def test_3d_candidates_fit_lds():
    for st in ("f32", "bf16"):
        for c in T.candidates(_q((64, 64, 64), st, "jacobi3d")):
            if c.variant == L.VARIANT_TUNED:
                lds = 2 * (4 * c.tile_y + 2) * (64 * c.tile_x + 2 * c.tile_x) * 4
                assert lds <= 64 * 1024
                assert (c.tile_x in (1, 2, 4)) if st == "f32" else (c.tile_x in (2, 4, 8))


def test_3d_fused_candidates_per_storage():
    """Two fused sweeps per launch are offered for the 7-point Jacobi with at
    least two sweeps: the shapes built for each storage (five for 16-bit, all
    also built for f32); three for f32 with at least three."""
    fused = lambda q: [c for c in T.candidates(q) if c.time_tile == 2]  # noqa: E731
    f32 = fused(_q((64, 64, 64), "f32", "jacobi3d", 10))
    b16 = fused(_q((64, 64, 64), "bf16", "jacobi3d", 10))
    assert len(f32) == 7 and len(b16) == 5
    assert {(c.tile_x, c.tile_y, c.block_x, c.block_y) for c in b16} <= \
        {(c.tile_x, c.tile_y, c.block_x, c.block_y) for c in f32}
    assert fused(_q((64, 64, 64), "f32", "jacobi3d", 1)) == []
    assert fused(_q((64, 64, 64), "f16", "jacobi3d", 1)) == []
    for storage in ("bf16", "f16"):
        h = T.heuristic_choice(_q((512, 512, 512), storage, "jacobi3d", 100))
        assert (h.time_tile, h.tile_x, h.tile_y, h.block_x, h.block_y) == (2, 4, 3, 768, 2)
    for storage in ("f32", "bf16", "f16"):
        assert T.heuristic_choice(_q((512, 512, 512), storage, "jacobi3d", 1)).time_tile == 1
    # three sweeps per launch: f32 only, one shape, from three sweeps up
    d3 = lambda q: [c for c in T.candidates(q) if c.time_tile == 3]  # noqa: E731
    assert [(c.tile_x, c.tile_y, c.block_x, c.block_y) for c in d3(_q((64, 64, 64), "f32", "jacobi3d", 3))] == \
        [(4, 4, 512, 2), (4, 2, 768, 2)]
    assert d3(_q((64, 64, 64), "f32", "jacobi3d", 2)) == [] and d3(_q((64, 64, 64), "bf16", "jacobi3d", 9)) == []


def test_3d_heuristic_fits_the_depth():
    """The heuristic takes three fused f32 sweeps when 512-cell tiles are 70%
    full, else the fused pair whose 512-, 256- or 128-cell rows the grid fills
    best, from 20% fill of its tiles up; below, one sweep per
    launch with a 64-lane row as wide as the depth needs and segments short
    enough for about 1024 blocks."""
    h = lambda shape, st: T.heuristic_choice(_q(shape, st, "jacobi3d", 100))  # noqa: E731
    key = lambda c: (c.time_tile, c.tile_x, c.tile_y, c.block_x, c.block_y)  # noqa: E731
    # f32: three sweeps per launch once 512-cell tiles are 70% full
    assert key(h((512, 512, 512), "f32")) == (3, 4, 4, 512, 2)
    assert key(h((479, 323, 406), "f32")) == (3, 4, 4, 512, 2)
    assert key(h((349, 507, 323), "f32"))[0] == 2
    assert key(h((410, 287, 517), "f32"))[0] == 2  # two 512-cell rows would be half empty
    assert key(h((512, 512, 512), "bf16")) == (2, 4, 3, 768, 2)
    assert key(h((401, 258, 252), "f16")) == (2, 4, 3, 768, 1)
    assert key(h((334, 359, 199), "f32")) == (2, 4, 3, 768, 1)
    assert key(h((211, 243, 66), "f32")) == (2, 2, 3, 768, 1)
    assert key(h((269, 487, 42), "f32")) == (2, 2, 3, 768, 1)
    c = h((269, 487, 24), "f32")  # 17% of a 128-cell row
    assert (c.time_tile, c.tile_x, c.tile_y, c.tile_z) == (1, 1, 1, 32)  # 122 tiles x 9 segments
    assert T.heuristic_choice(_q((429, 281, 449), "f16", "jacobi3d", 1)).time_tile == 1


def test_2d_heuristic():
    """Jacobi and 9-point stencils: the register march from 0.9M cells (5-point)
    or 4M cells (9-point), else LDS temporal tiles of 16 rows from 0.8M cells
    and of 8 rows below, at depth 4 (or the sweep count when smaller); life:
    the plain kernel on small grids, bit tiles on wide ones."""
    key = lambda c: (c.tile_x, c.tile_y, c.tile_z, c.time_tile)  # noqa: E731
    h = lambda shape, name, it: key(T.heuristic_choice(_q(shape, "f32", name, it)))  # noqa: E731
    assert h((125, 205), "jacobi2d", 158) == (0, 8, 0, 4)
    assert h((2226, 389), "jacobi2d", 24) == (0, 16, 0, 4)
    assert h((1006, 979), "jacobi2d", 219) == (4, 0, 0, 4)
    assert h((2048, 2048), "jacobi2d", 2) == (4, 0, 0, 2)
    assert h((1024, 1331), "diffusion9", 29) == (0, 16, 0, 4)
    assert h((2673, 908), "diffusion9", 58) == (0, 16, 0, 4)
    assert h((4203, 1196), "diffusion9", 18) == (4, 0, 0, 4)
    assert h((2903, 1983), "diffusion9", 3) == (4, 0, 0, 3)
    assert h((300, 700), "diffusion9", 2) == (0, 8, 0, 2)
    assert T.heuristic_choice(_q((229, 71), "u8", "life", 107)).variant == 0
    assert T.heuristic_choice(_q((8194, 8194), "u8", "life", 500)).time_tile == 8
    assert T.heuristic_choice(_q((1683, 1653), "u8", "life", 29)).time_tile == 1


def test_features_shape_and_determinism():
    q = _q((300, 700), "f32", "jacobi2d", 50)
    c = T.candidates(q)
    f1 = T.features(q, c)
    f2 = T.features(q, c)
    assert f1.shape == (len(c), L.TUNER_NFEAT) and np.array_equal(f1, f2)
    assert np.all(np.isfinite(f1))
    # candidates differ in their features
    assert len({tuple(r) for r in f1}) == len(c)
    # the query part encodes log2 dims and iterations
    assert abs(f1[0, 0] - np.log2(700)) < 1e-5 and abs(f1[0, 1] - np.log2(300)) < 1e-5
    assert abs(f1[0, 8] - np.log2(50)) < 1e-5


def test_heuristic_picks_a_candidate_on_the_workloads():
    for q in (_q((2048, 2048), "f32", "jacobi2d"), _q((512, 512, 512), "f32", "jacobi3d"),
              _q((64, 64), "u8", "life", 3)):
        h = T.heuristic_choice(q)
        assert T.cfg_key(h) in {T.cfg_key(c) for c in T.candidates(q)}


def test_random_queries_reproducible():
    a = T.random_queries(12, 7)
    b = T.random_queries(12, 7)
    assert a == b
    for q in a:
        assert all(s >= 32 for s in q.shape) and 10 <= q.iterations <= 1000


def test_heuristic_picks_a_candidate():
    """On random queries of every kind (the training distribution) the
    heuristic's launch is one of the tuner's own candidates, so the trainer
    can price it from the exhaustive table."""
    for q in T.random_queries(200, 7) + [_q((512, 512, 512), s, "jacobi3d", it) for s in ("f32", "bf16", "f16")
                                           for it in (1, 2, 1000)]:
        h = T.heuristic_choice(q)
        assert T.cfg_key(h) in {T.cfg_key(c) for c in T.candidates(q)}, (q, T.cfg_key(h))


def test_packed_ensemble_is_the_members_mean():
    """pack_params: the block-diagonal model's output equals the mean of the
    members' outputs (FP64 forward of the oracle), for 1, 3 and 5 members and
    two and three hidden layers."""
    import oracle
    from pskel_amd.tuner import pack_params

    rng = np.random.default_rng(5)
    for sizes in ([26, 64, 64, 1], [7, 5, 3, 4, 1]):
        npar = sum(sizes[l + 1] * sizes[l] + sizes[l + 1] for l in range(len(sizes) - 1))
        x = rng.standard_normal((40, sizes[0]))
        for M in (1, 3, 5):
            ps = [rng.standard_normal(npar).astype(np.float32) * 0.3 for _ in range(M)]
            big, flat = pack_params(ps, sizes)
            assert big == [sizes[0]] + [M * h for h in sizes[1:-1]] + [1]
            want = np.mean([oracle.mlp_predict(sizes, p.astype(np.float64), x) for p in ps], axis=0)
            got = oracle.mlp_predict(big, flat.astype(np.float64), x)
            np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


def test_saved_timing_rows_round_trip():
    """rows_to_json names each candidate by its launch configuration; loading
    maps it back to the current candidate index and lists what is missing."""
    qs = T.random_queries(8, 3)
    rows = [{"q": q, "ci": i, "ms": 1.0 + i} for q in qs for i in range(len(T.candidates(q)))]
    js = T.rows_to_json(rows)
    back, missing = T.rows_from_json(js)
    assert not missing and len(back) == len(rows)
    assert all(a["ci"] == b["ci"] and a["q"] == b["q"] and a["ms"] == b["ms"] for a, b in zip(rows, back))
    back, missing = T.rows_from_json(js[:-3])
    assert len(back) == len(rows) - 3 and len(missing) == 3
    assert {T.cfg_key(c) for _, c in missing} == {tuple(x["cfg"]) for x in js[-3:]}
    # a configuration that is no longer a candidate is dropped
    js[0] = dict(js[0], cfg=[99, 9, 9, 9, 9, 9, 9])
    back, missing = T.rows_from_json(js)
    assert len(back) == len(rows) - 1 and len(missing) == 1


def test_batch_family_choices():
    """The batch search's family choices and march segments (no device)."""
    from pskel_amd.tuner import batch as TB

    assert TB.FAMILY_CHOICES[0] == (None, None) and len(TB.FAMILY_CHOICES) == 13
    shapes = [(64, 64, "jacobi"), (300, 8, "jacobi"), (70, 1030, "life"), (129, 513, "diffusion9")]
    assert TB.march_segments(shapes, TB.MARCH) == [TB.MARCH_SEG, 0, 0, TB.MARCH_SEG]
    assert TB.march_segments(shapes, 3) == [0, 0, 0, 0]
    for kind in ("jacobi", "diffusion9", "life"):
        depths = sorted({max(1, int(c.time_tile)) for c in TB.candidates(kind)})
        assert depths == ([1, 3, 5, 7] if kind == "life" else [1, 3])
        assert TB.query(100, 200, kind, 50).storage == ("u8" if kind == "life" else "f32")
