"""Two or three fused 3D sweeps per launch (time_tile=2 or 3, 7-point Jacobi)
against the CPU oracle, bit-exact: the fused kernel evaluates every sweep with
the single sweep's expression, and with 16-bit storage rounds the first sweep to storage
(nearest, or stochastically with that sweep's own words) before the second
reads it, so any split of N sweeps into launches gives the same words.
Residuals of fused launches compare ||u^k - u^(k-1)|| against the oracle's to
1e-12 relative (different summation order)."""
import dataclasses

import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K
from pskel_amd._lib import PSkelError

pytestmark = pytest.mark.gpu

# (cells per lane, rows per wave, waves along d, waves along w, planes prefetched):
# the instantiated fused-tile shapes
SHAPES_T2 = [(2, 4, 4, 4, 1), (2, 4, 4, 2, 1), (4, 2, 2, 8, 1), (4, 2, 2, 4, 1), (2, 4, 4, 4, 2), (2, 4, 2, 4, 1),
             (4, 4, 2, 4, 1), (4, 3, 2, 6, 1), (4, 3, 1, 12, 1), (2, 3, 1, 12, 1), (4, 2, 2, 6, 1)]

# grids: tiny, one tile, ragged, many w tiles, several d tiles (D > 64 * 8)
GRIDS = [(1, 1, 1), (2, 3, 4), (3, 3, 3), (5, 9, 70), (17, 9, 70), (33, 40, 65), (20, 70, 300), (9, 23, 1100)]


def _field(shape, seed):
    H, W, D = shape
    a = inputs.uniform(seed, 0, H * W * D).reshape(shape)
    a[0] = a[-1] = 1.0
    a[:, 0] = a[:, -1] = 1.0
    a[:, :, 0] = a[:, :, -1] = 1.0
    return a


def _block(nwd, nww, pf):
    """cfg.block_x = threads, cfg.block_y = waves along d + 16 * (prefetch - 1)"""
    return (64 * nwd * nww, nwd + 16 * (pf - 1))


def _run(a, iters, tshape, per, **kw):
    vd, rw, nwd, nww, pf = tshape
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        st = g.run(K.jacobi3d(), iters, variant="tuned", block=_block(nwd, nww, pf), tile=(vd, rw, per),
                   time_tile=2, **kw)
        return g.download(), st


@pytest.mark.parametrize("tshape", SHAPES_T2, ids=["vd{}_rw{}_d{}_w{}_pf{}".format(*t) for t in SHAPES_T2])
@pytest.mark.parametrize("grid", GRIDS, ids=["x".join(map(str, g)) for g in GRIDS])
def test_fused_pair_matches_oracle(gpu, tshape, grid):
    a = _field(grid, sum(grid) + tshape[0])
    for per in (3, 0):
        for iters in (1, 2, 5):
            got, st = _run(a, iters, tshape, per)
            want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (tshape, grid, per, iters)
            assert st["launches"] == (iters + 1) // 2


@pytest.mark.parametrize("tshape", [(2, 4, 4, 4, 1), (4, 2, 2, 8, 1), (2, 4, 4, 4, 2), (4, 3, 2, 6, 1)])
def test_fused_many_tiles_and_segments(gpu, tshape):
    """Many w tiles, several h segments of odd length, N even and odd."""
    shape = (70, 130, 520)
    a = _field(shape, 11)
    # planes per block: fixed segments and the default, chosen from the
    # device's resident blocks
    for per, iters in ((7, 6), (32, 9), (100, 4), (0, 5)):
        got, _ = _run(a, iters, tshape, per)
        want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (per, iters)


def test_fused_wider_range(gpu):
    """A declared mask range of 2 (faces two cells deep) with the 7-point offsets."""
    s = dataclasses.replace(K.jacobi3d(), radius=2)
    shape = (21, 37, 90)
    a = _field(shape, 12)
    for iters in (2, 5):
        with api.Grid(shape, "f32") as g:
            g.upload_pair(a)
            g.run(s, iters, variant="tuned", time_tile=2)
            got = g.download()
        want, _ = oracle.run_3d(s, a, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), iters


@pytest.mark.parametrize("every", [2, 4, 5])
def test_fused_residuals(gpu, every):
    shape = (40, 36, 100)
    a = _field(shape, 6)
    iters = 20
    got_a, st = _run(a, iters, (2, 4, 4, 4, 1), 16, residual_every=every)
    want, want_res = oracle.run_3d(K.jacobi3d(), a, iters, residual_every=every)
    assert np.array_equal(got_a.view(np.uint32), want.view(np.uint32))
    got = np.array(st["residuals"])
    assert got.shape == want_res.shape == (iters // every,)
    np.testing.assert_allclose(got, want_res, rtol=1e-12, atol=0)


def test_fused_split_runs_and_graph(gpu):
    """run(3); run(4) and graph replays equal run(7), and equal single sweeps."""
    shape = (30, 50, 200)
    a = _field(shape, 13)
    want, _ = oracle.run_3d(K.jacobi3d(), a, 7)
    with api.Grid(shape, "f32") as g:
        g.upload_pair(a)
        g.run(K.jacobi3d(), 3, variant="tuned", time_tile=2)
        g.run(K.jacobi3d(), 4, variant="tuned", time_tile=2)
        assert np.array_equal(g.download().view(np.uint32), want.view(np.uint32))
    want14, _ = oracle.run_3d(K.jacobi3d(), a, 14)
    with api.Grid(shape, "f32") as g:
        g.upload_pair(a)
        for rep in range(2):  # capture, then replay
            g.run(K.jacobi3d(), 7, variant="tuned", time_tile=2, use_graph=True)
            w = want if rep == 0 else want14
            assert np.array_equal(g.download().view(np.uint32), w.view(np.uint32)), rep


def test_fused_full_size(gpu):
    """512^3 x 6 sweeps in three fused launches, bit-exact."""
    a = inputs.jacobi3d_field(512, seed=4)
    want, _ = oracle.run_3d(K.jacobi3d(), a, 6)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        st = g.run(K.jacobi3d(), 6, variant="tuned", time_tile=2)
        got = g.download()
    assert st["launches"] == 3
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_fused_rejects(gpu):
    a = _field((10, 10, 10), 14)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        with pytest.raises(PSkelError):
            g.run(K.jacobi3d(), 4, variant="tuned", time_tile=4)
        with pytest.raises(PSkelError):
            g.run(K.jacobi3d(), 4, variant="tuned", tile=(8, 3, 16), time_tile=2)
        with pytest.raises(PSkelError):  # not an instantiated shape
            g.run(K.jacobi3d(), 4, variant="tuned", block=_block(4, 2, 2), tile=(2, 4, 16), time_tile=2)
        with pytest.raises(PSkelError):  # threads not a whole count of waves along d
            g.run(K.jacobi3d(), 4, variant="tuned", block=(64 * 6, 4), tile=(2, 4, 16), time_tile=2)
        # depth 3 exists for two shapes (4 cells x 4 rows, 2 x 4 waves; 4 cells x 2 rows, 2 x 6 waves)
        with pytest.raises(PSkelError):
            g.run(K.jacobi3d(), 4, variant="tuned", time_tile=3, block=_block(2, 6, 1), tile=(4, 3, 0))
    with api.Grid(a.shape, "bf16") as g:
        g.upload_pair(a)
        with pytest.raises(PSkelError):
            g.run(K.jacobi3d(), 4, variant="tuned", time_tile=3)


# three fused sweeps per launch (f32): the instantiated shapes
SHAPES_T3 = [(4, 4, 2, 4, 1), (4, 2, 2, 6, 1)]
T3 = SHAPES_T3[0]


def _run3(a, iters, per, tshape=T3, **kw):
    vd, rw, nwd, nww, pf = tshape
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        st = g.run(K.jacobi3d(), iters, variant="tuned", block=_block(nwd, nww, pf), tile=(vd, rw, per),
                   time_tile=3, **kw)
        return g.download(), st


@pytest.mark.parametrize("tshape", SHAPES_T3, ids=["vd{}_rw{}_d{}_w{}_pf{}".format(*t) for t in SHAPES_T3])
@pytest.mark.parametrize("grid", GRIDS + [(40, 47, 515)], ids=["x".join(map(str, g)) for g in GRIDS + [(40, 47, 515)]])
def test_fused_depth3_matches_oracle(gpu, tshape, grid):
    """Three sweeps per launch; runs of 1 .. 7 sweeps end with a pair or a
    single sweep.  Tiles overlap by 4 rows along w and one lane vector along d."""
    a = _field(grid, sum(grid) + 3)
    for per in (3, 0):
        for iters in (1, 2, 3, 4, 5, 7):
            got, st = _run3(a, iters, per, tshape)
            want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (grid, per, iters)
            assert st["launches"] == -(-iters // 3)


def test_fused_depth3_default_shape(gpu):
    """time_tile=3 with no tile or block: the depth-3 default shape, and the
    depth-2 default for a closing pair."""
    shape = (23, 61, 140)
    a = _field(shape, 21)
    for iters in (5, 6):
        with api.Grid(shape, "f32") as g:
            g.upload_pair(a)
            g.run(K.jacobi3d(), iters, variant="tuned", time_tile=3)
            got = g.download()
        want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), iters


def test_fused_depth3_tiles_segments_range(gpu):
    """Many w tiles, h segments of odd length down to 5 planes (each reads 3
    planes beyond it on both sides), and a declared range of 2."""
    shape = (70, 130, 520)
    a = _field(shape, 17)
    for tshape in SHAPES_T3:
        for per, iters in ((5, 6), (7, 9), (32, 8), (100, 4), (0, 7)):
            got, _ = _run3(a, iters, per, tshape)
            want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (tshape, per, iters)
    s = dataclasses.replace(K.jacobi3d(), radius=2)
    b = _field((21, 37, 90), 12)
    with api.Grid(b.shape, "f32") as g:
        g.upload_pair(b)
        g.run(s, 7, variant="tuned", time_tile=3)
        got = g.download()
    want, _ = oracle.run_3d(s, b, 7)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("every", [1, 2, 3, 4, 5])
def test_fused_depth3_residuals(gpu, every):
    """Launches stop at every residual step; the residual of a depth-3 launch
    is its third sweep against its second."""
    shape = (40, 36, 100)
    a = _field(shape, 8)
    iters = 20
    got_a, st = _run3(a, iters, 16, SHAPES_T3[every % 2], residual_every=every)
    want, want_res = oracle.run_3d(K.jacobi3d(), a, iters, residual_every=every)
    assert np.array_equal(got_a.view(np.uint32), want.view(np.uint32))
    got = np.array(st["residuals"])
    assert got.shape == want_res.shape == (iters // every,)
    np.testing.assert_allclose(got, want_res, rtol=1e-12, atol=0)


def test_fused_depth3_split_runs_graph_full_size(gpu):
    """run(4); run(5) and graph replays equal single sweeps; 512^3 x 6 in two
    launches."""
    shape = (30, 50, 200)
    a = _field(shape, 19)
    want9, _ = oracle.run_3d(K.jacobi3d(), a, 9)
    with api.Grid(shape, "f32") as g:
        g.upload_pair(a)
        g.run(K.jacobi3d(), 4, variant="tuned", time_tile=3)
        g.run(K.jacobi3d(), 5, variant="tuned", time_tile=3)
        assert np.array_equal(g.download().view(np.uint32), want9.view(np.uint32))
    want18, _ = oracle.run_3d(K.jacobi3d(), a, 18)
    with api.Grid(shape, "f32") as g:
        g.upload_pair(a)
        for rep in range(2):
            g.run(K.jacobi3d(), 9, variant="tuned", time_tile=3, use_graph=True)
            api.synchronize()
            w = want9 if rep == 0 else want18
            assert np.array_equal(g.download().view(np.uint32), w.view(np.uint32)), rep
    big = inputs.jacobi3d_field(512, seed=4)
    want, _ = oracle.run_3d(K.jacobi3d(), big, 6)
    for tshape in SHAPES_T3:
        got, st = _run3(big, 6, 0, tshape)
        assert st["launches"] == 2
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), tshape


def test_fused_depth_is_a_hint(gpu):
    """A 3D stencil the fused kernel does not run takes time_tile=2 as a hint
    and sweeps singly, with the same result (bf16 storage runs fused)."""
    a = _field((12, 13, 70), 15)
    halo = dataclasses.replace(K.jacobi3d(), boundary=K.BOUNDARY_HALO, halo_value=0.5)
    with api.Grid(a.shape, "f32") as g:
        g.upload_pair(a)
        st = g.run(halo, 5, variant="tuned", time_tile=2)
        got = g.download()
    want, _ = oracle.run_3d(halo, a, 5)
    assert st["launches"] == 5
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    w = oracle.encode16(a, oracle.STORE_BF16)
    with api.Grid(a.shape, "bf16") as g:
        g.upload(0, w)
        g.upload(1, w)
        g.run(K.jacobi3d(), 5, variant="tuned", time_tile=2)
        got = g.download(raw=True)
    want16, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, w, 5)
    assert np.array_equal(got, want16)


SHAPES_16 = [(4, 3, 2, 6, 1), (4, 2, 2, 8, 1), (4, 3, 1, 12, 1), (2, 3, 1, 12, 1), (4, 2, 2, 6, 1)]


def _run16(words, iters, store, tshape, per=0, offset=0.5, rounding="nearest", seed=0, **kw):
    vd, rw, nwd, nww, pf = tshape
    with api.Grid(words.shape, store, store_offset=offset, rounding=rounding, seed=seed) as g:
        g.upload(0, words)
        g.upload(1, words)
        st = g.run(K.jacobi3d(), iters, variant="tuned", block=_block(nwd, nww, pf), tile=(vd, rw, per),
                   time_tile=2, **kw)
        return g.download(raw=True), st


@pytest.mark.parametrize("tshape", SHAPES_16, ids=["vd{}_rw{}_d{}_w{}_pf{}".format(*t) for t in SHAPES_16])
@pytest.mark.parametrize("store", ["bf16", "f16"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
@pytest.mark.parametrize("grid", [(3, 3, 3), (17, 9, 70), (20, 70, 300), (9, 23, 1100)],
                         ids=["3x3x3", "17x9x70", "20x70x300", "9x23x1100"])
def test_fused_16bit_matches_oracle(gpu, tshape, store, rounding, grid):
    code = oracle.STORE_BF16 if store == "bf16" else oracle.STORE_F16
    rnd = oracle.ROUND_STOCHASTIC if rounding == "stochastic" else oracle.ROUND_NEAREST
    a = _field(grid, sum(grid) + 3)
    for offset in (0.0, 0.5):
        words = oracle.encode16(a, code, offset)
        for per in (3, 0):
            for iters in (2, 5):
                got, st = _run16(words, iters, store, tshape, per, offset, rounding, seed=41)
                want, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, iters, store_offset=offset,
                                            rounding=rnd, seed=41)
                assert np.array_equal(got, want), (tshape, store, rounding, grid, offset, per, iters)
                assert st["launches"] == (iters + 1) // 2


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_fused_16bit_split_runs_residuals_and_graph(gpu, store):
    """Stochastic rounding continues its sweep index across runs and graph
    replays; residuals of fused launches match the oracle's."""
    code = oracle.STORE_BF16 if store == "bf16" else oracle.STORE_F16
    shape = (30, 50, 200)
    a = _field(shape, 16)
    words = oracle.encode16(a, code, 0.5)
    want7, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, 7, store_offset=0.5,
                                 rounding=oracle.ROUND_STOCHASTIC, seed=5)
    want14, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, 14, store_offset=0.5,
                                  rounding=oracle.ROUND_STOCHASTIC, seed=5)
    with api.Grid(shape, store, store_offset=0.5, rounding="stochastic", seed=5) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(K.jacobi3d(), 3, variant="tuned", time_tile=2)
        g.run(K.jacobi3d(), 4, variant="tuned", time_tile=2)
        assert np.array_equal(g.download(raw=True), want7)
        g.run(K.jacobi3d(), 7, variant="tuned", time_tile=2, use_graph=True)
        assert np.array_equal(g.download(raw=True), want14)
    _, want_res = oracle.run_3d_r16(K.jacobi3d(), code, words, 20, store_offset=0.5,
                                    rounding=oracle.ROUND_STOCHASTIC, seed=5, residual_every=4)
    with api.Grid(shape, store, store_offset=0.5, rounding="stochastic", seed=5) as g:
        g.upload(0, words)
        g.upload(1, words)
        st = g.run(K.jacobi3d(), 20, variant="tuned", time_tile=2, residual_every=4)
    np.testing.assert_allclose(np.array(st["residuals"]), want_res, rtol=1e-12, atol=0)


@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_fused_host_staged(gpu, store):
    """Host-staged slabs run fused pairs too; stochastic rounding uses the
    slab's global plane origin."""
    shape = (40, 24, 130)
    a = _field(shape, 17)
    if store == "f32":
        want, _ = oracle.run_3d(K.jacobi3d(), a, 10)
        got, _ = api.run_host_tiled(K.jacobi3d(), a, 10, 9, 4, time_tile=2)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    else:
        words = oracle.encode16(a, oracle.STORE_BF16, 0.5)
        want, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, words, 10, store_offset=0.5,
                                    rounding=oracle.ROUND_STOCHASTIC, seed=77)
        got, _ = api.run_host_tiled(K.jacobi3d(), words, 10, 9, 4, storage="bf16", store_offset=0.5,
                                    rounding="stochastic", seed=77, time_tile=2)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_automatic_storage_offset(gpu, store):
    """store_offset "auto" (NaN in the C ABI): an INPUT upload sets the offset
    to the mean of the uploaded values (summed in double in order, rounded
    once), and the run equals one with that offset given explicitly; raw
    storage words cannot set it."""
    a = _field((20, 30, 70), 31)
    mean = np.float32(np.cumsum(a.ravel().astype(np.float64))[-1] / a.size)
    with api.Grid(a.shape, store, store_offset="auto", rounding="stochastic", seed=3) as g:
        g.upload_pair(a)
        assert np.float32(g.info()["store_offset"]) == mean
        g.run(K.jacobi3d(), 7, variant="tuned", time_tile=2)
        got = g.download(raw=True)
        with pytest.raises(PSkelError):
            g.upload(0, np.zeros(a.shape, np.uint16))
    with api.Grid(a.shape, store, store_offset=float(mean), rounding="stochastic", seed=3) as g:
        g.upload_pair(a)
        g.run(K.jacobi3d(), 7, variant="tuned", time_tile=2)
        want = g.download(raw=True)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tshape,depth,store", [((4, 2, 2, 6, 1), 3, "f32"), ((4, 3, 2, 6, 1), 2, "f32"),
                                                ((4, 3, 2, 6, 1), 2, "bf16")])
def test_fused_even_split_over_tiles(gpu, tshape, depth, store):
    """Default planes per block on a grid whose tiles x planes split evenly over
    one block per CU (at 256 CUs a block's range of 9..18 tile-planes often
    ends in one tile and goes on in the next): bit-exact, with residuals at
    launch ends, against the oracle."""
    shape = (120, 300, 70)
    a = _field(shape, 23)
    vd, rw, nwd, nww, pf = tshape
    kw = dict(variant="tuned", block=_block(nwd, nww, pf), tile=(vd, rw, 0), time_tile=depth, residual_every=3)
    if store == "f32":
        with api.Grid(shape, "f32") as g:
            g.upload_pair(a)
            st = g.run(K.jacobi3d(), 9, **kw)
            got = g.download()
        want, want_res = oracle.run_3d(K.jacobi3d(), a, 9, residual_every=3)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    else:
        words = oracle.encode16(a, oracle.STORE_BF16, 0.5)
        got, st = _run16(words, 9, store, tshape, 0, 0.5, "stochastic", seed=5, residual_every=3)
        want, want_res = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, words, 9, store_offset=0.5,
                                           rounding=oracle.ROUND_STOCHASTIC, seed=5, residual_every=3)
        assert np.array_equal(got, want)
    np.testing.assert_allclose(np.array(st["residuals"]), want_res, rtol=1e-12, atol=0)
