"""GPU parity of the 3D sweeps against the CPU oracle.

f32 storage: bit-exact.  bf16/f16 storage: bit-exact against the oracle's
reduced-precision runner (same f32 arithmetic, same round-to-nearest-even, or
the same stochastic rounding bits for a given seed and sweep index).
FP64 residuals: the device sums in a different order, so they match to 1e-12
relative.
"""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (3, 3, 3), (4, 5, 6), (17, 9, 70), (33, 40, 65), (20, 70, 300), (64, 64, 64)]


def _field(shape, seed):
    H, W, D = shape
    a = inputs.uniform(seed, 0, H * W * D).reshape(shape)
    a[0] = a[-1] = 1.0
    a[:, 0] = a[:, -1] = 1.0
    a[:, :, 0] = a[:, :, -1] = 1.0
    return a


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("variant", ["plain", "tuned"])
def test_jacobi3d_f32(gpu, shape, variant):
    a = _field(shape, sum(shape))
    for iters in (0, 1, 4):
        with api.Grid(shape, "f32") as g:
            g.upload_pair(a)
            g.run(K.jacobi3d(), iters, variant=variant)
            got = g.download()
        want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (shape, variant, iters)


@pytest.mark.parametrize("tile", [(1, 1, 8), (1, 4, 16), (1, 8, 64), (2, 2, 32), (2, 8, 7), (4, 4, 64), (4, 1, 100)])
def test_jacobi3d_march_shapes(gpu, tile):
    shape = (70, 45, 300)
    a = _field(shape, 3)
    with api.Grid(shape, "f32") as g:
        g.upload_pair(a)
        g.run(K.jacobi3d(), 3, variant="tuned", tile=tile)
        got = g.download()
    want, _ = oracle.run_3d(K.jacobi3d(), a, 3)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), tile


@pytest.mark.parametrize("store", ["bf16", "f16"])
@pytest.mark.parametrize("variant", ["plain", "tuned"])
@pytest.mark.parametrize("offset", [0.0, 0.5])
def test_jacobi3d_reduced_storage(gpu, store, variant, offset):
    shape = (24, 33, 130)
    a = _field(shape, 4)
    code = oracle.STORE_BF16 if store == "bf16" else oracle.STORE_F16
    words = oracle.encode16(a, code, offset)
    with api.Grid(shape, store, store_offset=offset) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(K.jacobi3d(), 6, variant=variant)
        got = g.download(raw=True)
    want, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, 6, store_offset=offset)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tile", [(2, 4, 32), (4, 2, 16), (8, 2, 64), (8, 1, 8)])
def test_jacobi3d_reduced_march_shapes(gpu, tile):
    shape = (30, 21, 600)
    a = _field(shape, 5)
    words = oracle.encode16(a, oracle.STORE_BF16, 0.5)
    with api.Grid(shape, "bf16", store_offset=0.5) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(K.jacobi3d(), 4, variant="tuned", tile=tile)
        got = g.download(raw=True)
    want, _ = oracle.run_3d_r16(K.jacobi3d(), oracle.STORE_BF16, words, 4, store_offset=0.5)
    assert np.array_equal(got, want), tile


def _sr_case(shape, store, offset, seed):
    a = _field(shape, 4 + seed)
    code = oracle.STORE_BF16 if store == "bf16" else oracle.STORE_F16
    return a, code, oracle.encode16(a, code, offset)


@pytest.mark.parametrize("store", ["bf16", "f16"])
@pytest.mark.parametrize("variant", ["plain", "tuned"])
@pytest.mark.parametrize("offset", [0.0, 0.5])
def test_jacobi3d_stochastic_rounding(gpu, store, variant, offset):
    shape = (24, 33, 130)
    _, code, words = _sr_case(shape, store, offset, 1)
    want, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, 6, store_offset=offset,
                                rounding=oracle.ROUND_STOCHASTIC, seed=1234)
    with api.Grid(shape, store, store_offset=offset, rounding="stochastic", seed=1234) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(K.jacobi3d(), 6, variant=variant)
        got = g.download(raw=True)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tile", [(2, 4, 32), (4, 1, 32), (8, 2, 64), (4, 1, 100)])
@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_jacobi3d_stochastic_march_shapes(gpu, tile, store):
    shape = (30, 21, 600)
    _, code, words = _sr_case(shape, store, 0.5, 2)
    want, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, 4, store_offset=0.5,
                                rounding=oracle.ROUND_STOCHASTIC, seed=99)
    with api.Grid(shape, store, store_offset=0.5, rounding="stochastic", seed=99) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(K.jacobi3d(), 4, variant="tuned", tile=tile)
        got = g.download(raw=True)
    assert np.array_equal(got, want), tile


@pytest.mark.parametrize("use_graph", [False, True])
def test_stochastic_rounding_sweep_index(gpu, use_graph):
    """The random bits follow the sweep index since the last INPUT upload: split
    runs equal one run, graph replays draw fresh bits, re-uploading restarts the
    stream, and the seed changes it."""
    shape = (20, 26, 140)
    _, code, words = _sr_case(shape, "bf16", 0.5, 3)
    s = K.jacobi3d()
    want7, _ = oracle.run_3d_r16(s, code, words, 7, store_offset=0.5, rounding=oracle.ROUND_STOCHASTIC, seed=5)
    want4, _ = oracle.run_3d_r16(s, code, words, 4, store_offset=0.5, rounding=oracle.ROUND_STOCHASTIC, seed=5)
    want8, _ = oracle.run_3d_r16(s, code, words, 8, store_offset=0.5, rounding=oracle.ROUND_STOCHASTIC, seed=5)
    with api.Grid(shape, "bf16", store_offset=0.5, rounding="stochastic", seed=5) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(s, 3, variant="tuned", use_graph=use_graph)
        g.run(s, 4, variant="tuned", use_graph=use_graph)
        assert np.array_equal(g.download(raw=True), want7)
        g.upload(0, words)
        g.upload(1, words)
        g.run(s, 4, variant="tuned", use_graph=use_graph)
        assert np.array_equal(g.download(raw=True), want4)
        g.run(s, 4, variant="tuned", use_graph=use_graph)   # same launch shape, sweeps 4..7
        assert np.array_equal(g.download(raw=True), want8)
    with api.Grid(shape, "bf16", store_offset=0.5, rounding="stochastic", seed=6) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(s, 7, variant="tuned", use_graph=use_graph)
        assert not np.array_equal(g.download(raw=True), want7)


def test_stochastic_rounding_residuals(gpu):
    shape = (40, 36, 100)
    _, code, w = _sr_case(shape, "bf16", 0.5, 4)
    s = K.jacobi3d()
    want, want_res = oracle.run_3d_r16(s, code, w, 20, store_offset=0.5, residual_every=5,
                                       rounding=oracle.ROUND_STOCHASTIC, seed=8)
    with api.Grid(shape, "bf16", store_offset=0.5, rounding="stochastic", seed=8) as g:
        g.upload(0, w)
        g.upload(1, w)
        st = g.run(s, 20, variant="tuned", residual_every=5)
        got = g.download(raw=True)
    assert np.array_equal(got, want)
    np.testing.assert_allclose(np.array(st["residuals"]), want_res, rtol=1e-12, atol=0)


def test_stochastic_rounding_accuracy(gpu):
    """96^3 x 400 sweeps: stochastic bf16 stays within 1e-3 relative L2 of the
    FP64 run, nearest bf16 does not."""
    u = inputs.jacobi3d_field(96, seed=9)
    want, _ = oracle.run_3d(K.jacobi3d(), u.astype(np.float64), 400)
    rel = {}
    for mode in ("nearest", "stochastic"):
        w = oracle.encode16(u, oracle.STORE_BF16, 0.5)
        with api.Grid(u.shape, "bf16", store_offset=0.5, rounding=mode, seed=21) as g:
            g.upload(0, w)
            g.upload(1, w)
            g.run(K.jacobi3d(), 400, variant="tuned")
            got = oracle.decode16(g.download(raw=True), oracle.STORE_BF16, 0.5).astype(np.float64)
        rel[mode] = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel["stochastic"] < 1e-3, rel
    assert rel["nearest"] > 2.5 * rel["stochastic"], rel


@pytest.mark.parametrize("variant", ["plain", "tuned"])
@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_residuals(gpu, variant, store):
    shape = (40, 36, 100)
    a = _field(shape, 6)
    s = K.jacobi3d()
    with api.Grid(shape, store) as g:
        if store == "f32":
            g.upload_pair(a)
            want_res = oracle.run_3d(s, a, 20, residual_every=5)[1]
        else:
            w = oracle.encode16(a, oracle.STORE_BF16)
            g.upload(0, w)
            g.upload(1, w)
            want_res = oracle.run_3d_r16(s, oracle.STORE_BF16, w, 20, residual_every=5)[1]
        st = g.run(s, 20, variant=variant, residual_every=5)
    got = np.array(st["residuals"])
    assert got.shape == want_res.shape == (4,)
    np.testing.assert_allclose(got, want_res, rtol=1e-12, atol=0)


def test_halo_boundary_generic(gpu):
    """Mask3D halo semantics: every cell updated, outside reads the halo value."""
    s = K.Stencil(K.WEIGHTED_SUM, 3, [(0, 0, 0)] + list(K.J7_OFFSETS), weights=[0.4] + [0.1] * 6,
                  boundary=K.BOUNDARY_HALO, halo_value=2.0)
    shape = (9, 10, 11)
    a = _field(shape, 7)
    with api.Grid(shape, "f32") as g:
        g.upload_pair(a)
        g.run(s, 5)
        got = g.download()
    want, _ = oracle.run_3d(s, a, 5)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_jacobi3d_full_size_tuned(gpu):
    """512^3 x 6 sweeps, tuned and plain, bit-exact (the oracle needs ~10 s)."""
    a = inputs.jacobi3d_field(512, seed=4)
    want, _ = oracle.run_3d(K.jacobi3d(), a, 6)
    for kw in ({"variant": "tuned"}, {"variant": "plain"}):
        with api.Grid(a.shape, "f32") as g:
            g.upload_pair(a)
            g.run(K.jacobi3d(), 6, **kw)
            got = g.download()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), kw
