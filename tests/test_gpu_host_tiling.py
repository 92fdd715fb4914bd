"""Host-staged sweeps (runIterativeTilingGPU's out-of-core semantics) against the
oracle and the resident runner, bit-exact: slabs of rows/planes plus a
range*inner halo visit the device, the rest of the grid stays in host memory."""
import numpy as np
import pytest

import oracle
from pskel_amd import api, inputs
from pskel_amd import kernels as K
from pskel_amd._lib import PSkelError

pytestmark = pytest.mark.gpu


def _radius2():
    """A weighted mask with range 2 along h (plain kernel path)."""
    offs = [(-2, 0), (-1, 0), (0, -1), (0, 0), (0, 1), (1, 0), (2, 0)]
    return K.Stencil(K.WEIGHTED_SUM, 2, offs, weights=[0.1, 0.15, 0.15, 0.2, 0.15, 0.15, 0.1])


CASES_2D = [
    ("jacobi2d", K.jacobi2d, lambda: inputs.jacobi2d_field(97, 131, seed=1)),
    ("diffusion9", K.diffusion9, lambda: inputs.blobs_field(128, seed=3, nblobs=4, sigma=10.0)[:101, :127].copy()),
    ("life", K.life, lambda: inputs.life_field(90, seed=4)),
    ("radius2", _radius2, lambda: inputs.jacobi2d_field(60, 77, seed=5)),
]


@pytest.mark.parametrize("name,stencil,field", CASES_2D, ids=[c[0] for c in CASES_2D])
@pytest.mark.parametrize("tile,inner", [(1, 1), (7, 2), (16, 3), (33, 5), (1000, 4)])
def test_host_tiled_2d(gpu, name, stencil, field, tile, inner):
    a = field()
    b = a.copy()
    b[0, :] = b[0, :][::-1]        # a distinct output ring: it must survive
    for iters in (0, 1, 7, 12):
        want = oracle.run_2d(stencil(), a, iters, out=b)
        for variant, tt in (("plain", 1), ("tuned", 1), ("tuned", 2)):
            got, st = api.run_host_tiled(stencil(), a, iters, tile, inner, out=b, variant=variant, time_tile=tt)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (name, tile, inner, iters, variant, tt)
            assert st["sweeps"] == iters


@pytest.mark.parametrize("tile,inner", [(1, 1), (5, 2), (16, 3), (100, 6)])
@pytest.mark.parametrize("variant", ["plain", "tuned"])
def test_host_tiled_3d(gpu, tile, inner, variant):
    a = inputs.jacobi3d_field(33, seed=6)
    for iters in (1, 8, 13):
        want, _ = oracle.run_3d(K.jacobi3d(), a, iters)
        got, _ = api.run_host_tiled(K.jacobi3d(), a, iters, tile, inner, variant=variant)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (tile, inner, iters, variant)


def test_host_tiled_3d_halo_boundary(gpu):
    """3D HALO rule (every cell updated, outside reads the halo value)."""
    s = K.Stencil(K.WEIGHTED_SUM, 3, [(0, 0, 0)] + list(K.J7_OFFSETS), weights=[0.4] + [0.1] * 6,
                  boundary=K.BOUNDARY_HALO, halo_value=2.0)
    a = inputs.jacobi3d_field(21, seed=7)
    want, _ = oracle.run_3d(s, a, 9)
    got, _ = api.run_host_tiled(s, a, 9, 4, 3, variant="plain")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_host_tiled_16bit_stochastic(gpu, store):
    """Stochastic rounding draws from the global sweep index and plane: host
    tiling gives the resident run's words."""
    shape = (40, 24, 130)
    a = inputs.uniform(8, 0, 40 * 24 * 130).reshape(shape)
    a[0] = a[-1] = 1.0
    a[:, 0] = a[:, -1] = 1.0
    a[:, :, 0] = a[:, :, -1] = 1.0
    code = oracle.STORE_BF16 if store == "bf16" else oracle.STORE_F16
    words = oracle.encode16(a, code, 0.5)
    want, _ = oracle.run_3d_r16(K.jacobi3d(), code, words, 10, store_offset=0.5,
                                rounding=oracle.ROUND_STOCHASTIC, seed=77)
    got, _ = api.run_host_tiled(K.jacobi3d(), words, 10, 9, 4, storage=store, store_offset=0.5,
                                rounding="stochastic", seed=77)
    assert np.array_equal(got, want)
    with api.Grid(shape, store, store_offset=0.5, rounding="stochastic", seed=77) as g:
        g.upload(0, words)
        g.upload(1, words)
        g.run(K.jacobi3d(), 10, variant="tuned")
        assert np.array_equal(g.download(raw=True), got)


def test_host_tiled_rejects(gpu):
    a = inputs.jacobi2d_field(20, 30, seed=9)
    with pytest.raises(PSkelError):
        api.run_host_tiled(K.jacobi2d(), a, 3, 0, 1)          # tile 0
    with pytest.raises(PSkelError):
        api.run_host_tiled(K.jacobi2d(), a.astype(np.uint8), 3, 4, 1, storage="f32")
    t, wx, wy = inputs.cloudsim_fields(20, 30, seed=9)
    with pytest.raises(PSkelError):
        api.run_host_tiled(K.cloudsim(), t, 3, 4, 1)


def _host_search(a, stencil, iterations, budget=0):
    import ctypes as C

    from pskel_amd import _lib as L

    gd = L.GridDesc()
    gd.ndim = a.ndim
    gd.storage = L.DTYPE_F32 if a.dtype == np.float32 else L.DTYPE_U8
    gd.height, gd.width = a.shape[0], a.shape[1]
    gd.depth = a.shape[2] if a.ndim == 3 else 1
    d = stencil.desc()
    tile, inner, ms = C.c_size_t(), C.c_size_t(), C.c_double()
    L.check(L.lib().pskel_tuner_host_search(C.byref(gd), np.ascontiguousarray(a).ctypes.data, gd.storage,
                                            C.byref(d), iterations, budget, C.byref(tile), C.byref(inner),
                                            C.byref(ms)))
    return tile.value, inner.value


def test_out_of_core_search_picks_a_runnable_launch(gpu):
    """pskel_tuner_host_search (runIterativeAutoGPU's tiled branch): the slab
    height and sweeps per pass it returns are candidates, fit the budget, and
    give the oracle's result."""
    a = inputs.jacobi3d_field(96, seed=33)
    tile, inner = _host_search(a, K.jacobi3d(), 20)
    assert tile in (16, 32, 64) and inner in (2, 4, 8, 16)
    got, _ = api.run_host_tiled(K.jacobi3d(), a, 20, tile, inner)
    want, _ = oracle.run_3d(K.jacobi3d(), a, 20)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # a device budget too small for any slab is refused
    from pskel_amd._lib import PSkelError

    with pytest.raises(PSkelError):
        _host_search(a, K.jacobi3d(), 20, budget=1 << 16)
    b = inputs.life_field(600, seed=34)
    tile, inner = _host_search(b, K.life(), 12)
    got, _ = api.run_host_tiled(K.life(), b, 12, tile, inner)
    assert np.array_equal(got, oracle.run_2d(K.life(), b, 12))
