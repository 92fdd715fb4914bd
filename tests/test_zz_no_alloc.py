"""The allocation-free entry point (pskel_run_iterative_noalloc) through the C
ABI against pskel_run_iterative on a grid with the same launch configuration:
bit-identical results and residuals, on two calls in a row per case, the
second on a smaller grid (the static scratch array keeps the first call's
contents, which must not leak into the second).
"""
import ctypes as C

import numpy as np
import pytest

from pskel_amd import _lib as L
from pskel_amd import api, inputs
from pskel_amd import kernels as K

pytestmark = pytest.mark.gpu

_ELEM = {"u8": 1, "f32": 4, "bf16": 2, "f16": 2}
_NP = {"u8": np.uint8, "f32": np.float32, "bf16": np.uint16, "f16": np.uint16}


def _grid_run(host, storage, stencil, iters, cfg_kw, **gkw):
    """The existing entry point: a grid, both buffers uploaded from `host`."""
    with api.Grid(host.shape, storage, **gkw) as g:
        g.upload_pair(host)
        st = g.run(stencil, iters, variant="tuned", **cfg_kw)
        got = g.download(raw=storage in ("bf16", "f16"))
    return got, np.array(st["residuals"])


def _noalloc_run(host, storage, stencil, iters, cfg_kw, store_offset=0.0, rounding="nearest", seed=0):
    """The allocation-free entry point on two device buffers in the grid layout."""
    lib = L.lib()
    shape = host.shape
    ndim = len(shape)
    elem = _ELEM[storage]
    row = shape[-1]
    nrows = int(np.prod(shape[:-1]))
    pitch = -(-row * elem // 1024) * 1024 // elem
    buf = np.zeros((nrows, pitch), _NP[storage])
    buf[:, :row] = host.reshape(nrows, row)
    nbytes = buf.nbytes
    gd = L.GridDesc()
    gd.ndim = ndim
    gd.storage = api._STORAGE[storage]
    gd.height, gd.width = shape[0], shape[1]
    gd.depth = shape[2] if ndim == 3 else 1
    gd.store_offset = store_offset
    gd.device = 0
    gd.rounding = {"nearest": L.ROUND_NEAREST, "stochastic": L.ROUND_STOCHASTIC}[rounding]
    gd.seed = seed
    cfg = api.make_cfg("tuned", cfg_kw.get("block", (0, 0)), cfg_kw.get("tile", (0, 0, 0)),
                       cfg_kw.get("time_tile", 1), cfg_kw.get("residual_every", 0), False, 0)
    every = cfg_kw.get("residual_every", 0)
    nres = iters // every if every else 0
    res = (C.c_double * max(nres, 1))()
    st = L.RunStats()
    st.residuals = C.cast(res, C.POINTER(C.c_double))
    st.residual_capacity = nres
    din, dout = C.c_void_p(), C.c_void_p()
    L.check(lib.pskel_device_alloc(nbytes, C.byref(din)))
    L.check(lib.pskel_device_alloc(nbytes, C.byref(dout)))
    try:
        L.check(lib.pskel_memcpy(din, buf.ctypes.data, nbytes, 1))
        L.check(lib.pskel_memcpy(dout, buf.ctypes.data, nbytes, 1))
        d = stencil.desc()
        L.check(lib.pskel_run_iterative_noalloc(C.byref(gd), din, dout, pitch, C.byref(d), iters, C.byref(cfg),
                                                C.byref(st)))
        L.check(lib.pskel_device_synchronize())
        outb = np.empty_like(buf)
        L.check(lib.pskel_memcpy(outb.ctypes.data, dout, nbytes, 0))
    finally:
        lib.pskel_device_free(din)
        lib.pskel_device_free(dout)
    assert st.residual_count == nres
    return outb[:, :row].reshape(shape), np.array([res[i] for i in range(nres)])


def _life(n, seed):
    return inputs.life_field(n, seed=seed)


CASES = {
    # name: (storage, stencil, iterations, launch config, grid kwargs, [big input, small input])
    "jacobi3d_fused_pair": ("f32", K.jacobi3d, 11,
                            dict(tile=(4, 3, 0), block=(768, 2), time_tile=2, residual_every=3), {},
                            [lambda: inputs.jacobi3d_field(70, seed=31), lambda: inputs.jacobi3d_field(33, seed=32)]),
    "jacobi3d_one_sweep": ("f32", K.jacobi3d, 7, dict(residual_every=2), {},
                           [lambda: inputs.jacobi3d_field(64, seed=33), lambda: inputs.jacobi3d_field(20, seed=34)]),
    "jacobi2d_depth4": ("f32", K.jacobi2d, 13, dict(time_tile=4, tile=(0, 16, 0), residual_every=5), {},
                        [lambda: inputs.jacobi2d_field(700, 900, seed=35),
                         lambda: inputs.jacobi2d_field(130, 210, seed=36)]),
    "diffusion9_depth3": ("f32", K.diffusion9, 10, dict(time_tile=3, tile=(0, 16, 0), residual_every=4), {},
                          [lambda: inputs.blobs_field(512, seed=37, nblobs=4, sigma=20.0),
                           lambda: inputs.blobs_field(150, seed=38, nblobs=2, sigma=10.0)]),
    "life_bits": ("u8", K.life, 18, dict(time_tile=8, tile=(0, 32, 0), residual_every=6), {},
                  [lambda: _life(900, 39), lambda: _life(200, 40)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_noalloc_matches_grid_run(gpu, name):
    storage, stencil, iters, cfg_kw, gkw, makers = CASES[name]
    for make in makers:  # the second call runs a smaller grid
        host = make()
        want, want_res = _grid_run(host, storage, stencil(), iters, cfg_kw, **gkw)
        got, res = _noalloc_run(host, storage, stencil(), iters, cfg_kw, **gkw)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (name, host.shape)
        np.testing.assert_array_equal(res, want_res)


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_noalloc_16bit_stochastic(gpu, store):
    """16-bit words of u - 0.5 with stochastic rounding: the words and the
    residuals equal the grid's (both count sweeps from 0)."""
    cfg_kw = dict(tile=(4, 3, 0), block=(768, 2), time_tile=2, residual_every=2)
    gkw = dict(store_offset=0.5, rounding="stochastic", seed=9)
    for n in (72, 30):
        a = inputs.jacobi3d_field(n, seed=41 + n)
        with api.Grid(a.shape, store, **gkw) as g:
            g.upload_pair(a)
            words = g.download(raw=True)
        want, want_res = _grid_run(words, store, K.jacobi3d(), 8, cfg_kw, **gkw)
        got, res = _noalloc_run(words, store, K.jacobi3d(), 8, cfg_kw, **gkw)
        assert np.array_equal(got, want), (store, n)
        np.testing.assert_array_equal(res, want_res)


def test_noalloc_zero_iterations_and_limits(gpu):
    a = inputs.jacobi3d_field(24, seed=50)
    got, _ = _noalloc_run(a, "f32", K.jacobi3d(), 0, {})
    assert np.array_equal(got, a)
    # more residuals than the scratch array holds
    with pytest.raises(L.PSkelError):
        _noalloc_run(a, "f32", K.jacobi3d(), 4097, dict(residual_every=1))
