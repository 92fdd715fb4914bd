"""Host-side mirror of PSkel's stencil API over the C ABI.

`Grid` is the device-resident ping-pong pair (pskel_grid_t).  `Array2D`,
`Array3D`, `Mask2D`, `Mask3D`, `Stencil2D` and `Stencil3D` mirror the reference
classes (include/PSkelArray.h, include/PSkelMask.h, include/PSkelStencil.h) with
the same method names, argument meaning and index conventions:

* Array2D(width, height) holds height x width cells, (h, w) -> h*width + w
  (include/PSkelArray.hpp:668); Array3D(width, height, depth) holds
  (h*width + w)*depth + d (include/PSkelArray.hpp:642).
* runIterativeGPU(N, bx, by) leaves `input` untouched and the result in
  `output` (include/PSkelStencil.hpp:1047-1097).
* Only the GPU methods exist here.  runIterativeCPU belongs to the C++ drop-in
  headers (include/PSkel.h); this module never computes a sweep on the host.

The stencil kernel is given by the `args` object (a kernels.Stencil, or a
StencilArgs naming one), as the Args type selects the stencilKernel overload in
the C++ API.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .kernels import LIFE, SCALED_SUM, Stencil

_NP_TO_DTYPE = {
    np.dtype(np.uint8): L.DTYPE_U8,
    np.dtype(np.float32): L.DTYPE_F32,
    np.dtype(np.float64): L.DTYPE_F64,
}

_STORAGE = {"u8": L.DTYPE_U8, "f32": L.DTYPE_F32, "bf16": L.DTYPE_BF16, "f16": L.DTYPE_F16}


def _host_dtype(arr: np.ndarray, storage: int) -> int:
    if arr.dtype == np.uint16 and storage in (L.DTYPE_BF16, L.DTYPE_F16):
        return storage
    try:
        return _NP_TO_DTYPE[arr.dtype]
    except KeyError:
        raise TypeError(f"unsupported host dtype {arr.dtype}") from None


class Grid:
    """A device-resident grid pair (input/output ping-pong) on one GPU.
    store_offset: what 16-bit storage holds u - offset against; "auto": the
    mean of each INPUT upload (F32/F64 data)."""

    def __init__(self, shape: Sequence[int], storage: str = "f32", store_offset=0.0,
                 device: int = 0, rounding: str = "nearest", seed: int = 0):
        if len(shape) not in (2, 3):
            raise ValueError("shape is (height, width) or (height, width, depth)")
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(shape)
        self.storage = _STORAGE[storage]
        self.storage_name = storage
        gd = L.GridDesc()
        gd.ndim = self.ndim
        gd.storage = self.storage
        gd.height = self.shape[0]
        gd.width = self.shape[1]
        gd.depth = self.shape[2] if self.ndim == 3 else 1
        gd.store_offset = float("nan") if store_offset == "auto" else float(store_offset)
        gd.device = device
        gd.rounding = {"nearest": L.ROUND_NEAREST, "stochastic": L.ROUND_STOCHASTIC}[rounding]
        gd.seed = int(seed) & 0xFFFFFFFF
        h = C.c_void_p()
        L.check(L.lib().pskel_grid_create(C.byref(gd), C.byref(h)))
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            L.lib().pskel_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data -------------------------------------------------------------
    def _check_host(self, arr: np.ndarray) -> np.ndarray:
        arr = np.ascontiguousarray(arr)
        if tuple(arr.shape) != self.shape:
            raise ValueError(f"host array shape {arr.shape} != grid shape {self.shape}")
        return arr

    def upload(self, which: int, arr: np.ndarray):
        arr = self._check_host(arr)
        L.check(L.lib().pskel_grid_upload(self._h, which, arr.ctypes.data, _host_dtype(arr, self.storage)))

    def upload_pair(self, inp: np.ndarray, out: Optional[np.ndarray] = None):
        self.upload(L.BUFFER_INPUT, inp)
        self.upload(L.BUFFER_OUTPUT, inp if out is None else out)

    def _out_array(self, dtype) -> np.ndarray:
        if dtype is None:
            dtype = {L.DTYPE_U8: np.uint8, L.DTYPE_F32: np.float32}.get(self.storage, np.float32)
        return np.empty(self.shape, dtype)

    def download(self, dtype=None, raw: bool = False) -> np.ndarray:
        """Current state.  raw=True returns 16-bit storage words as uint16."""
        a = np.empty(self.shape, np.uint16) if raw else self._out_array(dtype)
        L.check(L.lib().pskel_grid_download(self._h, a.ctypes.data, _host_dtype(a, self.storage)))
        return a

    def download_previous(self, dtype=None, raw: bool = False) -> np.ndarray:
        a = np.empty(self.shape, np.uint16) if raw else self._out_array(dtype)
        L.check(L.lib().pskel_grid_download_previous(self._h, a.ctypes.data, _host_dtype(a, self.storage)))
        return a

    def info(self) -> Dict:
        gi = L.GridInfo()
        L.check(L.lib().pskel_grid_info(self._h, C.byref(gi)))
        return {f: getattr(gi, f) for f, _ in L.GridInfo._fields_}

    # -- sweeps -------------------------------------------------------------
    def run(self, stencil: Stencil, iterations: int, variant: str = "plain", block: Tuple[int, int] = (0, 0),
            tile: Tuple[int, int, int] = (0, 0, 0), time_tile: int = 1, residual_every: int = 0,
            use_graph: bool = False, stream: int = 0) -> Dict:
        cfg = make_cfg(variant, block, tile, time_tile, residual_every, use_graph, stream)
        st = L.RunStats()
        nres = iterations // residual_every if residual_every > 0 else 0
        res = (C.c_double * max(nres, 1))()
        st.residuals = res
        st.residual_capacity = nres
        d = stencil.desc()
        L.check(L.lib().pskel_run_iterative(self._h, C.byref(d), int(iterations), C.byref(cfg), C.byref(st)))
        return {
            "elapsed_ms": st.elapsed_ms,
            "launches": st.launches,
            "sweeps": st.sweeps,
            "residuals": [res[i] for i in range(min(st.residual_count, nres))],
        }

    def residual(self) -> float:
        v = C.c_double()
        L.check(L.lib().pskel_grid_residual(self._h, C.byref(v)))
        return v.value

    def set_field(self, slot: int, field: Optional["Field"]):
        """Attach a field argument (not owned; keep it alive while runs use it)."""
        L.check(L.lib().pskel_grid_set_field(self._h, int(slot), field._h if field is not None else None))


class Field:
    """A device-resident Args / Args2D / Args3D field (include/PSkelArgs.h:31-132).

    `host` is the Args host array in the reference's layout, x fastest: a numpy
    array of shape (width,), (height, width) or (depth, height, width), so
    host[y, x] is Args2D(x, y) (include/PSkelArgs.hpp:116-123, 155-160)."""

    def __init__(self, host: np.ndarray, device: int = 0):
        host = np.ascontiguousarray(host, np.float32)
        fd = L.FieldDesc()
        fd.ndim = host.ndim
        fd.dtype = L.DTYPE_F32
        fd.width = host.shape[-1]
        fd.height = host.shape[-2] if host.ndim >= 2 else 1
        fd.depth = host.shape[0] if host.ndim == 3 else 1
        fd.device = device
        h = C.c_void_p()
        L.check(L.lib().pskel_field_create(C.byref(fd), C.byref(h)))
        self._h = h
        self.shape = host.shape
        self.upload(host)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, np.float32)
        if host.shape != self.shape:
            raise ValueError(f"field shape {host.shape} != {self.shape}")
        L.check(L.lib().pskel_field_upload(self._h, host.ctypes.data))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, np.float32)
        L.check(L.lib().pskel_field_download(self._h, out.ctypes.data))
        return out

    def close(self):
        if getattr(self, "_h", None):
            L.lib().pskel_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many 2D grids swept together by one persistent kernel per launch
    (pskel_batch_t).  grids: list of Grid; stencils: one per grid; rows: rows
    per lane per grid (1, 2, 4 or 8) at depth 1, tile rows at depth > 1 (f32:
    16 or 32; life: 32); depths: sweeps fused per launch per grid (odd: 1 or 3
    for f32, 1, 3, 5 or 7 for life); march: per grid, 0 or, for an f32 grid at
    depth 3, the rows per segment of the register march that replaces its LDS
    tiles (tile (4, 0, rows))."""

    def __init__(self, grids, stencils, rows=None, depths=None, march=None):
        n = len(grids)
        self.grids = list(grids)
        hs = (C.c_void_p * n)(*[g._h for g in grids])
        ds = (L.StencilDesc * n)(*[s.desc() for s in stencils])
        cs = (L.LaunchCfg * n)()
        for i in range(n):
            cs[i].variant = L.VARIANT_TUNED
            k = int(depths[i]) if depths is not None else 1
            cs[i].time_tile = k
            cs[i].tile_y = int(rows[i]) if rows is not None and rows[i] else (4 if k == 1 else 0)
            if march is not None and march[i]:
                cs[i].tile_x, cs[i].tile_y, cs[i].tile_z = 4, 0, int(march[i])
        h = C.c_void_p()
        L.check(L.lib().pskel_batch_create(hs, ds, cs, n, C.byref(h)))
        self._h = h

    def run(self, iterations: int, use_graph: bool = True) -> Dict:
        st = L.RunStats()
        L.check(L.lib().pskel_batch_run(self._h, int(iterations), 1 if use_graph else 0, C.byref(st)))
        return {"elapsed_ms": st.elapsed_ms, "launches": st.launches, "sweeps": st.sweeps}

    def info(self) -> Dict:
        a, b = C.c_int32(), C.c_int32()
        L.check(L.lib().pskel_batch_info(self._h, C.byref(a), C.byref(b)))
        return {"items": a.value, "blocks": b.value}

    def close(self):
        if getattr(self, "_h", None):
            L.lib().pskel_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_host_tiled(stencil: Stencil, inp: np.ndarray, iterations: int, tile: int, inner: int,
                   out: Optional[np.ndarray] = None, storage: Optional[str] = None, variant: str = "tuned",
                   launch_tile: Tuple[int, int, int] = (0, 0, 0), time_tile: int = 1, store_offset: float = 0.0,
                   rounding: str = "nearest", seed: int = 0, device: int = 0, use_graph: bool = False):
    """Host-staged sweeps (runIterativeTilingGPU, include/PSkelStencil.hpp:1663-1760):
    the arrays stay in host memory, slabs of `tile` rows/planes plus a halo of
    range*inner visit the device for `inner` sweeps at a time.  Host arrays are
    in the storage type (raw uint16 words for bf16/f16).  Returns (result, stats)."""
    inp = np.ascontiguousarray(inp)
    out0 = np.ascontiguousarray(inp if out is None else out).copy()
    if storage is None:
        storage = {np.dtype(np.float32): "f32", np.dtype(np.uint8): "u8"}[inp.dtype]
    st_code = _STORAGE[storage]
    gd = L.GridDesc()
    gd.ndim = inp.ndim
    gd.storage = st_code
    gd.height = inp.shape[0]
    gd.width = inp.shape[1]
    gd.depth = inp.shape[2] if inp.ndim == 3 else 1
    gd.store_offset = store_offset
    gd.device = device
    gd.rounding = {"nearest": L.ROUND_NEAREST, "stochastic": L.ROUND_STOCHASTIC}[rounding]
    gd.seed = int(seed) & 0xFFFFFFFF
    cfg = make_cfg(variant, (0, 0), launch_tile, time_tile, 0, use_graph)
    stats = L.RunStats()
    d = stencil.desc()
    host_dtype = {np.dtype(np.float32): L.DTYPE_F32, np.dtype(np.uint8): L.DTYPE_U8,
                  np.dtype(np.uint16): st_code if storage in ("bf16", "f16") else 0}.get(inp.dtype, 0)
    L.check(L.lib().pskel_run_iterative_host(C.byref(gd), inp.ctypes.data, out0.ctypes.data, host_dtype, C.byref(d),
                                             int(iterations), int(tile), int(inner), C.byref(cfg), C.byref(stats)))
    return out0, {"elapsed_ms": stats.elapsed_ms, "launches": stats.launches, "sweeps": stats.sweeps}


def make_cfg(variant="plain", block=(0, 0), tile=(0, 0, 0), time_tile=1, residual_every=0,
             use_graph=False, stream=0) -> L.LaunchCfg:
    cfg = L.LaunchCfg()
    cfg.variant = {"plain": L.VARIANT_PLAIN, "tuned": L.VARIANT_TUNED}[variant]
    cfg.block_x, cfg.block_y = int(block[0]), int(block[1])
    cfg.tile_x, cfg.tile_y, cfg.tile_z = (int(t) for t in tile)
    cfg.time_tile = int(time_tile)
    cfg.residual_every = int(residual_every)
    cfg.use_graph = 1 if use_graph else 0
    cfg.stream = stream or None
    return cfg


def synchronize():
    L.check(L.lib().pskel_device_synchronize())


def device_info(device: int = 0) -> Dict:
    name = C.create_string_buffer(256)
    free, total = C.c_size_t(), C.c_size_t()
    cus = C.c_int32()
    L.check(L.lib().pskel_device_info(device, name, 256, C.byref(free), C.byref(total), C.byref(cus)))
    return {"name": name.value.decode(), "mem_free": free.value, "mem_total": total.value, "cus": cus.value}


# ---------------------------------------------------------------------------
# PSkel mirror
# ---------------------------------------------------------------------------
class Array2D:
    """PSkel::Array2D<T>(width, height) (include/PSkelArray.h:327-354)."""

    def __init__(self, width: int = 0, height: int = 0, dtype=np.float32):
        self.hostArray = np.zeros((height, width), dtype)

    def getWidth(self):
        return self.hostArray.shape[1]

    def getHeight(self):
        return self.hostArray.shape[0]

    def getDepth(self):
        return 1

    def size(self):
        return self.hostArray.size

    def typeSize(self):
        return self.hostArray.dtype.itemsize

    def memSize(self):
        return self.hostArray.nbytes

    def __call__(self, h, w):
        return self.hostArray[h, w]

    def __setitem__(self, hw, v):
        self.hostArray[hw] = v

    def __getitem__(self, hw):
        return self.hostArray[hw]


class Array3D(Array2D):
    """PSkel::Array3D<T>(width, height, depth), d contiguous (include/PSkelArray.h:294-321)."""

    def __init__(self, width: int = 0, height: int = 0, depth: int = 0, dtype=np.float32):
        self.hostArray = np.zeros((height, width, depth), dtype)

    def getDepth(self):
        return self.hostArray.shape[2]

    def __call__(self, h, w, d):
        return self.hostArray[h, w, d]


class _MaskBase:
    dimension = 2

    def __init__(self, size: int = 0, haloVal: float = 0.0, range: int = 0):
        self.size = size
        self.haloValue = haloVal
        self.range = range
        self.offsets = [(0,) * self.dimension] * size
        self.weights = [0.0] * size

    def getWeight(self, n):
        return self.weights[n]

    def getRange(self):
        """Chebyshev radius, or the range given at construction (include/PSkelMask.hpp:308-334)."""
        if self.range > 0:
            return self.range
        return max(max(abs(v) for v in o) for o in self.offsets)


class Mask2D(_MaskBase):
    """PSkel::Mask2D<T>(size, haloVal, range) (include/PSkelMask.h:177-214)."""

    dimension = 2

    def set(self, n, h, w, weight=0.0):
        self.offsets[n] = (int(h), int(w))
        self.weights[n] = float(weight)


class Mask3D(_MaskBase):
    """PSkel::Mask3D<T>(size, haloVal, range) (include/PSkelMask.h:127-171)."""

    dimension = 3

    def set(self, n, h, w, d, weight=0.0):
        self.offsets[n] = (int(h), int(w), int(d))
        self.weights[n] = float(weight)


class StencilArgs:
    """The kernel family a user's stencilKernel computes, plus its POD args.

    kind: kernels.SCALED_SUM / WEIGHTED_SUM / LIFE; coeff for scaled sums;
    boundary / halo for 3D.  Weights come from the mask (Mask::getWeight)."""

    def __init__(self, kind: int, coeff: float = 0.0, boundary: int = 0):
        self.kind = kind
        self.coeff = coeff
        self.boundary = boundary


class _StencilBase:
    ndim = 2

    def __init__(self, input: Array2D, output: Array2D, mask: _MaskBase, args):
        self.input = input
        self.output = output
        self.mask = mask
        self.args = args

    def _stencil(self) -> Stencil:
        if isinstance(self.args, Stencil):
            return self.args
        a = self.args
        return Stencil(a.kind, self.ndim, list(self.mask.offsets), weights=list(self.mask.weights),
                       coeff=a.coeff, coeff_f64=a.coeff, radius=self.mask.range, boundary=a.boundary,
                       halo_value=self.mask.haloValue)

    def _storage(self) -> str:
        return "u8" if self.input.hostArray.dtype == np.uint8 else "f32"

    def _run(self, iterations, **kw):
        shape = self.input.hostArray.shape
        with Grid(shape, self._storage()) as g:
            g.upload(L.BUFFER_INPUT, self.input.hostArray)
            g.upload(L.BUFFER_OUTPUT, self.output.hostArray)
            stats = g.run(self._stencil(), iterations, **kw)
            self.output.hostArray[...] = g.download(self.output.hostArray.dtype)
        return stats

    def runGPU(self, GPUBlockSizeX: int = 0, GPUBlockSizeY: int = 0):
        """One sweep (include/PSkelStencil.hpp:730-762)."""
        return self.runIterativeGPU(1, GPUBlockSizeX, GPUBlockSizeY)

    def runIterativeGPU(self, iterations: int, GPUBlockSizeX: int = 0, GPUBlockSizeY: int = 0):
        """The plain contract: one lane per cell, (bx, by) blocks
        (include/PSkelStencil.hpp:1047-1097)."""
        return self._run(iterations, variant="plain", block=(GPUBlockSizeX, GPUBlockSizeY))

    def runIterativeTilingGPU(self, iterations: int, tilingWidth: int, tilingHeight: int, tilingDepth: int,
                              innerIterations: int = 1, GPUBlockSizeX: int = 0, GPUBlockSizeY: int = 0):
        """Tiled run (include/PSkelStencil.hpp:1663-1760).  The grid is resident, so
        tiles are LDS tiles: innerIterations is the temporal depth (sweeps per
        launch, capped at 4 in 2D, 8 for life, 3 in 3D for the f32 7-point
        Jacobi and 2 for other 3D grids); the result is identical to
        runIterativeGPU."""
        st = self._stencil()
        if self.ndim == 3:
            cap = 3 if self._storage() == "f32" and st.kind == SCALED_SUM and len(st.offsets) == 6 else 2
        else:
            cap = 8 if st.kind == LIFE else 4
        k = max(1, min(int(innerIterations), cap))
        return self._run(iterations, variant="tuned", time_tile=k)

    def runTilingGPU(self, tilingWidth: int, tilingHeight: int, tilingDepth: int, GPUBlockSizeX: int = 0,
                     GPUBlockSizeY: int = 0):
        return self.runIterativeTilingGPU(1, tilingWidth, tilingHeight, tilingDepth, 1, GPUBlockSizeX, GPUBlockSizeY)

    def runIterativeAutoGPU(self, iterations: int, GPUBlockSize: int = 0):
        """Tuned run with the launch shape chosen by the cost model
        (include/PSkelStencil.hpp:2007-2076)."""
        from .tuner import choose_config  # local import: the tuner pulls in its model file

        shape = self.input.hostArray.shape
        cfg = choose_config(self._stencil(), shape, self._storage(), iterations)
        return self._run(iterations, **cfg)

    def runAutoGPU(self, GPUBlockSize: int = 0):
        return self.runIterativeAutoGPU(1, GPUBlockSize)


class Stencil2D(_StencilBase):
    """PSkel::Stencil2D<Array2D<T>, Mask2D<T>, Args> (include/PSkelStencil.h:349-377)."""

    ndim = 2


class Stencil3D(_StencilBase):
    """PSkel::Stencil3D<Array3D<T>, Mask3D<T>, Args> (include/PSkelStencil.h:317-343)."""

    ndim = 3
