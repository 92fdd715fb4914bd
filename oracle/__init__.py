"""CPU oracle of PSkel's runIterativeCPU -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker (cpu_baseline times it as the CPU port).
The product (pskel_amd, libpskel_amd.so, include/) never imports, links or runs
it.

Wraps oracle/liboracle.so (oracle/pskel_oracle.c, built by oracle/Makefile).
Every runner mirrors the reference in place: `inp` and `out` are numpy arrays
updated exactly like runIterativeCPU updates input/output
(include/PSkelStencil.hpp:880-938): `inp` ends up clobbered, the result is in
`out`.  The convenience functions return copies instead.

The reference ships no golden vectors or known-answer tests for this path,
so each function is pinned by what the mathematics fixes.  Pinning status per
function (see tests/test_oracle_*.py):
  run_2d f32/f64   closed-form eigenmode decay, fixed points, linearity,
                   reflection symmetry, moment drift of the 9-point automaton,
                   brute force on tiny grids
  run_2d u8 life   blinker / block / glider / beehive patterns, brute force
  run_3d f32/f64   separable eigenmode decay, u == 1 fixed point, halo
                   boundary closed form, residual closed form
  run_3d r16       conversion round trips against numpy's float16, RNE ties,
                   equality with the f32 runner on exactly representable data;
                   stochastic rounding: exhaustive probability counts, closed
                   forms, recorded gfx950 v_cvt_sr outputs, unbiasedness
  run_2d_cloud     zero-wind conduction eigenmode decay and fixed points, pure
                   advection of a linear ramp (closed form), upwind side chosen
                   by the wind sign, the Args2D (x, y) -> y*width + x field map
  mlp              finite differences of the loss (central, FP64)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle.so")

MAX_POINTS = 64
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2


class OracleDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("npoints", C.c_int32),
        ("offsets", C.c_int32 * (3 * MAX_POINTS)),
        ("weights", C.c_float * MAX_POINTS),
        ("coeff", C.c_float),
        ("coeff_f64", C.c_double),
        ("radius", C.c_int32),
        ("boundary", C.c_int32),
        ("halo_value", C.c_float),
        ("store_offset", C.c_float),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        P, SZ, I = C.c_void_p, C.c_size_t, C.c_int
        D = C.POINTER(OracleDesc)
        for name in ("oracle_run_iterative_2d_f32", "oracle_run_iterative_2d_f64", "oracle_run_iterative_2d_u8"):
            f = getattr(l, name)
            f.argtypes = [D, P, P, SZ, SZ, SZ, I]
            f.restype = None
        for name in ("oracle_run_iterative_3d_f32", "oracle_run_iterative_3d_f64"):
            f = getattr(l, name)
            f.argtypes = [D, P, P, SZ, SZ, SZ, SZ, I, SZ, P]
            f.restype = SZ
        l.oracle_run_iterative_3d_r16.argtypes = [D, I, P, P, SZ, SZ, SZ, SZ, I, SZ, P]
        l.oracle_run_iterative_3d_r16.restype = SZ
        l.oracle_run_iterative_3d_r16_mode.argtypes = [D, I, I, C.c_uint32, C.c_uint64, P, P, SZ, SZ, SZ, SZ, I, SZ, P]
        l.oracle_run_iterative_3d_r16_mode.restype = SZ
        l.oracle_sr_bits.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64]
        l.oracle_sr_bits.restype = C.c_uint32
        l.oracle_f32_to_bf16_sr.argtypes = [C.c_float, C.c_uint32]
        l.oracle_f32_to_bf16_sr.restype = C.c_uint16
        l.oracle_f32_to_f16_sr.argtypes = [C.c_float, C.c_uint32]
        l.oracle_f32_to_f16_sr.restype = C.c_uint16
        l.oracle_run_iterative_2d_cloud.argtypes = [P, P, SZ, SZ, P, P, SZ, C.c_float, SZ, I]
        l.oracle_run_iterative_2d_cloud.restype = None
        l.oracle_range.argtypes = [D, I]
        l.oracle_range.restype = C.c_int32
        l.oracle_f32_to_bf16.argtypes = [C.c_float]
        l.oracle_f32_to_bf16.restype = C.c_uint16
        l.oracle_bf16_to_f32.argtypes = [C.c_uint16]
        l.oracle_bf16_to_f32.restype = C.c_float
        l.oracle_f32_to_f16.argtypes = [C.c_float]
        l.oracle_f32_to_f16.restype = C.c_uint16
        l.oracle_f16_to_f32.argtypes = [C.c_uint16]
        l.oracle_f16_to_f32.restype = C.c_float
        l.oracle_mlp_loss_grad.argtypes = [P, I, P, P, P, SZ, P]
        l.oracle_mlp_loss_grad.restype = C.c_double
        l.oracle_mlp_predict.argtypes = [P, I, P, P, SZ, P]
        l.oracle_mlp_predict.restype = None
        l.oracle_last_seconds.restype = C.c_double
        l.oracle_num_procs.restype = C.c_int
        _lib = l
    return _lib


def desc_of(stencil, store_offset: float = 0.0) -> OracleDesc:
    """Oracle descriptor of a pskel_amd.kernels.Stencil (same fields, its own struct)."""
    d = OracleDesc()
    d.kind = stencil.kind
    d.npoints = stencil.npoints
    flat = [v for o in stencil.offsets for v in o]
    for i, v in enumerate(flat):
        d.offsets[i] = int(v)
    for i, v in enumerate(stencil.f32_weights()):
        d.weights[i] = float(v)
    d.coeff = float(np.float32(stencil.coeff))
    d.coeff_f64 = float(stencil.coeff_f64)
    d.radius = int(stencil.radius)
    d.boundary = int(stencil.boundary)
    d.halo_value = float(np.float32(stencil.halo_value))
    d.store_offset = float(np.float32(store_offset))
    return d


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data


def run_2d_inplace(stencil, inp: np.ndarray, out: np.ndarray, iterations: int, threads: int = 0):
    """runIterativeCPU on Array2D-shaped (height, width) arrays, in place."""
    l = lib()
    d = desc_of(stencil)
    H, W = inp.shape
    assert out.shape == inp.shape and out.dtype == inp.dtype
    fn = {np.dtype(np.float32): l.oracle_run_iterative_2d_f32, np.dtype(np.float64): l.oracle_run_iterative_2d_f64,
          np.dtype(np.uint8): l.oracle_run_iterative_2d_u8}[inp.dtype]
    fn(C.byref(d), _ptr(inp), _ptr(out), W, H, int(iterations), int(threads))


def run_2d(stencil, inp: np.ndarray, iterations: int, out: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    """Result of runIterativeCPU (copies; `out` defaults to a copy of `inp`)."""
    a = np.array(inp, copy=True, order="C")
    b = np.array(inp if out is None else out, copy=True, order="C")
    run_2d_inplace(stencil, a, b, iterations, threads)
    return b


def run_3d(stencil, inp: np.ndarray, iterations: int, out: np.ndarray | None = None, threads: int = 0,
           residual_every: int = 0):
    """(result, residuals) of runIterativeCPU on (height, width, depth) f32/f64 arrays."""
    l = lib()
    d = desc_of(stencil)
    a = np.array(inp, copy=True, order="C")
    b = np.array(inp if out is None else out, copy=True, order="C")
    H, W, D = a.shape
    nres = iterations // residual_every if residual_every else 0
    res = np.zeros(max(nres, 1), np.float64)
    fn = l.oracle_run_iterative_3d_f32 if a.dtype == np.float32 else l.oracle_run_iterative_3d_f64
    n = fn(C.byref(d), _ptr(a), _ptr(b), W, H, D, int(iterations), int(threads), int(residual_every), _ptr(res))
    return b, res[:n]


def run_2d_cloud(inp: np.ndarray, wind_x: np.ndarray, wind_y: np.ndarray, delta_t: float, iterations: int,
                 out: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    """Result of runIterativeCPU with the CLOUDSIM kernel on an (H, W) f32 grid.
    wind_x / wind_y are Args2D host arrays, shape (field_height, field_width),
    indexed [y, x]; the kernel reads wind(x=h, y=w)."""
    a = np.array(inp, dtype=np.float32, copy=True, order="C")
    b = np.array(inp if out is None else out, dtype=np.float32, copy=True, order="C")
    wx = np.ascontiguousarray(wind_x, np.float32)
    wy = np.ascontiguousarray(wind_y, np.float32)
    H, W = a.shape
    assert wx.shape == wy.shape and wx.shape[1] >= H and wx.shape[0] >= W
    lib().oracle_run_iterative_2d_cloud(_ptr(a), _ptr(b), W, H, _ptr(wx), _ptr(wy), wx.shape[1],
                                        float(np.float32(delta_t)), int(iterations), int(threads))
    return b


ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1


def run_3d_r16(stencil, store: int, inp_words: np.ndarray, iterations: int, store_offset: float = 0.0,
               out_words: np.ndarray | None = None, threads: int = 0, residual_every: int = 0,
               rounding: int = ROUND_NEAREST, seed: int = 0, first_sweep: int = 0):
    """(result words, residuals) with bf16/f16 storage of (u - store_offset)."""
    l = lib()
    d = desc_of(stencil, store_offset)
    a = np.array(inp_words, dtype=np.uint16, copy=True, order="C")
    b = np.array(inp_words if out_words is None else out_words, dtype=np.uint16, copy=True, order="C")
    H, W, D = a.shape
    nres = iterations // residual_every if residual_every else 0
    res = np.zeros(max(nres, 1), np.float64)
    n = l.oracle_run_iterative_3d_r16_mode(C.byref(d), int(store), int(rounding), int(seed), int(first_sweep),
                                           _ptr(a), _ptr(b), W, H, D, int(iterations), int(threads),
                                           int(residual_every), _ptr(res))
    return b, res[:n]


def sr_bits(seed: int, sweep: int, h: int, wd: int) -> int:
    """32-bit random word of cell (h, w, d) with wd = w*D + d."""
    return lib().oracle_sr_bits(seed, sweep, h, wd)


def encode16_sr(x: float, store: int, r: int) -> int:
    l = lib()
    f = l.oracle_f32_to_bf16_sr if store == STORE_BF16 else l.oracle_f32_to_f16_sr
    return f(float(np.float32(x)), int(r))


def encode16(x: np.ndarray, store: int, offset: float = 0.0) -> np.ndarray:
    l = lib()
    f = l.oracle_f32_to_bf16 if store == STORE_BF16 else l.oracle_f32_to_f16
    flat = np.asarray(x, np.float32).ravel()
    off = np.float32(offset)
    out = np.fromiter((f(float(np.float32(v - off))) for v in flat), dtype=np.uint16, count=flat.size)
    return out.reshape(np.shape(x))


def decode16(w: np.ndarray, store: int, offset: float = 0.0) -> np.ndarray:
    """Vectorised decode (u = decoded + offset, in f32)."""
    w = np.asarray(w, np.uint16)
    if store == STORE_BF16:
        v = (w.astype(np.uint32) << np.uint32(16)).view(np.float32)
    else:
        v = w.view(np.float16).astype(np.float32)
    return (v + np.float32(offset)).astype(np.float32)


def range_of(stencil) -> int:
    return lib().oracle_range(C.byref(desc_of(stencil)), stencil.ndim)


def mlp_loss_grad(sizes, params: np.ndarray, x: np.ndarray, y: np.ndarray, want_grad: bool = True):
    l = lib()
    s = np.asarray(sizes, np.int32)
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    yy = np.ascontiguousarray(y, np.float64)
    g = np.zeros_like(p) if want_grad else None
    loss = l.oracle_mlp_loss_grad(_ptr(s), len(sizes) - 1, _ptr(p), _ptr(xx), _ptr(yy), yy.size,
                                  _ptr(g) if want_grad else None)
    return loss, g


def mlp_predict(sizes, params: np.ndarray, x: np.ndarray) -> np.ndarray:
    l = lib()
    s = np.asarray(sizes, np.int32)
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    y = np.zeros(xx.shape[0], np.float64)
    l.oracle_mlp_predict(_ptr(s), len(sizes) - 1, _ptr(p), _ptr(xx), xx.shape[0], _ptr(y))
    return y


def last_seconds() -> float:
    return lib().oracle_last_seconds()


def num_procs() -> int:
    return lib().oracle_num_procs()
