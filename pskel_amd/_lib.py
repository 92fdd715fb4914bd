"""ctypes binding of the C ABI declared in include/pskel_amd.h.

The shared library libpskel_amd.so is built in-tree (pskel_amd/build.py).  There
is no fallback of any kind: if the library is missing, importing the GPU entry
points raises, and if the library loads but finds no device, every call returns
PSKEL_ERR_NO_DEVICE and raises PSkelError.
"""
from __future__ import annotations

import ctypes as C
import os

MAX_POINTS = 64

# pskel_status_t
OK = 0
SEARCH_EXHAUSTIVE, SEARCH_GENETIC = 0, 1
ERR_INVALID_ARGUMENT = 1
ERR_OUT_OF_MEMORY = 2
ERR_NO_DEVICE = 3
ERR_LAUNCH = 4
ERR_UNSUPPORTED = 5
ERR_RUNTIME = 6

# pskel_dtype_t
DTYPE_U8 = 1
DTYPE_F32 = 2
DTYPE_F64 = 3
DTYPE_BF16 = 4
DTYPE_F16 = 5

# pskel_kind_t
KIND_SCALED_SUM = 1
KIND_WEIGHTED_SUM = 2
KIND_LIFE = 3
KIND_CLOUDSIM = 4

BOUNDARY_COPY = 0
BOUNDARY_HALO = 1

VARIANT_PLAIN = 0
VARIANT_TUNED = 1

ROUND_NEAREST = 0
ROUND_STOCHASTIC = 1

BUFFER_INPUT = 0
BUFFER_OUTPUT = 1


class StencilDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("npoints", C.c_int32),
        ("offsets", C.c_int32 * (3 * MAX_POINTS)),
        ("weights", C.c_float * MAX_POINTS),
        ("coeff", C.c_float),
        ("radius", C.c_int32),
        ("boundary", C.c_int32),
        ("halo_value", C.c_float),
        ("reserved", C.c_int32 * 4),
    ]


class LaunchCfg(C.Structure):
    _fields_ = [
        ("variant", C.c_int32),
        ("block_x", C.c_int32),
        ("block_y", C.c_int32),
        ("tile_x", C.c_int32),
        ("tile_y", C.c_int32),
        ("tile_z", C.c_int32),
        ("time_tile", C.c_int32),
        ("residual_every", C.c_int32),
        ("use_graph", C.c_int32),
        ("reserved", C.c_int32 * 3),
        ("stream", C.c_void_p),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("elapsed_ms", C.c_double),
        ("launches", C.c_int64),
        ("sweeps", C.c_int64),
        ("residuals", C.POINTER(C.c_double)),
        ("residual_capacity", C.c_int64),
        ("residual_count", C.c_int64),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("storage", C.c_int32),
        ("width", C.c_size_t),
        ("height", C.c_size_t),
        ("depth", C.c_size_t),
        ("store_offset", C.c_float),
        ("device", C.c_int32),
        ("rounding", C.c_int32),
        ("seed", C.c_uint32),
    ]


class GridInfo(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("storage", C.c_int32),
        ("elem_bytes", C.c_int32),
        ("parity", C.c_int32),
        ("width", C.c_size_t),
        ("height", C.c_size_t),
        ("depth", C.c_size_t),
        ("pitch", C.c_size_t),
        ("bytes_per_buffer", C.c_size_t),
        ("store_offset", C.c_float),
        ("current", C.c_void_p),
        ("previous", C.c_void_p),
        ("stream", C.c_void_p),
    ]


class TuneQuery(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("storage", C.c_int32),
        ("width", C.c_size_t),
        ("height", C.c_size_t),
        ("depth", C.c_size_t),
        ("iterations", C.c_size_t),
    ]


class FieldDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("dtype", C.c_int32),
        ("width", C.c_size_t),
        ("height", C.c_size_t),
        ("depth", C.c_size_t),
        ("device", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


TUNER_MAX_CANDIDATES = 256
TUNER_NFEAT = 26

STRUCTS = [StencilDesc, LaunchCfg, RunStats, GridDesc, GridInfo, FieldDesc]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpskel_amd.so")


class PSkelError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"pskel status {status}: {msg}")
        self.status = status


_lib = None


def _sig(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)


def _declare(lib):
    P, I32, SZ, D = C.c_void_p, C.c_int32, C.c_size_t, C.c_double
    st = C.c_int
    _sig(lib, "pskel_abi_version", I32)
    _sig(lib, "pskel_abi_struct_size", SZ, I32)
    _sig(lib, "pskel_status_string", C.c_char_p, st)
    _sig(lib, "pskel_last_error", C.c_char_p)
    _sig(lib, "pskel_device_count", st, C.POINTER(I32))
    _sig(lib, "pskel_device_synchronize", st)
    _sig(lib, "pskel_device_info", st, I32, C.c_char_p, SZ, C.POINTER(SZ), C.POINTER(SZ), C.POINTER(I32))
    _sig(lib, "pskel_grid_create", st, C.POINTER(GridDesc), C.POINTER(P))
    _sig(lib, "pskel_grid_destroy", st, P)
    _sig(lib, "pskel_grid_upload", st, P, I32, P, I32)
    _sig(lib, "pskel_grid_download", st, P, P, I32)
    _sig(lib, "pskel_grid_download_previous", st, P, P, I32)
    _sig(lib, "pskel_grid_info", st, P, C.POINTER(GridInfo))
    _sig(lib, "pskel_run_iterative", st, P, C.POINTER(StencilDesc), SZ, C.POINTER(LaunchCfg), C.POINTER(RunStats))
    _sig(lib, "pskel_run_iterative_noalloc", st, C.POINTER(GridDesc), P, P, SZ, C.POINTER(StencilDesc), SZ,
         C.POINTER(LaunchCfg), C.POINTER(RunStats))
    _sig(lib, "pskel_grid_residual", st, P, C.POINTER(D))
    _sig(lib, "pskel_grid_swap", st, P)
    _sig(lib, "pskel_grid_set_field", st, P, I32, P)
    _sig(lib, "pskel_run_iterative_host", st, C.POINTER(GridDesc), P, P, I32, C.POINTER(StencilDesc), SZ, SZ, SZ,
         C.POINTER(LaunchCfg), C.POINTER(RunStats))
    _sig(lib, "pskel_field_create", st, C.POINTER(FieldDesc), C.POINTER(P))
    _sig(lib, "pskel_field_destroy", st, P)
    _sig(lib, "pskel_field_upload", st, P, P)
    _sig(lib, "pskel_field_download", st, P, P)
    _sig(lib, "pskel_field_info", st, P, C.POINTER(P), C.POINTER(SZ))
    F = C.POINTER(C.c_float)
    _sig(lib, "pskel_batch_create", st, C.POINTER(P), C.POINTER(StencilDesc), C.POINTER(LaunchCfg), SZ, C.POINTER(P))
    _sig(lib, "pskel_batch_run", st, P, SZ, I32, C.POINTER(RunStats))
    _sig(lib, "pskel_batch_info", st, P, C.POINTER(I32), C.POINTER(I32))
    _sig(lib, "pskel_batch_destroy", st, P)
    _sig(lib, "pskel_tuner_candidates", st, C.POINTER(TuneQuery), C.POINTER(StencilDesc), C.POINTER(LaunchCfg), SZ,
         C.POINTER(SZ))
    _sig(lib, "pskel_tuner_features", st, C.POINTER(TuneQuery), C.POINTER(StencilDesc), C.POINTER(LaunchCfg), F)
    _sig(lib, "pskel_tuner_choose", st, P, C.POINTER(TuneQuery), C.POINTER(StencilDesc), C.POINTER(LaunchCfg))
    _sig(lib, "pskel_tuner_heuristic", st, C.POINTER(TuneQuery), C.POINTER(StencilDesc), C.POINTER(LaunchCfg))
    _sig(lib, "pskel_tuner_set_default_model", st, P)
    _sig(lib, "pskel_tuner_measure", st, P, C.POINTER(StencilDesc), SZ, C.POINTER(LaunchCfg), I32, C.POINTER(D))
    _sig(lib, "pskel_tuner_has_model", st, C.POINTER(I32))
    _sig(lib, "pskel_tuner_host_search", st, C.POINTER(GridDesc), P, I32, C.POINTER(StencilDesc), SZ, SZ,
         C.POINTER(SZ), C.POINTER(SZ), C.POINTER(D))
    _sig(lib, "pskel_tuner_search", st, P, C.POINTER(StencilDesc), SZ, I32, SZ, I32, C.POINTER(LaunchCfg),
         C.POINTER(D), C.POINTER(I32))
    _sig(lib, "pskel_mlp_create", st, C.POINTER(I32), I32, C.c_uint64, I32, C.POINTER(P))
    _sig(lib, "pskel_mlp_destroy", st, P)
    _sig(lib, "pskel_mlp_num_params", st, P, C.POINTER(SZ))
    _sig(lib, "pskel_mlp_get_params", st, P, F, SZ)
    _sig(lib, "pskel_mlp_set_params", st, P, F, SZ)
    _sig(lib, "pskel_mlp_set_normalization", st, P, F, F, C.c_float, C.c_float)
    _sig(lib, "pskel_mlp_loss_grad", st, P, F, F, SZ, F, F)
    _sig(lib, "pskel_mlp_train", st, P, F, F, SZ, I32, I32, C.c_float, C.c_uint64, F)
    _sig(lib, "pskel_mlp_predict", st, P, F, SZ, F)
    _sig(lib, "pskel_mlp_save", st, P, C.c_char_p)
    _sig(lib, "pskel_mlp_load", st, C.c_char_p, I32, C.POINTER(P))
    _sig(lib, "pskel_device_copy", st, P, P, SZ, P, C.POINTER(D))
    _sig(lib, "pskel_device_alloc", st, SZ, C.POINTER(P))
    _sig(lib, "pskel_device_free", st, P)
    _sig(lib, "pskel_memcpy", st, P, P, SZ, I32)


def lib():
    """Load libpskel_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m pskel_amd.build`; there is no non-HIP path"
            )
        l = C.CDLL(LIB_PATH)
        _declare(l)
        for i, s in enumerate(STRUCTS):
            got = l.pskel_abi_struct_size(i)
            if got != C.sizeof(s):
                raise ImportError(f"ABI mismatch for {s.__name__}: C {got} bytes, ctypes {C.sizeof(s)}")
        _lib = l
    return _lib


def check(status: int):
    if status != OK:
        l = lib()
        msg = l.pskel_last_error().decode(errors="replace")
        raise PSkelError(status, f"{l.pskel_status_string(status).decode()}: {msg}")


def device_count() -> int:
    n = C.c_int32(0)
    st = lib().pskel_device_count(C.byref(n))
    return n.value if st == OK else 0
