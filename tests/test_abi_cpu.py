"""The C ABI library without a GPU: it loads, its struct layouts match the
ctypes mirror, and it exports every function include/pskel_amd.h declares
(nothing here calls into a kernel)."""
import ctypes as C
import os
import re

from pskel_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pskel_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w \*]*?\b(pskel_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ("pskel_run_iterative", "pskel_grid_create", "pskel_run_iterative_noalloc", "pskel_batch_run",
              "pskel_tuner_choose", "pskel_mlp_train"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_and_version():
    lib = L.lib()  # checks every ABI struct size against the ctypes mirror
    assert lib.pskel_abi_version() == 1
    assert lib.pskel_abi_struct_size(99) == 0
    assert lib.pskel_status_string(5) == b"unsupported configuration"
