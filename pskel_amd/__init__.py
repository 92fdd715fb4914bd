"""MI355X (gfx950) implementation of PSkel's iterative stencil path.

The product is libpskel_amd.so (C ABI in include/pskel_amd.h) plus the C++
drop-in headers in include/.  This package is the Python view of the same
library: ctypes bindings (_lib), the resident Grid and the PSkel mirror (api),
stencil descriptors (kernels), workload generators (inputs) and the launch
tuner (tuner).
"""
from . import kernels, inputs  # noqa: F401  (pure Python, no device needed)

__all__ = ["kernels", "inputs", "api", "tuner"]


def __getattr__(name):
    # api/tuner load the shared library; import them lazily so that the
    # descriptor and input modules work on machines without the build.
    if name in ("api", "tuner"):
        import importlib

        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
