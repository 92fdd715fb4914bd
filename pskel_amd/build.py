"""Build libpskel_amd.so (gfx950) and the CPU oracle in-tree.

    python -m pskel_amd.build          # rebuild everything
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pskel_amd", "csrc")


def build_library(force: bool = True, jobs: int = 8) -> str:
    cmd = ["make", f"-j{jobs}", "-C", CSRC]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(ROOT, "pskel_amd", "libpskel_amd.so")


def build_oracle(force: bool = True) -> str:
    cmd = ["make", "-C", os.path.join(ROOT, "oracle")]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_dropin_drivers(force: bool = True) -> str:
    """The C++ programs that exercise the drop-in headers (tests/cpp)."""
    d = os.path.join(ROOT, "tests", "cpp")
    cmd = ["make", "-C", d]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(d, "bin")


def main(argv=None):
    build_library()
    build_oracle()
    build_dropin_drivers()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
