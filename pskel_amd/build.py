"""Build libpskel_amd.so (gfx950) in-tree.

    python -m pskel_amd.build          # rebuild the library

(The test oracle and the drop-in drivers are built by __graft_entry__.build()
and by the Makefiles in oracle/ and tests/cpp/.)
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


def main(argv=None):
    build_library()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
