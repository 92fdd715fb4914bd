"""Helpers to run the drop-in header drivers (tests/cpp/drive.cpp)."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def driver(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), f"bin/{name}"], check=True)
    return p


def run(name, kernel, method, arr, iters, out0=None, timeout=300, wind=None):
    """Run `drive <kernel> <method>` on a host array (h, w) or (h, w, d);
    `wind` = (wind_x, wind_y) Args2D host arrays for the CLOUDSIM kernels.
    Returns (returncode, result array or None, stderr)."""
    arr = np.ascontiguousarray(arr)
    out0 = arr if out0 is None else np.ascontiguousarray(out0)
    dims = arr.shape
    W = dims[1]
    H = dims[0]
    D = dims[2] if arr.ndim == 3 else 1
    with tempfile.TemporaryDirectory() as t:
        fi, fo, fr = (os.path.join(t, n) for n in ("in.bin", "out0.bin", "res.bin"))
        arr.tofile(fi)
        out0.tofile(fo)
        extra = []
        if wind is not None:
            extra = [os.path.join(t, "wx.bin"), os.path.join(t, "wy.bin")]
            np.ascontiguousarray(wind[0], np.float32).tofile(extra[0])
            np.ascontiguousarray(wind[1], np.float32).tofile(extra[1])
        p = subprocess.run([driver(name), kernel, method, str(W), str(H), str(D), str(iters), fi, fo, fr] + extra,
                           capture_output=True, text=True, timeout=timeout)
        res = None
        if p.returncode == 0:
            res = np.fromfile(fr, dtype=arr.dtype).reshape(arr.shape)
        return p.returncode, res, p.stderr
