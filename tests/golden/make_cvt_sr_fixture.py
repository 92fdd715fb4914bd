"""Writes tests/golden/cvt_sr_gfx950.npz from the output of cvt_sr_probe.hip.

On a gfx950 machine:
    hipcc -O2 --offload-arch=gfx950 -o cvt_sr_probe tests/golden/cvt_sr_probe.hip
    ./cvt_sr_probe probe.bin
then anywhere:
    python tests/golden/make_cvt_sr_fixture.py probe.bin

Keeps every special case (the first 360 records) and a fixed-seed sample of
3736 of the random ones: 4096 (value, random word) pairs with the device's
bf16 and f16 results.
"""
import os
import sys

import numpy as np


def main(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    x = np.frombuffer(raw[4:4 + 4 * n], np.uint32)
    r = np.frombuffer(raw[4 + 4 * n:4 + 8 * n], np.uint32)
    out = np.frombuffer(raw[4 + 8 * n:], np.uint32).reshape(n, 3)
    assert np.array_equal(out[:, 2] >> 16, out[:, 0] & 0xFFFF), "word_sel halves disagree"
    nspec = 360
    rng = np.random.default_rng(2026)
    pick = np.concatenate([np.arange(nspec), nspec + np.sort(rng.choice(n - nspec, 4096 - nspec, replace=False))])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cvt_sr_gfx950.npz")
    np.savez_compressed(dst, x_bits=x[pick], r=r[pick], bf16=(out[pick, 0] & 0xFFFF).astype(np.uint16),
                        f16=(out[pick, 1] & 0xFFFF).astype(np.uint16))
    print("wrote", dst, len(pick), "cases")


if __name__ == "__main__":
    main(sys.argv[1])
