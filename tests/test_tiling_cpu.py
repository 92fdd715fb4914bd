"""StencilTiling::tile (include/PSkelStencilTiling.h) against what k sweeps of a
range-r stencil need, found by brute force.

A logical tile [off, off+len) along an axis of n cells (cut at the array's
end) is computed exactly by k sweeps only from the cells within r*k of it:
the expanded tile must be that set, clipped to [0, n), and the core the
logical tile inside it.  The expected set is found by dilating the logical
tile k times by r on a boolean array, not from a formula.  Includes the
reference's A9 case, 0 < offset < r*k, where the reference
(include/PSkelStencilTiling.h:108-136 of the PSkel tree) zeroes the offset
before computing its halo correction and leaves the far halo short by
`offset` cells.
"""
import subprocess

import numpy as np
import pytest

import dropin


def _dilated(n, off, length, r, k):
    need = np.zeros(n, bool)
    need[off:min(n, off + length)] = True
    for _ in range(k):
        grown = need.copy()
        for s in range(1, r + 1):
            grown[s:] |= need[:-s]
            grown[:-s] |= need[s:]
        need = grown
    idx = np.flatnonzero(need)
    return int(idx[0]), int(idx[-1] + 1)


def _cases(rng):
    fixed = [
        (100, 50, 30, 1, 3, 0, 0, 0, 10, 10, 10),      # offset 0
        (100, 50, 30, 1, 3, 2, 1, 2, 10, 10, 10),      # A9: 0 < offset < r*k
        (100, 50, 30, 2, 4, 7, 5, 3, 10, 10, 10),      # A9 with range 2
        (100, 50, 30, 1, 3, 3, 3, 3, 10, 10, 10),      # offset == r*k
        (100, 50, 30, 2, 4, 95, 45, 29, 10, 10, 10),   # tile cut at the array's end
        (64, 64, 64, 1, 1, 0, 0, 0, 64, 64, 64),       # whole array
        (40, 40, 40, 3, 5, 20, 20, 20, 1, 1, 1),       # one cell, halo past both ends
    ]
    rand = []
    for _ in range(400):
        n = rng.integers(1, 120, size=3)
        r = int(rng.integers(1, 4))
        k = int(rng.integers(1, 9))
        off = [int(rng.integers(0, x)) for x in n]
        ln = [int(rng.integers(1, x + 1)) for x in n]
        rand.append((int(n[0]), int(n[1]), int(n[2]), r, k, *off, *ln))
    return fixed + rand


def test_tile_geometry_matches_brute_force():
    cases = _cases(np.random.default_rng(7))
    text = "\n".join(" ".join(map(str, c)) for c in cases) + "\n"
    out = subprocess.run([dropin.driver("tile_geometry")], input=text, capture_output=True, text=True,
                         check=True).stdout.split("\n")
    for c, line in zip(cases, out):
        W, H, D, r, k, wo, ho, do, w, h, d = c
        got = list(map(int, line.split()))
        for axis, (n, off, ln) in enumerate(((W, wo, w), (H, ho, h), (D, do, d))):
            t_off, t_len, c_off, c_len = got[4 * axis:4 * axis + 4]
            lo, hi = _dilated(n, off, ln, r, k)
            assert (t_off, t_off + t_len) == (lo, hi), (c, axis)
            assert c_len == min(ln, n - off), (c, axis)
            assert t_off + c_off == off, (c, axis)
            assert c_off + c_len <= t_len, (c, axis)


def test_a9_far_halo_is_complete():
    """The reference's expanded tile for offset 2, halo 3, length 10 ends at
    cell 13 (length 10 + 2*3 - 3); the cells the core needs run to 15."""
    out = subprocess.run([dropin.driver("tile_geometry")], input="100 50 30 1 3 2 0 0 10 10 10\n",
                         capture_output=True, text=True, check=True).stdout.split()
    t_off, t_len, c_off, c_len = map(int, out[:4])
    assert (t_off, t_len, c_off, c_len) == (0, 15, 2, 10)
