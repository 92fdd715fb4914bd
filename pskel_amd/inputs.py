"""Deterministic synthetic inputs for the benchmark workloads.

Counter-based: cell i of a field seeded with s draws splitmix64((s << 40) ^ i),
so any slice of any field can be regenerated without the rest and the same
numbers come out on every machine.  Uniform floats take the top 24 bits
times 2^-24 (exactly representable in f32).
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + _M1
        z = (z ^ (z >> np.uint64(30))) * _M2
        z = (z ^ (z >> np.uint64(27))) * _M3
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, start: int, n: int) -> np.ndarray:
    """Uniform [0,1) float32 for counters start .. start+n-1."""
    idx = np.arange(start, start + n, dtype=np.uint64)
    z = splitmix64(idx ^ np.uint64(seed << 40))
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def uniform64(seed: int, n: int) -> np.ndarray:
    """Uniform [0,1) float64 (53 bits) for counters 0..n-1."""
    z = splitmix64(np.arange(n, dtype=np.uint64) ^ np.uint64(seed << 40))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def jacobi2d_field(height: int, width: int, seed: int = 1) -> np.ndarray:
    """Interior uniform [0,1), fixed zero ring (Dirichlet)."""
    a = uniform(seed, 0, height * width).reshape(height, width)
    a[0, :] = 0
    a[-1, :] = 0
    a[:, 0] = 0
    a[:, -1] = 0
    return a


def life_field(n: int, seed: int = 2, p: float = 0.3) -> np.ndarray:
    """n x n logical cells alive with probability p, stored with a dead ring:
    an (n+2) x (n+2) u8 array whose interior is the n x n board, so the
    interior-only 2D sweep updates every logical cell and out-of-bounds
    neighbours read dead cells."""
    out = np.zeros((n + 2, n + 2), np.uint8)
    for r0 in range(0, n, 1024):
        r1 = min(n, r0 + 1024)
        u = uniform(seed, r0 * n, (r1 - r0) * n).reshape(r1 - r0, n)
        out[1 + r0:1 + r1, 1:-1] = (u < np.float32(p)).astype(np.uint8)
    return out


def blobs_field(n: int, seed: int = 3, nblobs: int = 32, sigma: float = 40.0,
                background: float = 0.1) -> np.ndarray:
    """Background plus Gaussian blobs (centres uniform over the grid,
    amplitudes uniform in [0.5, 2]); computed in f64, stored f32."""
    u = uniform64(seed, 3 * nblobs).reshape(nblobs, 3)
    cy = u[:, 0] * n
    cx = u[:, 1] * n
    amp = 0.5 + 1.5 * u[:, 2]
    out = np.empty((n, n), np.float32)
    xs = np.arange(n, dtype=np.float64)
    inv = 1.0 / (2.0 * sigma * sigma)
    for r0 in range(0, n, 256):
        r1 = min(n, r0 + 256)
        ys = np.arange(r0, r1, dtype=np.float64)
        acc = np.full((r1 - r0, n), background, np.float64)
        for k in range(nblobs):
            gy = np.exp(-((ys - cy[k]) ** 2) * inv)
            gx = np.exp(-((xs - cx[k]) ** 2) * inv)
            acc += amp[k] * np.outer(gy, gx)
        out[r0:r1] = acc.astype(np.float32)
    return out


def jacobi3d_field(n: int, seed: int = 4, face: float = 1.0, dtype=np.float32) -> np.ndarray:
    """n^3 uniform [0,1) with every boundary face fixed at `face`;
    array order (h, w, d) = Array3D's (h*W + w)*D + d."""
    out = np.empty((n, n, n), dtype)
    plane = n * n
    for h in range(n):
        out[h] = uniform(seed, h * plane, plane).reshape(n, n)
    out[0] = face
    out[-1] = face
    out[:, 0] = face
    out[:, -1] = face
    out[:, :, 0] = face
    out[:, :, -1] = face
    return out


def batch_shapes(count: int = 256, seed: int = 5, lo: int = 64, hi: int = 2048) -> List[Tuple[int, int, str]]:
    """(height, width, kind) for the batched workload: each side log-uniform in
    [lo, hi], independent per axis; kinds alternate jacobi / life."""
    u = uniform64(seed, 2 * count).reshape(count, 2)
    a, b = math.log(lo), math.log(hi)
    out = []
    for i in range(count):
        h = int(round(math.exp(a + (b - a) * u[i, 0])))
        w = int(round(math.exp(a + (b - a) * u[i, 1])))
        out.append((h, w, "jacobi" if i % 2 == 0 else "life"))
    return out


def cloudsim_fields(height: int, width: int, seed: int = 6):
    """CLOUDSIM inputs: an (height, width) temperature field (interior uniform
    in [15, 25), ring fixed at 20) and the wind components as Args2D host
    arrays of shape (width, height), indexed [y, x] with x = grid row h and
    y = grid column w (the map the kernel reads them through).  The wind is a
    smooth swirl of magnitude <= 1 plus a small uniform perturbation, so both
    signs occur (and with them both upwind sides)."""
    t = (np.float32(15.0) + np.float32(10.0) * uniform(seed, 0, height * width)).reshape(height, width)
    t[0, :] = t[-1, :] = 20.0
    t[:, 0] = t[:, -1] = 20.0
    h = np.arange(height, dtype=np.float64)[None, :]   # x = h
    w = np.arange(width, dtype=np.float64)[:, None]    # y = w
    jitter_x = uniform(seed + 1, 0, width * height).reshape(width, height).astype(np.float64) - 0.5
    jitter_y = uniform(seed + 2, 0, width * height).reshape(width, height).astype(np.float64) - 0.5
    wind_x = (0.8 * np.sin(2 * math.pi * w / width) + 0.2 * jitter_x).astype(np.float32)
    wind_y = (0.8 * np.cos(2 * math.pi * h / height) + 0.2 * jitter_y).astype(np.float32)
    return t.astype(np.float32), wind_x, wind_y
