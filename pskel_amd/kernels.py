"""Stencil descriptors: the stock stencil kernels and user-defined masks.

A PSkel user writes `stencilKernel(input, output, mask, args, h, w[, d])`
(include/PSkelStencil.h:121-156) and reads neighbours with Mask::get in mask
order (include/PSkelMask.hpp:186-199, 292-305).  The kernels this library
runs on the GPU are the three families such a kernel computes:

  scaled sum    out = coeff * (((p0 + p1) + p2) + ...)
  weighted sum  out = ((w0*p0 + w1*p1) + w2*p2) + ...
  life          n = p0 + ... ; out = n == 3 or (n == 2 and centre == 1)

A `Stencil` carries the family, the mask (offsets, weights, range) and the
3D boundary rule.  The stock constructors below define the point order the
tuned kernels are specialised for; the C++ drop-in header PSkelKernels.h
defines the same stencils with the same orders.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib as L

SCALED_SUM = L.KIND_SCALED_SUM
WEIGHTED_SUM = L.KIND_WEIGHTED_SUM
LIFE = L.KIND_LIFE
CLOUDSIM = L.KIND_CLOUDSIM

BOUNDARY_COPY = L.BOUNDARY_COPY
BOUNDARY_HALO = L.BOUNDARY_HALO

# canonical point orders (the tuned kernels' accumulation order)
J5_OFFSETS = [(-1, 0), (1, 0), (0, -1), (0, 1)]
BOX9_OFFSETS = [(dh, dw) for dh in (-1, 0, 1) for dw in (-1, 0, 1)]
MOORE8_OFFSETS = [(dh, dw) for dh in (-1, 0, 1) for dw in (-1, 0, 1) if (dh, dw) != (0, 0)]
J7_OFFSETS = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


@dataclass
class Stencil:
    kind: int
    ndim: int
    offsets: List[Tuple[int, ...]]
    weights: List[float] = field(default_factory=list)
    coeff: float = 0.0
    coeff_f64: float = 0.0       # exact coefficient for FP64 reference runs
    radius: int = 0              # 0 -> Mask::getRange()
    boundary: int = BOUNDARY_COPY
    halo_value: float = 0.0
    name: str = "custom"

    def __post_init__(self):
        if not 1 <= len(self.offsets) <= L.MAX_POINTS:
            raise ValueError("a mask has 1..64 points")
        for o in self.offsets:
            if len(o) != self.ndim:
                raise ValueError(f"offset {o} is not {self.ndim}-dimensional")
        if self.weights and len(self.weights) != len(self.offsets):
            raise ValueError("one weight per point")

    @property
    def npoints(self) -> int:
        return len(self.offsets)

    def range(self) -> int:
        """Mask::getRange(): the Chebyshev radius (include/PSkelMask.hpp:308-334)."""
        if self.radius > 0:
            return self.radius
        return max(max(abs(v) for v in o) for o in self.offsets)

    def f32_weights(self) -> np.ndarray:
        w = np.zeros(len(self.offsets), np.float32)
        if self.weights:
            w[:] = np.asarray(self.weights, np.float64).astype(np.float32)
        return w

    def desc(self) -> L.StencilDesc:
        d = L.StencilDesc()
        d.kind = self.kind
        d.npoints = self.npoints
        flat = [v for o in self.offsets for v in o]
        for i, v in enumerate(flat):
            d.offsets[i] = int(v)
        for i, v in enumerate(self.f32_weights()):
            d.weights[i] = float(v)
        d.coeff = float(np.float32(self.coeff))
        d.radius = int(self.radius)
        d.boundary = int(self.boundary)
        d.halo_value = float(np.float32(self.halo_value))
        return d


def jacobi2d(coeff: float = 0.25) -> Stencil:
    """out(h,w) = coeff*(((in(h-1,w) + in(h+1,w)) + in(h,w-1)) + in(h,w+1))."""
    return Stencil(SCALED_SUM, 2, list(J5_OFFSETS), coeff=coeff, coeff_f64=coeff, name="jacobi2d")


def weighted9(weights: Sequence[float]) -> Stencil:
    """Weighted 3x3 box, weights row-major over (dh, dw) in {-1,0,1}^2."""
    if len(weights) != 9:
        raise ValueError("nine weights")
    return Stencil(WEIGHTED_SUM, 2, list(BOX9_OFFSETS), weights=list(weights), name="weighted9")


def diffusion9_weights(centre=0.5, edge=0.1, corner=0.025, wind=(0.02, -0.01)) -> List[float]:
    """Weights of the 9-point diffusion/advection automaton.

    Base weights (centre, edge, corner) are skewed by the wind (wx along w,
    wy along h): w(dh,dw) = base(dh,dw) - g(dh,dw)*(dw*wx + dh*wy) with g = 1 on
    edges, 1/4 on corners, 0 at the centre.  The skew is antisymmetric, so the
    weights still sum to centre + 4 edge + 4 corner (= 1 for the defaults) and
    the first moment drifts by -sum_k w_k*o_k per sweep (downwind).
    """
    wx, wy = wind
    out = []
    for dh, dw in BOX9_OFFSETS:
        if dh == 0 and dw == 0:
            base, g = centre, 0.0
        elif dh == 0 or dw == 0:
            base, g = edge, 1.0
        else:
            base, g = corner, 0.25
        out.append(base - g * (dw * wx + dh * wy))
    return out


def diffusion9(centre=0.5, edge=0.1, corner=0.025, wind=(0.02, -0.01)) -> Stencil:
    s = weighted9(diffusion9_weights(centre, edge, corner, wind))
    s.name = "diffusion9"
    return s


def cloudsim(delta_t: float = 0.01) -> Stencil:
    """Wind advection + conduction (the reference's CLOUDSIM_KERNEL formula,
    include/PSkelStencil.hpp:349-387) with wind_x / wind_y field arguments in
    grid field slots 0 and 1; coeff carries deltaT."""
    return Stencil(CLOUDSIM, 2, list(J5_OFFSETS), coeff=delta_t, name="cloudsim")


def life() -> Stencil:
    """Game of Life B3/S23 over the 8 Moore neighbours (the reference's GOL_KERNEL)."""
    return Stencil(LIFE, 2, list(MOORE8_OFFSETS), name="life")


def jacobi3d(coeff: float | None = None) -> Stencil:
    """7-point Jacobi heat sweep (r = 1/6): out = c*(((((hm+hp)+wm)+wp)+dm)+dp);
    cells on the faces keep their value (boundary COPY)."""
    c64 = 1.0 / 6.0 if coeff is None else float(coeff)
    return Stencil(SCALED_SUM, 3, list(J7_OFFSETS), coeff=float(np.float32(c64)), coeff_f64=c64,
                   boundary=BOUNDARY_COPY, name="jacobi3d")
