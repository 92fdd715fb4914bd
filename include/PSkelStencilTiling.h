//-------------------------------------------------------------------------------
// PSkelStencilTiling.h -- tile geometry of temporal tiling (drop-in for the
// reference's include/PSkelStencilTiling.h; paths relative to the PSkel tree).
//
// tile(iterations, offsets, sizes) expands a logical tile by a halo of
// range * iterations cells on every side, clamped to the array, and records the
// core (the logical tile) inside the expanded tile (:85-148).  The library's
// gfx950 temporal-tile kernels use the same geometry on LDS tiles (the halo
// cells are recomputed redundantly each level).  Fixed on purpose: the
// reference zeroes the offset before computing the halo clamp (:108-136), so
// halos come out short when 0 < offset < range * iterations; here the halo is
// min(offset, range * iterations).
//-------------------------------------------------------------------------------
#ifndef PSKEL_STENCIL_TILING_H
#define PSKEL_STENCIL_TILING_H

#include <cstddef>

namespace PSkel {

template <class Array, class Mask>
class StencilTiling {
public:
    Array input;
    Array output;
    Mask mask;
    size_t iterations = 0;
    size_t widthOffset = 0, heightOffset = 0, depthOffset = 0;
    size_t width = 0, height = 0, depth = 0;
    size_t coreWidthOffset = 0, coreHeightOffset = 0, coreDepthOffset = 0;
    size_t coreWidth = 0, coreHeight = 0, coreDepth = 0;

    StencilTiling(Array in, Array out, Mask m) : input(in), output(out), mask(m) {}

    void tile(size_t iters, size_t wOff, size_t hOff, size_t dOff, size_t w, size_t h, size_t d)
    {
        iterations = iters;
        clampAxis(wOff, w, input.getWidth(), widthOffset, width, coreWidthOffset, coreWidth);
        clampAxis(hOff, h, input.getHeight(), heightOffset, height, coreHeightOffset, coreHeight);
        clampAxis(dOff, d, input.getDepth(), depthOffset, depth, coreDepthOffset, coreDepth);
    }

private:
    void clampAxis(size_t off, size_t len, size_t n, size_t& tOff, size_t& tLen, size_t& cOff, size_t& cLen)
    {
        if (off >= n) {
            tOff = off;
            tLen = 0;
            cOff = 0;
            cLen = 0;
            return;
        }
        if (off + len > n) len = n - off;
        const size_t halo = mask.getRange() * iterations;
        const size_t lo = off >= halo ? off - halo : 0;
        size_t hi = off + len + halo;
        if (hi > n) hi = n;
        tOff = lo;
        tLen = hi - lo;
        cOff = off - lo;
        cLen = len;
    }
};

}  // namespace PSkel

#endif
