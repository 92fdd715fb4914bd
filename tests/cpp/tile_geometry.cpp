// StencilTiling::tile (include/PSkelStencilTiling.h) for the geometry test
// (tests/test_tiling_cpu.py).  Reads lines
//     W H D range iterations wOff hOff dOff w h d
// from stdin and prints, per line,
//     widthOffset width coreWidthOffset coreWidth  (the same for height and depth)
// for an Array3D(W, H, D) and a Mask3D whose only point sits `range` cells away.
#include <cstdio>

#include "PSkel.h"

using namespace PSkel;

int main()
{
    size_t W, H, D, r, k, wo, ho, dof, w, h, d;
    while (std::scanf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu", &W, &H, &D, &r, &k, &wo, &ho, &dof, &w, &h,
                      &d) == 11) {
        Array3D<float> in(W, H, D), out(W, H, D);
        Mask3D<float> mask(1);
        mask.set(0, (int)r, 0, 0);
        StencilTiling<Array3D<float>, Mask3D<float>> t(in, out, mask);
        t.tile(k, wo, ho, dof, w, h, d);
        std::printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", t.widthOffset, t.width, t.coreWidthOffset,
                    t.coreWidth, t.heightOffset, t.height, t.coreHeightOffset, t.coreHeight, t.depthOffset, t.depth,
                    t.coreDepthOffset, t.coreDepth);
        in.hostFree();
        out.hostFree();
        mask.hostFree();
    }
    return 0;
}
