//-------------------------------------------------------------------------------
// PSkelKernels.h -- stock stencil kernels and their device descriptors.
//
// Each Args type below selects a stencilKernel overload (the reference's
// mechanism for user kernels, include/PSkelStencil.h:121-156) and specialises
// StencilTraits, so the GPU methods run it on the library's gfx950 kernels.
// The CPU overloads evaluate exactly the expression the descriptor states, in
// mask order, so runIterativeCPU and runIterativeGPU agree bit for bit when the
// TU is built without floating-point contraction (-ffp-contract=off).
//
//   ScaledSumArgs{coeff, boundary}   out = coeff * (((p0 + p1) + p2) + ...)
//                                    (Jacobi: the mask lists the 4 / 6 neighbours)
//   WeightedSumArgs{boundary}        out = ((w0*p0 + w1*p1) + w2*p2) + ...
//                                    (w_i = mask.getWeight(i))
//   LifeArgs{}                       n = p0 + ... ; out = n == 3 || (n == 2 && c == 1)
//                                    (B3/S23, the reference's GOL_KERNEL rule,
//                                    include/PSkelStencil.hpp:389-440)
// 3D boundary rule: BOUNDARY_COPY keeps cells within the mask range of a face;
// BOUNDARY_HALO updates every cell and reads Mask3D's halo value outside
// (include/PSkelMask.hpp:186-199).
//-------------------------------------------------------------------------------
#ifndef PSKEL_KERNELS_H
#define PSKEL_KERNELS_H

#include "PSkelArray.h"
#include "PSkelDefs.h"
#include "PSkelMask.h"
#include "PSkelArgs.h"
#include "PSkelStencil.h"

namespace PSkel {

enum : int { BOUNDARY_COPY = 0, BOUNDARY_HALO = 1 };

struct ScaledSumArgs {
    double coeff;
    int boundary;
};

struct WeightedSumArgs {
    int boundary;
};

struct LifeArgs {
    int unused;
};

// Wind advection + conduction (the reference's CLOUDSIM_KERNEL formula,
// include/PSkelStencil.hpp:349-387) with the wind as field arguments, read as
// wind_x(h, w) / wind_y(h, w) like the reference's kernel reads them.
struct CloudsimArgs {
    Args2D<float> wind_x, wind_y;
    float deltaT;
};

// ---- 2D -----------------------------------------------------------------------
template <typename T, typename TM>
__parallel__ void stencilKernel(Array2D<T> input, Array2D<T> output, Mask2D<TM> mask, ScaledSumArgs args, size_t h,
                                size_t w)
{
    T acc = mask.get(0, input, h, w);
    for (size_t i = 1; i < mask.size; ++i) acc = acc + mask.get(i, input, h, w);
    output(h, w) = (T)args.coeff * acc;
}

template <typename T, typename TM>
__parallel__ void stencilKernel(Array2D<T> input, Array2D<T> output, Mask2D<TM> mask, WeightedSumArgs, size_t h,
                                size_t w)
{
    T acc = (T)mask.getWeight(0) * mask.get(0, input, h, w);
    for (size_t i = 1; i < mask.size; ++i) acc = acc + (T)mask.getWeight(i) * mask.get(i, input, h, w);
    output(h, w) = acc;
}

template <typename T, typename TM>
__parallel__ void stencilKernel(Array2D<T> input, Array2D<T> output, Mask2D<TM> mask, LifeArgs, size_t h, size_t w)
{
    int n = 0;
    for (size_t i = 0; i < mask.size; ++i) n += (int)mask.get(i, input, h, w);
    output(h, w) = (T)((n == 3 || (n == 2 && input(h, w) == 1)) ? 1 : 0);
}

template <typename TM>
__parallel__ void stencilKernel(Array2D<float> input, Array2D<float> output, Mask2D<TM> mask, CloudsimArgs args,
                                size_t h, size_t w)
{
    (void)mask;
    const float c = input(h, w);
    const float n = input(h - 1, w);
    const float e = input(h, w + 1);
    const float ww = input(h, w - 1);
    const float s = input(h + 1, w);
    const float sum = 4 * c - (n + e + ww + s);
    const float xwind = args.wind_x((int)h, (int)w);
    const float ywind = args.wind_y((int)h, (int)w);
    const int xfactor = (xwind > 0) ? 1 : -1;
    const int yfactor = (ywind > 0) ? 1 : -1;
    const float comp_y = yfactor * ywind;
    const float comp_x = xfactor * xwind;
    const float nb_x = input(h, (size_t)((long)w + xfactor));
    const float nb_y = input((size_t)((long)h + yfactor), w);
    const float temp_wind = (-comp_x * ((c - nb_x) * 10.0f)) - (comp_y * ((c - nb_y) * 10.0f));
    const float conduction = -0.0243f * (sum * 0.25f) * args.deltaT;
    const float result = c + conduction;
    output(h, w) = result + temp_wind * args.deltaT;
}

// ---- 3D -----------------------------------------------------------------------
template <typename T>
__parallel__ bool onFace(const Array3D<T>& a, size_t r, size_t h, size_t w, size_t d)
{
    return h < r || h + r >= a.getHeight() || w < r || w + r >= a.getWidth() || d < r || d + r >= a.getDepth();
}

template <typename T, typename TM>
__parallel__ void stencilKernel(Array3D<T> input, Array3D<T> output, Mask3D<TM> mask, ScaledSumArgs args, size_t h,
                                size_t w, size_t d)
{
    if (args.boundary == BOUNDARY_COPY && onFace(input, mask.range, h, w, d)) {
        output(h, w, d) = input(h, w, d);
        return;
    }
    T acc = (T)mask.get(0, input, h, w, d);
    for (size_t i = 1; i < mask.size; ++i) acc = acc + (T)mask.get(i, input, h, w, d);
    output(h, w, d) = (T)args.coeff * acc;
}

template <typename T, typename TM>
__parallel__ void stencilKernel(Array3D<T> input, Array3D<T> output, Mask3D<TM> mask, WeightedSumArgs args,
                                size_t h, size_t w, size_t d)
{
    if (args.boundary == BOUNDARY_COPY && onFace(input, mask.range, h, w, d)) {
        output(h, w, d) = input(h, w, d);
        return;
    }
    T acc = (T)mask.getWeight(0) * (T)mask.get(0, input, h, w, d);
    for (size_t i = 1; i < mask.size; ++i) acc = acc + (T)mask.getWeight(i) * (T)mask.get(i, input, h, w, d);
    output(h, w, d) = acc;
}

#ifdef PSKEL_GPU
namespace detail {
template <class Mask>
pskel_stencil_desc_t describeMask(Mask& mask, int kind, int ndim)
{
    pskel_stencil_desc_t d{};
    d.kind = kind;
    d.npoints = (int32_t)mask.size;
    for (size_t n = 0; n < mask.size; ++n) {
        for (int k = 0; k < ndim; ++k) d.offsets[ndim * n + k] = mask.hostOffset(n, (size_t)k);
        d.weights[n] = (float)mask.getWeight(n);
    }
    d.radius = (int32_t)mask.getRange();
    d.halo_value = (float)mask.getHaloValue();
    return d;
}
}  // namespace detail

template <>
struct StencilTraits<ScaledSumArgs> {
    static constexpr bool has_descriptor = true;
    template <class Mask>
    static pskel_stencil_desc_t describe(Mask& mask, const ScaledSumArgs& a, int ndim)
    {
        pskel_stencil_desc_t d = detail::describeMask(mask, PSKEL_KIND_SCALED_SUM, ndim);
        d.coeff = (float)a.coeff;
        d.boundary = a.boundary;
        return d;
    }
};

template <>
struct StencilTraits<WeightedSumArgs> {
    static constexpr bool has_descriptor = true;
    template <class Mask>
    static pskel_stencil_desc_t describe(Mask& mask, const WeightedSumArgs& a, int ndim)
    {
        pskel_stencil_desc_t d = detail::describeMask(mask, PSKEL_KIND_WEIGHTED_SUM, ndim);
        d.boundary = a.boundary;
        return d;
    }
};

template <>
struct StencilTraits<CloudsimArgs> {
    static constexpr bool has_descriptor = true;
    static constexpr int num_fields = 2;  // grid field slots 0, 1 = wind_x, wind_y
    template <class Mask>
    static pskel_stencil_desc_t describe(Mask& mask, const CloudsimArgs& a, int ndim)
    {
        pskel_stencil_desc_t d = detail::describeMask(mask, PSKEL_KIND_CLOUDSIM, ndim);
        d.coeff = a.deltaT;
        return d;
    }
    static void fields(const CloudsimArgs& a, pskel_field_t** out)
    {
        out[0] = a.wind_x.deviceField();
        out[1] = a.wind_y.deviceField();
    }
};

template <>
struct StencilTraits<LifeArgs> {
    static constexpr bool has_descriptor = true;
    template <class Mask>
    static pskel_stencil_desc_t describe(Mask& mask, const LifeArgs&, int ndim)
    {
        return detail::describeMask(mask, PSKEL_KIND_LIFE, ndim);
    }
};
#endif

// ---- stock masks in the point orders the tuned kernels accumulate in ------------
// von Neumann 4-neighbourhood (-1,0) (1,0) (0,-1) (0,1): Jacobi 5-point
template <typename T>
Mask2D<T> jacobiMask2D()
{
    Mask2D<T> m(4);
    m.set(0, -1, 0);
    m.set(1, 1, 0);
    m.set(2, 0, -1);
    m.set(3, 0, 1);
    return m;
}

// 3x3 box, row-major over (dh, dw), with the given weights
template <typename T>
Mask2D<T> boxMask2D(const T (&weights)[9])
{
    Mask2D<T> m(9);
    for (int i = 0; i < 9; ++i) m.set(i, i / 3 - 1, i % 3 - 1, weights[i]);
    return m;
}

// the 8 Moore neighbours, row-major
template <typename T>
Mask2D<T> mooreMask2D()
{
    Mask2D<T> m(8);
    int n = 0;
    for (int dh = -1; dh <= 1; ++dh)
        for (int dw = -1; dw <= 1; ++dw)
            if (dh || dw) m.set(n++, dh, dw, T(1));
    return m;
}

// the 6 face neighbours (-1,0,0) (1,0,0) (0,-1,0) (0,1,0) (0,0,-1) (0,0,1): Jacobi 7-point
template <typename T>
Mask3D<T> jacobiMask3D()
{
    Mask3D<T> m(6);
    m.set(0, -1, 0, 0);
    m.set(1, 1, 0, 0);
    m.set(2, 0, -1, 0);
    m.set(3, 0, 1, 0);
    m.set(4, 0, 0, -1);
    m.set(5, 0, 0, 1);
    return m;
}

}  // namespace PSkel

#endif
