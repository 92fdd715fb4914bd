//-------------------------------------------------------------------------------
// PSkelMap.h -- the map skeleton (drop-in for the reference's include/PSkelMap.h
// and include/PSkelMap.hpp; paths relative to the PSkel tree).
//
// The user writes mapKernel overloads in namespace PSkel with the reference's
// signatures (include/PSkelMap.h:53-59):
//   mapKernel(Array<T> in, Array<T> out, Args args, size_t i)
//   mapKernel(Array2D<T> in, Array2D<T> out, Args args, size_t h, size_t w)
//   mapKernel(Array3D<T> in, Array3D<T> out, Args args, size_t h, size_t w, size_t d)
// and runs them with Map / Map2D / Map3D over every cell.
//
// CPU methods restate the reference's loops (include/PSkelMap.hpp:120-350):
// the iterative ones ping-pong between a clone of `input` and `output`, so
// `input` is left as it was and the result is in `output` (:185-218).  GPU
// methods keep the arrays resident in HBM for all iterations (a grid of the
// library), run the user's mapKernel one lane per cell with lanes on the
// contiguous axis, and copy the result back once; the reference copies the
// input in and the result out around the loop the same way (:222-243).  The
// user's mapKernel is compiled for the device only in a TU built by hipcc;
// elsewhere the GPU methods fail with PSKEL_ERR_UNSUPPORTED.
//-------------------------------------------------------------------------------
#ifndef PSKEL_MAP_H
#define PSKEL_MAP_H

#include <cstddef>
#include <type_traits>
#include <utility>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#endif

#include "PSkelArgs.h"
#include "PSkelArray.h"
#include "PSkelDefs.h"

namespace PSkel {

namespace detail {

inline size_t mapThreads(size_t n)
{
#if defined(_OPENMP)
    return n == 0 ? (size_t)omp_get_num_procs() : n;
#else
    return n == 0 ? 1 : n;
#endif
}

#if defined(PSKEL_GPU) && defined(PSKEL_DEVICE_COMPILER)
template <class Arr, class Args>
__global__ void mapKernel1D(Arr in, Arr out, Args args, size_t W)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < W) mapKernel(in, out, args, i);
}

template <class Arr, class Args>
__global__ void mapKernel2D(Arr in, Arr out, Args args, size_t W, size_t H)
{
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t h = (size_t)blockIdx.y * blockDim.y + threadIdx.y;
    if (w < W && h < H) mapKernel(in, out, args, h, w);
}

template <class Arr, class Args>
__global__ void mapKernel3D(Arr in, Arr out, Args args, size_t W, size_t H, size_t D)
{
    const size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t w = (size_t)blockIdx.y * blockDim.y + threadIdx.y;
    const size_t h = blockIdx.z;
    if (d < D && w < W && h < H) mapKernel(in, out, args, h, w, d);
}
#endif

}  // namespace detail

template <class Arrays, class Args = int>
class MapBase {
protected:
    Arrays input;
    Arrays output;
    Args args;

    double lastMs = 0.0;

    virtual void runSeq(Arrays in, Arrays out) = 0;
    virtual void runOpenMP(Arrays in, Arrays out, size_t numThreads) = 0;

public:
    virtual ~MapBase() {}
    // device time of the last GPU run's sweeps (an extension, as on Stencil)
    double lastGpuMillis() const { return lastMs; }

    void runSequential() { runSeq(input, output); }
    void runCPU(size_t numThreads = 0) { runOpenMP(input, output, detail::mapThreads(numThreads)); }

    void runIterativeSequential(size_t iterations)
    {
        Arrays inputCopy;
        inputCopy.hostClone(input);
        for (size_t it = 0; it < iterations; it++) {
            if (it % 2 == 0) runSeq(inputCopy, output);
            else runSeq(output, inputCopy);
        }
        if ((iterations % 2) == 0) output.hostMemCopy(inputCopy);
        inputCopy.hostFree();
    }

    void runIterativeCPU(size_t iterations, size_t numThreads = 0)
    {
        const size_t th = detail::mapThreads(numThreads);
        Arrays inputCopy;
        inputCopy.hostClone(input);
        for (size_t it = 0; it < iterations; it++) {
            if (it % 2 == 0) runOpenMP(inputCopy, output, th);
            else runOpenMP(output, inputCopy, th);
        }
        if ((iterations % 2) == 0) output.hostMemCopy(inputCopy);
        inputCopy.hostFree();
    }

#ifdef PSKEL_GPU
    void runGPU(size_t blockSize = 0) { runIterativeGPU(1, blockSize); }

    // blockSize: threads along the contiguous axis (0: 64)
    void runIterativeGPU(size_t iterations, size_t blockSize = 0)
    {
#if defined(PSKEL_DEVICE_COMPILER)
        using T = std::remove_reference_t<decltype(*std::declval<Arrays&>().hostArray)>;
        constexpr int nd = Arrays::kDims;
        // f32 arrays (u8 grids of the library hold life cells, 0 or 1)
        const int hdt = std::is_same_v<T, float> ? PSKEL_DTYPE_F32 : 0;
        if (hdt == 0) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
        detail::syncFields();
        pskel_grid_desc_t gd{};
        gd.ndim = nd == 3 ? 3 : 2;
        gd.storage = hdt;
        gd.width = input.getWidth();
        gd.height = nd == 1 ? 1 : input.getHeight();
        gd.depth = nd == 3 ? input.getDepth() : 1;
        pskel_grid_t* g = nullptr;
        gpuErrchk(pskel_grid_create(&gd, &g));
        std::vector<T> buf(input.size());
        input.packTo(buf.data());
        gpuErrchk(pskel_grid_upload(g, PSKEL_BUFFER_INPUT, buf.data(), hdt));
        gpuErrchk(pskel_grid_upload(g, PSKEL_BUFFER_OUTPUT, buf.data(), hdt));
        pskel_grid_info_t gi{};
        gpuErrchk(pskel_grid_info(g, &gi));
        const unsigned bx = blockSize > 0 ? (unsigned)blockSize : 64u;
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, (hipStream_t)gi.stream);
        for (size_t it = 0; it < iterations; ++it) {
            gpuErrchk(pskel_grid_info(g, &gi));
            hipStream_t stream = (hipStream_t)gi.stream;
            Arrays in = input, out = output;
            in.setDeviceView((T*)gi.current, gi.pitch);
            out.setDeviceView((T*)gi.previous, gi.pitch);
            if constexpr (nd == 1) {
                hipLaunchKernelGGL((detail::mapKernel1D<Arrays, Args>), dim3((unsigned)((gi.width + bx - 1) / bx)),
                                   dim3(bx), 0, stream, in, out, args, gi.width);
            } else if constexpr (nd == 2) {
                const dim3 grid((unsigned)((gi.width + bx - 1) / bx), (unsigned)((gi.height + 3) / 4));
                hipLaunchKernelGGL((detail::mapKernel2D<Arrays, Args>), grid, dim3(bx, 4), 0, stream, in, out, args,
                                   gi.width, gi.height);
            } else {
                const dim3 grid((unsigned)((gi.depth + bx - 1) / bx), (unsigned)((gi.width + 3) / 4),
                                (unsigned)gi.height);
                hipLaunchKernelGGL((detail::mapKernel3D<Arrays, Args>), grid, dim3(bx, 4), 0, stream, in, out, args,
                                   gi.width, gi.height, gi.depth);
            }
            if (hipGetLastError() != hipSuccess) {
                pskel_grid_destroy(g);
                gpuErrchk((pskel_status_t)PSKEL_ERR_LAUNCH);
            }
            gpuErrchk(pskel_grid_swap(g));
        }
        (void)hipEventRecord(e1, (hipStream_t)gi.stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        lastMs = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        gpuErrchk(pskel_grid_download(g, buf.data(), hdt));
        output.unpackFrom(buf.data());
        pskel_grid_destroy(g);
#else
        (void)iterations;
        (void)blockSize;
        gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);  // compile the TU with hipcc
#endif
    }
#endif
};

template <class Arrays, class Args>
class Map3D : public MapBase<Arrays, Args> {
protected:
    void runSeq(Arrays in, Arrays out) override
    {
        for (size_t h = 0; h < in.getHeight(); h++)
            for (size_t w = 0; w < in.getWidth(); w++)
                for (size_t d = 0; d < in.getDepth(); d++) mapKernel(in, out, this->args, h, w, d);
    }
    void runOpenMP(Arrays in, Arrays out, size_t numThreads) override
    {
        (void)numThreads;
        const long long H = (long long)in.getHeight();
#pragma omp parallel for num_threads(numThreads) schedule(static)
        for (long long h = 0; h < H; h++)
            for (size_t w = 0; w < in.getWidth(); w++)
                for (size_t d = 0; d < in.getDepth(); d++) mapKernel(in, out, this->args, (size_t)h, w, d);
    }

public:
    Map3D() {}
    Map3D(Arrays _input, Arrays _output, Args _args)
    {
        this->input = _input;
        this->output = _output;
        this->args = _args;
    }
};

template <class Arrays, class Args>
class Map2D : public MapBase<Arrays, Args> {
protected:
    void runSeq(Arrays in, Arrays out) override
    {
        for (size_t h = 0; h < in.getHeight(); h++)
            for (size_t w = 0; w < in.getWidth(); w++) mapKernel(in, out, this->args, h, w);
    }
    void runOpenMP(Arrays in, Arrays out, size_t numThreads) override
    {
        (void)numThreads;
        const long long H = (long long)in.getHeight();
#pragma omp parallel for num_threads(numThreads) schedule(static)
        for (long long h = 0; h < H; h++)
            for (size_t w = 0; w < in.getWidth(); w++) mapKernel(in, out, this->args, (size_t)h, w);
    }

public:
    Map2D() {}
    Map2D(Arrays _input, Arrays _output, Args _args)
    {
        this->input = _input;
        this->output = _output;
        this->args = _args;
    }
};

template <class Arrays, class Args>
class Map : public MapBase<Arrays, Args> {
protected:
    void runSeq(Arrays in, Arrays out) override
    {
        for (size_t i = 0; i < in.getWidth(); i++) mapKernel(in, out, this->args, i);
    }
    void runOpenMP(Arrays in, Arrays out, size_t numThreads) override
    {
        (void)numThreads;
        const long long W = (long long)in.getWidth();
#pragma omp parallel for num_threads(numThreads) schedule(static)
        for (long long i = 0; i < W; i++) mapKernel(in, out, this->args, (size_t)i);
    }

public:
    Map() {}
    Map(Arrays _input, Arrays _output, Args _args)
    {
        this->input = _input;
        this->output = _output;
        this->args = _args;
    }
};

}  // namespace PSkel

#endif
