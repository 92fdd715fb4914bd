//-------------------------------------------------------------------------------
// PSkelStencil.h -- the stencil skeleton (drop-in for the reference's
// include/PSkelStencil.h and include/PSkelStencil.hpp; paths relative to the
// PSkel tree).
//
// The user writes stencilKernel overloads in namespace PSkel with the
// reference's signatures (include/PSkelStencil.h:121-156):
//   stencilKernel(Array2D<T> in, Array2D<T> out, Mask2D<TM> mask, Args args, size_t h, size_t w)
//   stencilKernel(Array3D<T> in, Array3D<T> out, Mask3D<TM> mask, Args args, size_t h, size_t w, size_t d)
//   stencilKernel(Array<T> in, Array<T> out, Mask<TM> mask, Args args, size_t i)
// and runs them with Stencil2D / Stencil3D / Stencil.
//
// CPU methods restate the reference's loops (include/PSkelStencil.hpp:864-938,
// 2095-2113, 2184-2257).  GPU methods run on the MI355X through libpskel_amd.so:
//   * when StencilTraits<Args> describes the kernel (the stock kernels of
//     PSkelKernels.h do, and a user can specialise it for their own Args),
//     the library's gfx950 kernels run it: plain, tuned or temporal tiles;
//   * otherwise, in a TU compiled by hipcc, the user's stencilKernel is
//     compiled for the device and launched one lane per cell on the same
//     resident buffers (the reference's stencilTilingCU contract,
//     include/PSkelStencil.hpp:572-652);
//   * otherwise the GPU method fails with PSKEL_ERR_UNSUPPORTED.
// Semantics kept: 2D sweeps update the interior [r, N-r) only, 3D sweeps every
// cell; the result of N sweeps is in `output`; runIterativeCPU clobbers
// `input`, the GPU methods leave it untouched (include/PSkelStencil.hpp:880-938,
// 1047-1097).  Fixed on purpose: numThreads == 0 means all processors (the
// reference passes num_threads(0), :882); block sizes of 0 mean 64 x 4 (the
// reference leaves by = 0, :1048-1054); the device output is initialised from
// the host output (the reference never uploads it, :1076).
//-------------------------------------------------------------------------------
#ifndef PSKEL_STENCIL_H
#define PSKEL_STENCIL_H

#include <algorithm>
#include <chrono>
#include <cmath>
#include <exception>
#include <cstdint>
#include <iostream>
#include <memory>
#include <type_traits>
#include <thread>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#endif

#include "PSkelArray.h"
#include "PSkelArgs.h"
#include "PSkelDefs.h"
#include "PSkelMask.h"

namespace PSkel {

// ---------------------------------------------------------------------------
// Kernel descriptors.  Specialise StencilTraits<Args> with
//   static constexpr bool has_descriptor = true;
//   template <class Mask> static pskel_stencil_desc_t describe(Mask& m, const Args& a, int ndim);
// to run a stencilKernel on the library's tuned kernels.  describe() must state
// exactly what the kernel computes (see PSkelKernels.h).
// ---------------------------------------------------------------------------
template <class Args>
struct StencilTraits {
    static constexpr bool has_descriptor = false;
};

namespace detail {

// StencilTraits that name field arguments (num_fields, fields(args, out))
template <class Tr, class = void>
struct hasFields : std::false_type {};
template <class Tr>
struct hasFields<Tr, std::void_t<decltype(Tr::num_fields)>> : std::true_type {};

inline size_t cpuThreads(size_t n)
{
#if defined(_OPENMP)
    return n == 0 ? (size_t)omp_get_num_procs() : n;
#else
    return n == 0 ? 1 : n;
#endif
}

#ifdef PSKEL_GPU
template <typename T>
constexpr int hostDtype()
{
    if constexpr (std::is_same_v<T, float>) return PSKEL_DTYPE_F32;
    else if constexpr (std::is_same_v<T, double>) return PSKEL_DTYPE_F64;
    else if constexpr (std::is_same_v<T, unsigned char> || std::is_same_v<T, char> ||
                       std::is_same_v<T, signed char> || std::is_same_v<T, bool>)
        return PSKEL_DTYPE_U8;
    else return 0;
}

// Resident device state of a Stencil: the grid pair and the run options.
struct GpuState {
    pskel_grid_t* grid = nullptr;
    int device = 0;
    int storage = 0;  // 0: the host element type
    float storeOffset = 0.0f;
    int rounding = PSKEL_ROUND_NEAREST;
    uint32_t seed = 0;
    size_t dims[3] = {0, 0, 0};
    int gridStorage = 0;
    int gridRounding = 0;
    uint32_t gridSeed = 0;
    float gridOffset = 0.0f;
    size_t residualEvery = 0;
    std::vector<double> residuals;
    bool haveCfg = false;
    bool hostTiling = false;
    pskel_launch_cfg_t cfg{};
    double lastMs = 0.0;
    ~GpuState()
    {
        if (grid) pskel_grid_destroy(grid);
    }
};

#if defined(PSKEL_DEVICE_COMPILER)
// One lane per interior cell: the reference's stencilTilingCU
// (include/PSkelStencil.hpp:572-652), lanes on the contiguous axis.
template <class Arr, class Msk, class Args>
__global__ void genericKernel2D(Arr in, Arr out, Msk mask, Args args, size_t r, size_t W, size_t H)
{
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t h = (size_t)blockIdx.y * blockDim.y + threadIdx.y;
    if (w >= r && w + r < W && h >= r && h + r < H) stencilKernel(in, out, mask, args, h, w);
}

// Every cell (Stencil3D visits all cells, include/PSkelStencil.hpp:2095-2113);
// lanes on d, the contiguous axis (the reference maps them to w, :687).
template <class Arr, class Msk, class Args>
__global__ void genericKernel3D(Arr in, Arr out, Msk mask, Args args, size_t W, size_t H, size_t D)
{
    const size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t w = (size_t)blockIdx.y * blockDim.y + threadIdx.y;
    const size_t h = blockIdx.z;
    if (d < D && w < W && h < H) stencilKernel(in, out, mask, args, h, w, d);
}
#endif
#endif  // PSKEL_GPU

}  // namespace detail

//*******************************************************************************************
// Stencil Base
//*******************************************************************************************
template <class Array, class Mask, class Args = int>
class StencilBase {
protected:
    Array input;
    Array output;
    Args args;
    Mask mask;
#ifdef PSKEL_GPU
    std::shared_ptr<detail::GpuState> gpu = std::make_shared<detail::GpuState>();
#endif

    // a stencil of the same kind, mask and args on other arrays
    virtual std::unique_ptr<StencilBase> cloneWith(Array in, Array out) const = 0;
    virtual void runSeq(Array in, Array out, size_t width, size_t height, size_t maskRange) = 0;
    virtual void runOpenMP(Array in, Array out, size_t width, size_t height, size_t depth, size_t maskRange,
                           size_t numThreads) = 0;
    virtual int dims() const = 0;

public:
    virtual ~StencilBase() = default;

    // ---- CPU (include/PSkelStencil.hpp:708-727, 864-938) ------------------
    void runSequential()
    {
        runSeq(input, output, input.getWidth(), input.getHeight(), mask.getRange());
    }

    void runCPU(size_t numThreads = 0)
    {
        runOpenMP(input, output, input.getWidth(), input.getHeight(), input.getDepth(), mask.getRange(),
                  detail::cpuThreads(numThreads));
    }

    void runIterativeSequential(size_t iterations)
    {
        const size_t width = input.getWidth(), height = input.getHeight(), maskRange = mask.getRange();
        for (size_t it = 0; it < iterations; it++) {
            if (it % 2 == 0) runSeq(input, output, width, height, maskRange);
            else runSeq(output, input, width, height, maskRange);
        }
        if ((iterations % 2) == 0) output.hostMemCopy(input);
    }

    void runIterativeCPU(size_t iterations, size_t numThreads = 0)
    {
        const size_t width = input.getWidth(), height = input.getHeight(), depth = input.getDepth();
        const size_t maskRange = mask.getRange();
        const size_t th = detail::cpuThreads(numThreads);
        const auto t0 = std::chrono::steady_clock::now();
        for (size_t it = 0; it < iterations; it++) {
            if (it % 2 == 0) runOpenMP(input, output, width, height, depth, maskRange, th);
            else runOpenMP(output, input, width, height, depth, maskRange, th);
        }
        if ((iterations % 2) == 0) output.hostMemCopy(input);
#ifdef PSKEL_PRINT_TIMES
        std::cout << "CPU_time\t"
                  << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << std::endl;
#else
        (void)t0;
#endif
    }

#ifdef PSKEL_GPU
    // ---- GPU (include/PSkelStencil.hpp:730-861, 943-1097, 1663-2076) --------
    // One sweep (runGPU, :730-762).
    void runGPU(size_t GPUBlockSizeX = 0, size_t GPUBlockSizeY = 0)
    {
        runIterativeGPU(1, GPUBlockSizeX, GPUBlockSizeY);
    }

    // N sweeps on the resident grid (:1047-1097).  "If GPUBlockSize is 0, the
    // block size is automatically chosen" (include/PSkelStencil.h:252-256): with
    // both block sizes 0 the launch comes from setLaunchConfig, else from the
    // tuner (tuned kernels, temporal tiles), as runIterativeAutoGPU picks it for
    // a resident grid.  Explicit block sizes keep the plain contract: one lane
    // per cell in (bx, by) blocks (:1765-1792).
    void runIterativeGPU(size_t iterations, size_t GPUBlockSizeX = 0, size_t GPUBlockSizeY = 0)
    {
        pskel_launch_cfg_t cfg{};
        if (GPUBlockSizeX == 0 && GPUBlockSizeY == 0) {
            cfg.variant = PSKEL_VARIANT_TUNED;
            gpuRun(iterations, cfg, false, /*autotune=*/true);
            return;
        }
        cfg.variant = PSKEL_VARIANT_PLAIN;
        cfg.block_x = (int)(GPUBlockSizeX ? GPUBlockSizeX : 64);
        cfg.block_y = (int)(GPUBlockSizeY ? GPUBlockSizeY : 1);
        gpuRun(iterations, cfg, true);
    }

    // The reference's shared-memory pyramid (:943-1041): pyramidHeight sweeps
    // fused per launch -- here overlapped temporal tiles of that depth: up to 4
    // in 2D (8 for life), 2 in 3D (the fused 7-point pair; other 3D stencils
    // sweep singly).
    void runIterativeGPU(size_t iterations, size_t pyramidHeight, size_t GPUBlockSizeX, size_t GPUBlockSizeY)
    {
        (void)GPUBlockSizeX;
        (void)GPUBlockSizeY;
        pskel_launch_cfg_t cfg{};
        cfg.variant = PSKEL_VARIANT_TUNED;
        cfg.time_tile = (int)std::max<size_t>(1, std::min<size_t>(pyramidHeight, maxTemporalDepth()));
        gpuRun(iterations, cfg, false);
    }

    // Tiled run (:1663-1760).  A grid that fits in device memory stays resident
    // and `innerIterations` becomes the temporal-tile depth (same result).  A
    // grid that does not fit (or any grid after setHostTiling(true)) runs the
    // reference's out-of-core scheme: slabs of `tilingHeight` rows (2D) or
    // h-planes (3D) plus a range*inner halo visit the device, inner sweeps at a
    // time, and the grid stays in host memory.
    void runIterativeTilingGPU(size_t iterations, size_t tilingWidth, size_t tilingHeight, size_t tilingDepth,
                               size_t innerIterations = 1, size_t GPUBlockSizeX = 0, size_t GPUBlockSizeY = 0)
    {
        (void)tilingWidth;
        (void)tilingDepth;
        (void)GPUBlockSizeX;
        (void)GPUBlockSizeY;
        if (gpu->hostTiling || !fitsDevice()) {
            hostTiledRun(iterations, std::max<size_t>(1, tilingHeight), std::max<size_t>(1, innerIterations));
            return;
        }
        pskel_launch_cfg_t cfg{};
        cfg.variant = PSKEL_VARIANT_TUNED;
        cfg.time_tile = (int)std::max<size_t>(1, std::min<size_t>(innerIterations, maxTemporalDepth()));
        gpuRun(iterations, cfg, false);
    }

    void runTilingGPU(size_t tilingWidth, size_t tilingHeight, size_t tilingDepth, size_t GPUBlockSizeX = 0,
                      size_t GPUBlockSizeY = 0)
    {
        runIterativeTilingGPU(1, tilingWidth, tilingHeight, tilingDepth, 1, GPUBlockSizeX, GPUBlockSizeY);
    }

    // Auto run (:2007-2076): the launch shape and temporal depth come from the
    // library's tuner (cost model over the grid shape, mask, type and
    // iteration count), not from a GA over an analytic objective.
    void runIterativeAutoGPU(size_t iterations, size_t GPUBlockSize = 0)
    {
        (void)GPUBlockSize;
        if (gpu->hostTiling || !fitsDevice()) {
            // out of core (:2014-2074): slab height and sweeps per pass from a
            // measured search over the pipeline's candidates that fit in half
            // the free device memory (pskel_tuner_host_search)
            size_t tile = 0, inner = 0;
            hostTiledSearch(iterations, tile, inner);
            hostTiledRun(iterations, tile, inner);
            return;
        }
        pskel_launch_cfg_t cfg{};
        cfg.variant = PSKEL_VARIANT_TUNED;
        gpuRun(iterations, cfg, false, /*autotune=*/true);
    }

    void runAutoGPU(size_t GPUBlockSize = 0) { runIterativeAutoGPU(1, GPUBlockSize); }

    // CPU+GPU split (:1288-1659): the first ceil(H * gpuFactor) rows (2D) or
    // h-planes (3D) are computed on the GPU and the rest on the CPU, at the same
    // time, each part on a private copy of its rows plus a ghost zone of
    // range * iterations (:1323-1325), so each part's core is exact; the cores
    // then go to `output`.  gpuFactor 1 runs on the GPU, 0 on the CPU
    // (:1428-1431).  `input` is left as it was.
    void runIterativePartition(size_t iterations, float gpuFactor, size_t numThreads = 0,
                               size_t GPUBlockSizeX = 32, size_t GPUBlockSizeY = 4)
    {
        const size_t H = input.getHeight();
        const size_t gh = (size_t)std::ceil((double)H * (double)gpuFactor);
        if (gpuFactor >= 1.0f || gh >= H) return runIterativeGPU(iterations, GPUBlockSizeX, GPUBlockSizeY);
        if (gpuFactor <= 0.0f || gh == 0 || Array::kDims == 1) return runIterativeCPU(iterations, numThreads);
        const size_t halo = mask.getRange() * iterations;
        const size_t g1 = std::min(H, gh + halo), c0 = gh > halo ? gh - halo : 0;
        const size_t W = input.getWidth(), D = input.getDepth();
        Array sl, gin, gout, cin, cout;
        sl.hostSlice(input, 0, 0, 0, W, g1, D);
        gin.hostClone(sl);
        sl.hostSlice(output, 0, 0, 0, W, g1, D);
        gout.hostClone(sl);
        sl.hostSlice(input, 0, c0, 0, W, H - c0, D);
        cin.hostClone(sl);
        sl.hostSlice(output, 0, c0, 0, W, H - c0, D);
        cout.hostClone(sl);
        auto gpuPart = cloneWith(gin, gout);
        auto cpuPart = cloneWith(cin, cout);
        detail::GpuState& mine = *gpu;
        detail::GpuState& theirs = *gpuPart->gpu;
        theirs.device = mine.device;
        theirs.storage = mine.storage;
        theirs.storeOffset = mine.storeOffset;
        theirs.rounding = mine.rounding;
        theirs.seed = mine.seed;
        theirs.haveCfg = mine.haveCfg;
        theirs.cfg = mine.cfg;
        std::exception_ptr err;
        std::thread gpuThread([&]() {
            try {
                gpuPart->runIterativeGPU(iterations, GPUBlockSizeX, GPUBlockSizeY);
            } catch (...) {
                err = std::current_exception();
            }
        });
        cpuPart->runIterativeCPU(iterations, numThreads);
        gpuThread.join();
        if (err) std::rethrow_exception(err);
        Array core;
        sl.hostSlice(output, 0, 0, 0, W, gh, D);
        core.hostSlice(gout, 0, 0, 0, W, gh, D);
        sl.hostMemCopy(core);
        sl.hostSlice(output, 0, gh, 0, W, H - gh, D);
        core.hostSlice(cout, 0, gh - c0, 0, W, H - gh, D);
        sl.hostMemCopy(core);
        gin.hostFree();
        gout.hostFree();
        cin.hostFree();
        cout.hostFree();
    }

    // ---- MI355X extensions ------------------------------------------------
    void setDevice(int device) { gpu->device = device; }
    // Run runIterativeTilingGPU / runIterativeAutoGPU out of core (host-staged
    // slabs) even when the grid fits in device memory.
    void setHostTiling(bool on) { gpu->hostTiling = on; }
    // Device storage of the grid: PSKEL_DTYPE_BF16 / F16 keep (u - offset) in
    // 16 bits with f32 arithmetic (3D); 0 = the host element type.  The offset
    // defaults to PSKEL_STORE_OFFSET_AUTO: the mean of `input` at each run.
    // `rounding` is PSKEL_ROUND_NEAREST or PSKEL_ROUND_STOCHASTIC
    // (reproducible from `seed`).
    void setDeviceStorage(int storage, float offset = PSKEL_STORE_OFFSET_AUTO, int rounding = PSKEL_ROUND_NEAREST,
                          uint32_t seed = 0)
    {
        gpu->storage = storage;
        gpu->storeOffset = offset;
        gpu->rounding = rounding;
        gpu->seed = seed;
    }
    // FP64 residual ||u^k - u^(k-1)||_2 every n sweeps of the next GPU runs.
    void setResidualInterval(size_t n) { gpu->residualEvery = n; }
    const std::vector<double>& residuals() const { return gpu->residuals; }
    // Launch shape for the tuned methods (overrides the tuner).
    void setLaunchConfig(const pskel_launch_cfg_t& cfg)
    {
        gpu->cfg = cfg;
        gpu->haveCfg = true;
    }
    double lastGpuMillis() const { return gpu->lastMs; }

protected:
    using T = std::remove_reference_t<decltype(*std::declval<Array&>().hostArray)>;

    // Sweeps the library fuses per launch at most: in 3D 3 for the 7-point
    // Jacobi on f32 storage, 2 on 16-bit storage (other 3D stencils sweep
    // singly); 4 in 2D, 8 for life.
    size_t maxTemporalDepth()
    {
        size_t kmax = dims() == 2 ? 4 : 2;
        if constexpr (StencilTraits<Args>::has_descriptor) {
            const pskel_stencil_desc_t d = StencilTraits<Args>::describe(mask, args, dims());
            if (dims() == 2 && d.kind == PSKEL_KIND_LIFE) kmax = 8;
            const int storage = gpu->storage ? gpu->storage : detail::hostDtype<T>();
            if (dims() == 3 && d.kind == PSKEL_KIND_SCALED_SUM && d.npoints == 6 && storage == PSKEL_DTYPE_F32)
                kmax = 3;
        }
        return kmax;
    }

    void ensureGrid()
    {
        detail::GpuState& s = *gpu;
        int storage = s.storage ? s.storage : detail::hostDtype<T>();
        if (storage == PSKEL_DTYPE_F64 || storage == 0)
            gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);  // no FP64 sweeps on the device
        const size_t d0 = input.getWidth(), d1 = input.getHeight(), d2 = input.getDepth();
        const bool sameOffset = s.gridOffset == s.storeOffset || (std::isnan(s.gridOffset) && std::isnan(s.storeOffset));
        if (s.grid && (s.dims[0] != d0 || s.dims[1] != d1 || s.dims[2] != d2 || s.gridStorage != storage ||
                       s.gridRounding != s.rounding || s.gridSeed != s.seed || !sameOffset)) {
            pskel_grid_destroy(s.grid);
            s.grid = nullptr;
        }
        if (!s.grid) {
            pskel_grid_desc_t gd{};
            gd.ndim = dims();
            gd.storage = storage;
            gd.width = d0;
            gd.height = d1;
            gd.depth = dims() == 3 ? d2 : 1;
            gd.store_offset = s.storeOffset;
            gd.device = s.device;
            gd.rounding = s.rounding;
            gd.seed = s.seed;
            gpuErrchk(pskel_grid_create(&gd, &s.grid));
            s.dims[0] = d0;
            s.dims[1] = d1;
            s.dims[2] = d2;
            s.gridStorage = storage;
            s.gridRounding = s.rounding;
            s.gridSeed = s.seed;
            s.gridOffset = s.storeOffset;
        }
    }

    // Would the resident grid (two buffers plus a spare) fit in free device memory?
    bool fitsDevice() const
    {
        size_t freeb = 0, totalb = 0;
        gpuErrchk(pskel_device_info(gpu->device, nullptr, 0, &freeb, &totalb, nullptr));
        const size_t bytes = input.size() * sizeof(T);
        return 3 * bytes + (size_t(64) << 20) < freeb;
    }

    // The out-of-core launch (slab height, sweeps per pass) for this grid.
    void hostTiledSearch(size_t iterations, size_t& tile, size_t& inner)
    {
        if constexpr (Array::kDims == 1 || !StencilTraits<Args>::has_descriptor) {
            gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);  // user kernels stay resident
        } else {
            detail::GpuState& s = *gpu;
            const int hdt = detail::hostDtype<T>();
            if (hdt == 0 || (s.storage != 0 && s.storage != hdt)) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
            pskel_grid_desc_t gd{};
            gd.ndim = dims();
            gd.storage = hdt;
            gd.width = input.getWidth();
            gd.height = input.getHeight();
            gd.depth = dims() == 3 ? input.getDepth() : 1;
            gd.device = s.device;
            pskel_stencil_desc_t desc = StencilTraits<Args>::describe(mask, args, dims());
            std::vector<T> in(input.size());
            input.packTo(in.data());
            gpuErrchk(pskel_tuner_host_search(&gd, in.data(), hdt, &desc, iterations, 0, &tile, &inner, nullptr));
        }
    }

    // Out-of-core sweeps through the library (descriptor kernels only).
    void hostTiledRun(size_t iterations, size_t tile, size_t inner)
    {
        if constexpr (Array::kDims == 1 || !StencilTraits<Args>::has_descriptor) {
            gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);  // user kernels stay resident
        } else {
            detail::GpuState& s = *gpu;
            s.residuals.clear();
            // host-staged passes reduce no residuals: refuse rather than
            // return none for an interval that was asked for
            if (s.residualEvery) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
            const int hdt = detail::hostDtype<T>();
            if (hdt == 0 || (s.storage != 0 && s.storage != hdt)) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
            pskel_grid_desc_t gd{};
            gd.ndim = dims();
            gd.storage = hdt;
            gd.width = input.getWidth();
            gd.height = input.getHeight();
            gd.depth = dims() == 3 ? input.getDepth() : 1;
            gd.device = s.device;
            pskel_stencil_desc_t desc = StencilTraits<Args>::describe(mask, args, dims());
            // whole arrays are read and written in place, slices through packed copies
            std::vector<T> in, out;
            const T* inHost = input.denseHost();
            T* outHost = output.denseHost();
            if (!inHost) {
                in.resize(input.size());
                input.packTo(in.data());
                inHost = in.data();
            }
            if (!outHost) {
                out.resize(output.size());
                output.packTo(out.data());
                outHost = out.data();
            }
            pskel_launch_cfg_t cfg{};
            cfg.variant = PSKEL_VARIANT_TUNED;
            if (s.haveCfg) cfg = s.cfg;
            cfg.residual_every = 0;
            pskel_run_stats_t st{};
            gpuErrchk(pskel_run_iterative_host(&gd, inHost, outHost, hdt, &desc, iterations, tile, inner, &cfg, &st));
            s.lastMs = st.elapsed_ms;
            if (!out.empty()) output.unpackFrom(out.data());
        }
    }

    void gpuRun(size_t iterations, pskel_launch_cfg_t cfg, bool plain, bool autotune = false)
    {
        if constexpr (Array::kDims == 1) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);  // 1D GPU does not build in the reference either
        detail::GpuState& s = *gpu;
        ensureGrid();
        const int hdt = detail::hostDtype<T>();
        if (hdt == 0) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
        // upload both buffers: the device state starts as the host state
        // (whole arrays go from their own host memory, slices are packed first)
        std::vector<T> buf, obuf;
        const T* inHost = input.denseHost();
        const T* outHost = output.denseHost();
        if (!inHost) {
            buf.resize(input.size());
            input.packTo(buf.data());
            inHost = buf.data();
        }
        if (!outHost) {
            obuf.resize(output.size());
            output.packTo(obuf.data());
            outHost = obuf.data();
        }
        auto upload = [&]() {
            gpuErrchk(pskel_grid_upload(s.grid, PSKEL_BUFFER_INPUT, inHost, hdt));
            gpuErrchk(pskel_grid_upload(s.grid, PSKEL_BUFFER_OUTPUT, outHost, hdt));
        };
        upload();
        if (!plain && s.haveCfg) cfg = s.cfg;
        cfg.residual_every = (int)s.residualEvery;
        s.residuals.clear();

        // field arguments (Args/Args2D/Args3D): their host arrays go to HBM once per call
        detail::syncFields();
        if constexpr (StencilTraits<Args>::has_descriptor) {
            pskel_stencil_desc_t desc = StencilTraits<Args>::describe(mask, args, dims());
            if constexpr (detail::hasFields<StencilTraits<Args>>::value) {
                pskel_field_t* fl[StencilTraits<Args>::num_fields];
                StencilTraits<Args>::fields(args, fl);
                for (int k = 0; k < StencilTraits<Args>::num_fields; ++k)
                    gpuErrchk(pskel_grid_set_field(s.grid, k, fl[k]));
            }
            if (autotune && !s.haveCfg && iterations > 0) {
                pskel_launch_cfg_t tuned{};
                int32_t haveModel = 0;
                gpuErrchk(pskel_tuner_has_model(&haveModel));
                if (haveModel) {
                    pskel_tune_query_t q{};
                    q.ndim = dims();
                    q.width = input.getWidth();
                    q.height = input.getHeight();
                    q.depth = dims() == 3 ? input.getDepth() : 1;
                    q.storage = s.gridStorage;
                    q.iterations = iterations;
                    gpuErrchk(pskel_tuner_choose(nullptr, &q, &desc, &tuned));
                } else {
                    // no cost model: time the candidates on this grid (once per
                    // shape in the process), then start again from the host state
                    gpuErrchk(pskel_tuner_search(s.grid, &desc, iterations, PSKEL_SEARCH_EXHAUSTIVE, 0, 3, &tuned,
                                                 nullptr, nullptr));
                    upload();
                }
                tuned.residual_every = cfg.residual_every;
                cfg = tuned;
            }
            pskel_run_stats_t st{};
            std::vector<double> res(s.residualEvery ? iterations / s.residualEvery : 0);
            st.residuals = res.empty() ? nullptr : res.data();
            st.residual_capacity = (int64_t)res.size();
            gpuErrchk(pskel_run_iterative(s.grid, &desc, iterations, &cfg, &st));
            s.residuals.assign(res.begin(), res.begin() + std::min<int64_t>(st.residual_count, (int64_t)res.size()));
            s.lastMs = st.elapsed_ms;
        } else {
#if defined(PSKEL_DEVICE_COMPILER)
            runGeneric(iterations, cfg);
#else
            gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);  // compile the TU with hipcc or describe the kernel
#endif
        }
        if (T* dst = output.denseHost()) {
            gpuErrchk(pskel_grid_download(s.grid, dst, hdt));
        } else {
            obuf.resize(output.size());
            gpuErrchk(pskel_grid_download(s.grid, obuf.data(), hdt));
            output.unpackFrom(obuf.data());
        }
#ifdef PSKEL_PRINT_TIMES
        std::cout << "GPU_time\t" << s.lastMs * 1e-3 << std::endl;
#endif
    }

#if defined(PSKEL_DEVICE_COMPILER)
    // The user's stencilKernel compiled for the device, on the grid's buffers.
    void runGeneric(size_t iterations, const pskel_launch_cfg_t& cfg)
    {
        detail::GpuState& s = *gpu;
        if (s.gridStorage != detail::hostDtype<T>()) gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
        pskel_grid_info_t gi{};
        gpuErrchk(pskel_grid_info(s.grid, &gi));
        hipStream_t stream = (hipStream_t)gi.stream;
        Mask m = mask;
        m.deviceAlloc();
        m.copyToDevice();
        const size_t r = mask.getRange();
        const unsigned bx = cfg.block_x > 0 ? (unsigned)cfg.block_x : 64u;
        const unsigned by = cfg.block_y > 0 ? (unsigned)cfg.block_y : 4u;
        const dim3 block(bx, by);
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, stream);
        for (size_t it = 0; it < iterations; ++it) {
            gpuErrchk(pskel_grid_info(s.grid, &gi));
            Array in = input, out = output;
            in.setDeviceView((T*)gi.current, gi.pitch);
            out.setDeviceView((T*)gi.previous, gi.pitch);
            if constexpr (Array::kDims == 2) {
                const dim3 grid((unsigned)((gi.width + bx - 1) / bx), (unsigned)((gi.height + by - 1) / by));
                hipLaunchKernelGGL((detail::genericKernel2D<Array, Mask, Args>), grid, block, 0, stream, in, out, m,
                                   args, r, gi.width, gi.height);
            } else if constexpr (Array::kDims == 3) {
                const dim3 grid((unsigned)((gi.depth + bx - 1) / bx), (unsigned)((gi.width + by - 1) / by),
                                (unsigned)gi.height);
                hipLaunchKernelGGL((detail::genericKernel3D<Array, Mask, Args>), grid, block, 0, stream, in, out, m,
                                   args, gi.width, gi.height, gi.depth);
            } else {
                gpuErrchk((pskel_status_t)PSKEL_ERR_UNSUPPORTED);
            }
            if (hipGetLastError() != hipSuccess) gpuErrchk((pskel_status_t)PSKEL_ERR_LAUNCH);
            gpuErrchk(pskel_grid_swap(s.grid));
            // FP64 ||u^k - u^(k-1)|| after every n-th sweep (on the grid's stream)
            if (s.residualEvery && (it + 1) % s.residualEvery == 0) {
                double r2 = 0.0;
                gpuErrchk(pskel_grid_residual(s.grid, &r2));
                s.residuals.push_back(r2);
            }
        }
        (void)hipEventRecord(e1, stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        s.lastMs = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        m.deviceFree();
    }
#endif
#endif  // PSKEL_GPU
};

//*******************************************************************************************
// Stencil 3D (include/PSkelStencil.h:317-343, include/PSkelStencil.hpp:2084-2113)
//*******************************************************************************************
template <class Array, class Mask, class Args>
class Stencil3D : public StencilBase<Array, Mask, Args> {
protected:
    int dims() const override { return 3; }

    void runSeq(Array in, Array out, size_t, size_t, size_t) override
    {
        for (size_t h = 0; h < in.getHeight(); h++)
            for (size_t w = 0; w < in.getWidth(); w++)
                for (size_t d = 0; d < in.getDepth(); d++) stencilKernel(in, out, this->mask, this->args, h, w, d);
    }

    void runOpenMP(Array in, Array out, size_t width, size_t height, size_t depth, size_t, size_t numThreads) override
    {
        (void)numThreads;
#pragma omp parallel for num_threads(numThreads) schedule(static)
        for (long long h = 0; h < (long long)height; h++)
            for (size_t w = 0; w < width; w++)
                for (size_t d = 0; d < depth; d++)
                    stencilKernel(in, out, this->mask, this->args, (size_t)h, w, d);
    }

    std::unique_ptr<StencilBase<Array, Mask, Args>> cloneWith(Array in, Array out) const override
    {
        return std::make_unique<Stencil3D>(in, out, this->mask, this->args);
    }

public:
    Stencil3D() {}
    Stencil3D(Array _input, Array _output, Mask _mask, Args _args)
    {
        this->input = _input;
        this->output = _output;
        this->args = _args;
        this->mask = _mask;
    }
};

//*******************************************************************************************
// Stencil 2D (include/PSkelStencil.h:349-377, include/PSkelStencil.hpp:2149-2257)
//*******************************************************************************************
template <class Array, class Mask, class Args>
class Stencil2D : public StencilBase<Array, Mask, Args> {
protected:
    int dims() const override { return 2; }

    void runSeq(Array in, Array out, size_t width, size_t height, size_t maskRange) override
    {
        if (height < 2 * maskRange || width < 2 * maskRange) return;
        const size_t hrange = height - maskRange, wrange = width - maskRange;
        for (size_t h = maskRange; h < hrange; h++)
            for (size_t w = maskRange; w < wrange; w++) stencilKernel(in, out, this->mask, this->args, h, w);
    }

    void runOpenMP(Array in, Array out, size_t width, size_t height, size_t, size_t maskRange,
                   size_t numThreads) override
    {
        (void)numThreads;
        if (height < 2 * maskRange || width < 2 * maskRange) return;
        const long long hrange = (long long)(height - maskRange);
        const size_t wrange = width - maskRange;
#pragma omp parallel for num_threads(numThreads) schedule(static)
        for (long long h = (long long)maskRange; h < hrange; h++)
            for (size_t w = maskRange; w < wrange; w++) stencilKernel(in, out, this->mask, this->args, (size_t)h, w);
    }

    std::unique_ptr<StencilBase<Array, Mask, Args>> cloneWith(Array in, Array out) const override
    {
        return std::make_unique<Stencil2D>(in, out, this->mask, this->args);
    }

public:
    Stencil2D() {}
    Stencil2D(Array _input, Array _output, Mask _mask, Args _args)
    {
        this->input = _input;
        this->output = _output;
        this->args = _args;
        this->mask = _mask;
    }
};

//*******************************************************************************************
// Stencil 1D (include/PSkelStencil.h:384-407, include/PSkelStencil.hpp:2300-2328); CPU only
//*******************************************************************************************
template <class Array, class Mask, class Args>
class Stencil : public StencilBase<Array, Mask, Args> {
protected:
    int dims() const override { return 1; }

    void runSeq(Array in, Array out, size_t, size_t, size_t) override
    {
        for (size_t i = 0; i < in.getWidth(); i++) stencilKernel(in, out, this->mask, this->args, i);
    }

    void runOpenMP(Array in, Array out, size_t width, size_t, size_t, size_t, size_t numThreads) override
    {
        (void)numThreads;
#pragma omp parallel for num_threads(numThreads) schedule(static)
        for (long long i = 0; i < (long long)width; i++) stencilKernel(in, out, this->mask, this->args, (size_t)i);
    }

    std::unique_ptr<StencilBase<Array, Mask, Args>> cloneWith(Array in, Array out) const override
    {
        return std::make_unique<Stencil>(in, out, this->mask, this->args);
    }

public:
    Stencil() {}
    Stencil(Array _input, Array _output, Mask _mask, Args _args)
    {
        this->input = _input;
        this->output = _output;
        this->args = _args;
        this->mask = _mask;
    }
};

}  // namespace PSkel

#endif
