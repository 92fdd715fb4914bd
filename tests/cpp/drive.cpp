// Driver for the PSkel drop-in headers, used by tests/test_dropin_*.py.
//
//   drive <kernel> <method> <W> <H> <D> <iters> <input.bin> <output0.bin> <result.bin>
//
// kernel: jacobi2d | diffusion9 | life | jacobi3d | user2d | user3d
// method: cpu | seq | gpu (32 x 4 blocks) | default (runIterativeGPU(N): no
//         block sizes) | pyramid | tiling | auto | ...
// Arrays are raw little-endian host arrays in Array2D/Array3D order
// (float, or uint8 for life).  The program is written the way a PSkel user
// writes one: Arrays, a Mask, an Args, Stencil2D/3D, one run method.
//
// jacobi3d also takes `gen:<seed>` for both input files (the field generated
// here: uniform [0,1) from splitmix64((seed << 40) ^ i), top 24 bits, faces at
// 1.0) and `-` for the result file (not written).
// Environment: PSKEL_DRIVE_RESIDUAL=n sets setResidualInterval(n) and prints
// "residuals r1 r2 ..." on stderr; PSKEL_DRIVE_TIMES prints "gpu_ms <device
// time of the sweeps> wall_ms <whole run method, host copies included>";
// PSKEL_DRIVE_REPEAT=n runs the method n times (the GPU methods give the same
// result each time: they leave `input` alone), timing each.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "PSkel.h"

using namespace PSkel;

// A user kernel with no descriptor: the GPU methods run it on the device only
// when this file is compiled by hipcc.
struct UserJacobiArgs {
    float c;
};

namespace PSkel {
__parallel__ void stencilKernel(Array2D<float> in, Array2D<float> out, Mask2D<float> mask, UserJacobiArgs a,
                                size_t h, size_t w)
{
    (void)mask;
    out(h, w) = a.c * (((in(h - 1, w) + in(h + 1, w)) + in(h, w - 1)) + in(h, w + 1));
}

__parallel__ void stencilKernel(Array3D<float> in, Array3D<float> out, Mask3D<float> mask, UserJacobiArgs a,
                                size_t h, size_t w, size_t d)
{
    (void)mask;
    if (h == 0 || w == 0 || d == 0 || h + 1 == in.getHeight() || w + 1 == in.getWidth() || d + 1 == in.getDepth()) {
        out(h, w, d) = in(h, w, d);
        return;
    }
    float acc = in(h - 1, w, d) + in(h + 1, w, d);
    acc = acc + in(h, w - 1, d);
    acc = acc + in(h, w + 1, d);
    acc = acc + in(h, w, d - 1);
    acc = acc + in(h, w, d + 1);
    out(h, w, d) = a.c * acc;
}
// A user kernel reading field arguments of its own Args type (no descriptor):
// on the GPU it runs only when this file is compiled by hipcc, with the fields
// read from their device copies.
struct UserCloudArgs {
    Args2D<float> wx, wy;
    float dt;
};

__parallel__ void stencilKernel(Array2D<float> in, Array2D<float> out, Mask2D<float> mask, UserCloudArgs a, size_t h,
                                size_t w)
{
    (void)mask;
    const float c = in(h, w);
    const float sum = 4 * c - (in(h - 1, w) + in(h, w + 1) + in(h, w - 1) + in(h + 1, w));
    const float xwind = a.wx((int)h, (int)w), ywind = a.wy((int)h, (int)w);
    const int xf = xwind > 0 ? 1 : -1, yf = ywind > 0 ? 1 : -1;
    const float cy = yf * ywind, cx = xf * xwind;
    const float tw = (-cx * ((c - in(h, (size_t)((long)w + xf))) * 10.0f)) -
                     (cy * ((c - in((size_t)((long)h + yf), w)) * 10.0f));
    const float result = c + -0.0243f * (sum * 0.25f) * a.dt;
    out(h, w) = result + tw * a.dt;
}
// A user map kernel (Map / Map2D / Map3D, include/PSkelMap.h:53-59).
struct MapArgs {
    float c;
};

__parallel__ void mapKernel(Array<float> in, Array<float> out, MapArgs a, size_t i)
{
    const float x = in(i);
    out(i) = 0.5f * x * x + a.c;
}

__parallel__ void mapKernel(Array2D<float> in, Array2D<float> out, MapArgs a, size_t h, size_t w)
{
    const float x = in(h, w);
    out(h, w) = 0.5f * x * x + a.c;
}

__parallel__ void mapKernel(Array3D<float> in, Array3D<float> out, MapArgs a, size_t h, size_t w, size_t d)
{
    const float x = in(h, w, d);
    out(h, w, d) = 0.5f * x * x + a.c;
}
}  // namespace PSkel

template <typename T>
static bool readf(const char* p, T* dst, size_t n)
{
    FILE* f = std::fopen(p, "rb");
    if (!f) return false;
    const bool ok = std::fread(dst, sizeof(T), n, f) == n;
    std::fclose(f);
    return ok;
}

template <typename T>
static bool writef(const char* p, const T* src, size_t n)
{
    FILE* f = std::fopen(p, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(src, sizeof(T), n, f) == n;
    std::fclose(f);
    return ok;
}

template <class St>
static int runMethodOnce(St& st, const std::string& method, size_t iters);

template <class St>
static int runMethod(St& st, const std::string& method, size_t iters)
{
    const char* rep = std::getenv("PSKEL_DRIVE_REPEAT");
    const long n = rep ? std::max(1L, std::strtol(rep, nullptr, 10)) : 1L;
    for (long i = 0; i < n; ++i) {
        const int rc = runMethodOnce(st, method, iters);
        if (rc) return rc;
    }
    return 0;
}

template <class St>
static int runMethodOnce(St& st, const std::string& method, size_t iters)
{
#ifdef PSKEL_GPU
    if (const char* r = std::getenv("PSKEL_DRIVE_RESIDUAL")) st.setResidualInterval(std::strtoull(r, nullptr, 10));
#endif
    const auto t0 = std::chrono::steady_clock::now();
    if (method == "cpu") st.runIterativeCPU(iters, 0);
    else if (method == "seq") st.runIterativeSequential(iters);
#ifdef PSKEL_GPU
    else if (method == "gpu") st.runIterativeGPU(iters, 32, 4);
    else if (method == "default") st.runIterativeGPU(iters);
    else if (method == "pyramid") st.runIterativeGPU(iters, 2, 32, 4);
    else if (method == "tiling") st.runIterativeTilingGPU(iters, 256, 256, 1, 3);
    else if (method == "auto") st.runIterativeAutoGPU(iters);
    else if (method == "partition") st.runIterativePartition(iters, 0.4f, 0, 32, 4);
    else if (method == "partition_small") st.runIterativePartition(iters, 0.05f, 0, 32, 4);
    else if (method == "hosttiling") {
        st.setHostTiling(true);
        st.runIterativeTilingGPU(iters, 256, 7, 1, 3);
    } else if (method == "hostauto") {
        st.setHostTiling(true);
        st.runIterativeAutoGPU(iters);
    }
#endif
    else {
        std::fprintf(stderr, "unknown method %s\n", method.c_str());
        return 2;
    }
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
#ifdef PSKEL_GPU
    if (std::getenv("PSKEL_DRIVE_TIMES")) std::fprintf(stderr, "gpu_ms %.6f wall_ms %.6f\n", st.lastGpuMillis(), wall);
    if (std::getenv("PSKEL_DRIVE_RESIDUAL")) {
        std::fprintf(stderr, "residuals");
        for (double r : st.residuals()) std::fprintf(stderr, " %.17g", r);
        std::fprintf(stderr, "\n");
    }
#else
    if (std::getenv("PSKEL_DRIVE_TIMES")) std::fprintf(stderr, "wall_ms %.6f\n", wall);
#endif
    return 0;
}

template <typename T, class Mask, class Args>
static int run2d(Mask mask, Args args, const std::string& method, size_t W, size_t H, size_t iters, char** files)
{
    Array2D<T> in(W, H), out(W, H);
    if (!readf(files[0], in.hostArray, W * H) || !readf(files[1], out.hostArray, W * H)) return 3;
    Stencil2D<Array2D<T>, Mask, Args> st(in, out, mask, args);
    const int rc = runMethod(st, method, iters);
    if (rc) return rc;
    if (!writef(files[2], out.hostArray, W * H)) return 4;
    in.hostFree();
    out.hostFree();
    mask.hostFree();  // handles are shallow: the program frees what it allocated
    return 0;
}

// `gen:<seed>`: the Jacobi-3D field, uniform [0,1) with every face at 1.0
static bool genJacobi3D(const char* spec, float* dst, size_t W, size_t H, size_t D)
{
    if (std::strncmp(spec, "gen:", 4) != 0) return false;
    const uint64_t seed = std::strtoull(spec + 4, nullptr, 10);
    for (size_t h = 0; h < H; ++h)
        for (size_t w = 0; w < W; ++w)
            for (size_t d = 0; d < D; ++d) {
                const size_t i = (h * W + w) * D + d;
                uint64_t z = ((seed << 40) ^ (uint64_t)i) + 0x9E3779B97F4A7C15ull;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                const bool face = h == 0 || w == 0 || d == 0 || h + 1 == H || w + 1 == W || d + 1 == D;
                dst[i] = face ? 1.0f : (float)(z >> 40) * (1.0f / 16777216.0f);
            }
    return true;
}

template <class Mask, class Args>
static int run3d(Mask mask, Args args, const std::string& method, size_t W, size_t H, size_t D, size_t iters,
                 char** files)
{
    Array3D<float> in(W, H, D), out(W, H, D);
    const size_t n = W * H * D;
    if (!genJacobi3D(files[0], in.hostArray, W, H, D) && !readf(files[0], in.hostArray, n)) return 3;
    if (!genJacobi3D(files[1], out.hostArray, W, H, D) && !readf(files[1], out.hostArray, n)) return 3;
    Stencil3D<Array3D<float>, Mask, Args> st(in, out, mask, args);
    const int rc = runMethod(st, method, iters);
    if (rc) return rc;
    if (std::strcmp(files[2], "-") != 0 && !writef(files[2], out.hostArray, n)) return 4;
    in.hostFree();
    out.hostFree();
    mask.hostFree();
    return 0;
}

// Wind fields for the CLOUDSIM kernels: Args2D(width = H, height = W), host
// layout from the files (x fastest), read as wind(h, w).
template <class A>
static int runCloud(const std::string& method, size_t W, size_t H, size_t iters, char** files, char** wind)
{
    A a;
    Args2D<float> wx((int)H, (int)W), wy((int)H, (int)W);
    if (!readf(wind[0], wx.hostArray, W * H) || !readf(wind[1], wy.hostArray, W * H)) return 3;
    if constexpr (std::is_same_v<A, CloudsimArgs>) {
        a.wind_x = wx;
        a.wind_y = wy;
        a.deltaT = 0.01f;
    } else {
        a.wx = wx;
        a.wy = wy;
        a.dt = 0.01f;
    }
    const int rc = run2d<float>(jacobiMask2D<float>(), a, method, W, H, iters, files);
    wx.hostFree();
    wy.hostFree();
    return rc;
}

template <class M, class A>
static int runMap(M& m, A& in, A& out, const std::string& method, size_t iters, size_t n, char** files)
{
    if (!readf(files[0], in.hostArray, n) || !readf(files[1], out.hostArray, n)) return 3;
    if (method == "cpu") m.runIterativeCPU(iters, 0);
    else if (method == "seq") m.runIterativeSequential(iters);
#ifdef PSKEL_GPU
    else if (method == "gpu") {
        m.runIterativeGPU(iters);
        if (std::getenv("PSKEL_DRIVE_TIMES")) std::fprintf(stderr, "gpu_ms %.6f\n", m.lastGpuMillis());
    }
#endif
    else return 2;
    if (!writef(files[2], out.hostArray, n)) return 4;
    in.hostFree();
    out.hostFree();
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 10 && argc != 12) {
        std::fprintf(stderr, "usage: drive kernel method W H D iters input output0 result [wind_x wind_y]\n");
        return 2;
    }
    const std::string kernel = argv[1], method = argv[2];
    const size_t W = std::strtoull(argv[3], nullptr, 10), H = std::strtoull(argv[4], nullptr, 10),
                 D = std::strtoull(argv[5], nullptr, 10), iters = std::strtoull(argv[6], nullptr, 10);
    char** files = argv + 7;
    try {
        if (kernel == "jacobi2d") return run2d<float>(jacobiMask2D<float>(), ScaledSumArgs{0.25, 0}, method, W, H, iters, files);
        if (kernel == "map1d") {
            Array<float> in(W), out(W);
            Map<Array<float>, MapArgs> m(in, out, MapArgs{0.125f});
            return runMap(m, in, out, method, iters, W, files);
        }
        if (kernel == "map2d") {
            Array2D<float> in(W, H), out(W, H);
            Map2D<Array2D<float>, MapArgs> m(in, out, MapArgs{0.125f});
            return runMap(m, in, out, method, iters, W * H, files);
        }
        if (kernel == "map3d") {
            Array3D<float> in(W, H, D), out(W, H, D);
            Map3D<Array3D<float>, MapArgs> m(in, out, MapArgs{0.125f});
            return runMap(m, in, out, method, iters, W * H * D, files);
        }
        if (kernel == "cloudsim" || kernel == "usercloud") {
            if (argc != 12) return 2;
            if (kernel == "cloudsim") return runCloud<CloudsimArgs>(method, W, H, iters, files, argv + 10);
            return runCloud<UserCloudArgs>(method, W, H, iters, files, argv + 10);
        }
        if (kernel == "user2d") return run2d<float>(jacobiMask2D<float>(), UserJacobiArgs{0.25f}, method, W, H, iters, files);
        if (kernel == "diffusion9") {
            // centre 0.5, edge 0.1, corner 0.025, skewed by the wind (0.02, -0.01)
            const double wx = 0.02, wy = -0.01;
            float wts[9];
            for (int i = 0; i < 9; ++i) {
                const int dh = i / 3 - 1, dw = i % 3 - 1;
                double base, g;
                if (!dh && !dw) base = 0.5, g = 0.0;
                else if (!dh || !dw) base = 0.1, g = 1.0;
                else base = 0.025, g = 0.25;
                wts[i] = (float)(base - g * (dw * wx + dh * wy));
            }
            return run2d<float>(boxMask2D<float>(wts), WeightedSumArgs{0}, method, W, H, iters, files);
        }
        if (kernel == "life") return run2d<unsigned char>(mooreMask2D<unsigned char>(), LifeArgs{0}, method, W, H, iters, files);
        if (kernel == "jacobi3d") return run3d(jacobiMask3D<float>(), ScaledSumArgs{1.0 / 6.0, BOUNDARY_COPY}, method, W, H, D, iters, files);
        if (kernel == "user3d") return run3d(jacobiMask3D<float>(), UserJacobiArgs{(float)(1.0 / 6.0)}, method, W, H, D, iters, files);
    } catch (const PSkel::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 10 + e.status;
    }
    std::fprintf(stderr, "unknown kernel %s\n", kernel.c_str());
    return 2;
}
