// Utility kernels: plain copy (bandwidth yardstick), zero fill, 2D ring copy.
#include "k_common.hpp"

namespace pskel {

namespace {

// 16-byte vector copy, one vector per lane and one 4 KiB chunk per block: the
// shape that reaches the device's copy bandwidth (a grid-stride loop over a
// capped grid does not); tail bytes handled one by one.
__global__ __launch_bounds__(256) void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                              size_t n16, uint8_t* __restrict__ dtail,
                                              const uint8_t* __restrict__ stail, size_t ntail)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ dst, size_t n16,
                                                   uint8_t* __restrict__ tail, size_t ntail)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// Copy the 2D ring (cells outside [r, N-r)^2) of src into dst, element size esz.
__global__ __launch_bounds__(256) void k_copy_ring2d(Sweep2D s, int esz)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (w >= s.W || h >= s.H) return;
    const int r = s.radius;
    if (h >= r && h < s.H - r && w >= r && w < s.W - r) return;
    const size_t i = ((size_t)h * s.pitch + w) * esz;
    const uint8_t* a = (const uint8_t*)s.src + i;
    uint8_t* b = (uint8_t*)s.dst + i;
    for (int k = 0; k < esz; ++k) b[k] = a[k];
}

// FP64 partial sums of (dec(a) - dec(b))^2 over rows; block b takes rows
// b, b + nb, ...  (fixed order: deterministic for a given nb).
template <int ST>
__global__ __launch_bounds__(256) void k_diff_partials(const void* a_, const void* b_, size_t nrows,
                                                       size_t rowlen, size_t pitch,
                                                       double* __restrict__ partials)
{
    using Tr = Stor<ST>;
    using T = typename Tr::T;
    __shared__ double red[4];
    const T* a = (const T*)a_;
    const T* b = (const T*)b_;
    double acc = 0.0;
    for (size_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        const size_t base = row * pitch;
        for (size_t i = threadIdx.x; i < rowlen; i += blockDim.x) {
            const double dv = (double)Tr::dec_raw(a[base + i]) - (double)Tr::dec_raw(b[base + i]);
            acc += dv * dv;
        }
    }
    const double t = block_sum<4>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_diff_partials_u8(const uint8_t* a, const uint8_t* b, size_t nrows,
                                                          size_t rowlen, size_t pitch,
                                                          double* __restrict__ partials)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (size_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        const size_t base = row * pitch;
        for (size_t i = threadIdx.x; i < rowlen; i += blockDim.x) {
            const double dv = (double)a[base + i] - (double)b[base + i];
            acc += dv * dv;
        }
    }
    const double t = block_sum<4>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

unsigned copy_blocks(size_t n16)
{
    const size_t want = (n16 + 255) / 256;
    const size_t cap = 256 * 32;  // 32 blocks per CU on 256 CUs, then grid-stride
    return (unsigned)(want < 1 ? 1 : (want > cap ? cap : want));
}

}  // namespace

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    const size_t n16 = bytes / 16, tail = bytes % 16;
    const size_t nb = (n16 + 255) / 256;
    if (nb > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)(nb ? nb : 1)), dim3(256), 0, st, (uint4*)dst, (const uint4*)src,
                       n16, (uint8_t*)dst + n16 * 16, (const uint8_t*)src + n16 * 16, tail);
    return hipGetLastError();
}

hipError_t launch_fill_zero(void* dst, size_t bytes, hipStream_t st)
{
    const size_t n16 = bytes / 16, tail = bytes % 16;
    hipLaunchKernelGGL(k_fill_zero, dim3(copy_blocks(n16)), dim3(256), 0, st, (uint4*)dst, n16,
                       (uint8_t*)dst + n16 * 16, tail);
    return hipGetLastError();
}

size_t diff_blocks(size_t nrows) { return nrows < 2048 ? (nrows ? nrows : 1) : 2048; }

hipError_t launch_diff_partials(const void* a, const void* b, int storage, size_t nrows, size_t rowlen,
                                size_t pitch, double* partials, size_t nblocks, hipStream_t st)
{
    const dim3 grid((unsigned)nblocks), block(256);
    switch (storage) {
    case PSKEL_DTYPE_F32:
        hipLaunchKernelGGL(k_diff_partials<PSKEL_DTYPE_F32>, grid, block, 0, st, a, b, nrows, rowlen, pitch, partials);
        break;
    case PSKEL_DTYPE_BF16:
        hipLaunchKernelGGL(k_diff_partials<PSKEL_DTYPE_BF16>, grid, block, 0, st, a, b, nrows, rowlen, pitch, partials);
        break;
    case PSKEL_DTYPE_F16:
        hipLaunchKernelGGL(k_diff_partials<PSKEL_DTYPE_F16>, grid, block, 0, st, a, b, nrows, rowlen, pitch, partials);
        break;
    case PSKEL_DTYPE_U8:
        hipLaunchKernelGGL(k_diff_partials_u8, grid, block, 0, st, (const uint8_t*)a, (const uint8_t*)b, nrows, rowlen,
                           pitch, partials);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_copy_ring2d(const Sweep2D& s, int elem, hipStream_t st)
{
    const dim3 grid((s.W + 255) / 256, s.H);
    hipLaunchKernelGGL(k_copy_ring2d, grid, dim3(256), 0, st, s, elem);
    return hipGetLastError();
}

}  // namespace pskel
