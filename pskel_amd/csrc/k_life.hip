// Game of Life (B3/S23) sweeps on u8 grids for gfx950, bit-exact integer work.
//
// Reference rule: the GOL_KERNEL formula (include/PSkelStencil.hpp:389-440),
// sum of the 8 Moore neighbours, alive if sum == 3 or (sum == 2 and alive).
// Interior-only like every 2D runner (include/PSkelStencil.hpp:576, 2235-2255).
//
//   k_life_plain  one lane per cell, the caller's block: the runCUDA shape.
//   k_life_swar   16 cells per lane packed in two u64 words; horizontal
//                 neighbours by byte shifts plus one cross-lane exchange per row,
//                 counts and rule evaluated bytewise (SWAR), RY rows per lane.
//   k_life_bits   K generations per launch on
//                 cells packed one bit each: each wave evolves its own tile
//                 in registers.
#include "k_tiles2d.hpp"

namespace pskel {

namespace {

// ---------------------------------------------------------------------------
// plain
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(1024) void k_life_plain(Sweep2D s, Points p)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y * blockDim.y + threadIdx.y;
    const int r = s.radius;
    if (w < r || w >= s.W - r || h < r || h >= s.H - r) return;
    const uint8_t* __restrict__ c = (const uint8_t*)s.src + (size_t)h * s.pitch + w;
    const ptrdiff_t pitch = (ptrdiff_t)s.pitch;
    int n = 0;
    const int np = NP > 0 ? NP : p.np;
#pragma unroll
    for (int i = 0; i < (NP > 0 ? NP : kMaxPoints); ++i) {
        if (NP == 0 && i >= np) break;
        n += c[p.dh[i] * pitch + p.dw[i]];
    }
    ((uint8_t*)s.dst)[(size_t)h * s.pitch + w] = (n == 3 || (n == 2 && c[0] == 1)) ? 1 : 0;
}

template <int RY>
__global__ __launch_bounds__(256) void k_life_swar(Sweep2D s)
{
    tile_life<RY>(s, blockIdx.x, blockIdx.y);
}

template <int K, int R>
__global__ __launch_bounds__(256) void k_life_bits(Sweep2D s, int ntx, int ntiles)
{
    const int t = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (t >= ntiles) return;  // whole waves; the tile body does not synchronise
    life_bits_tile<K, R>(s, ntx, t, (int)threadIdx.x % kWave, K);
}

}  // namespace

hipError_t launch_life_bits(const Sweep2D& s, int k, const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    // rows per lane: cfg.tile_y rounded up to 24, 32, 40, 48 or 64, and at
    // least 2k + 8 (8 core rows)
    int r = std::max(cfg.tile_y > 0 ? cfg.tile_y : 32, 2 * k + 8);
    r = r <= 24 ? 24 : (r <= 32 ? 32 : (r <= 40 ? 40 : (r <= 48 ? 48 : 64)));
    constexpr int CW = 62 * 32;
    const int ntx = std::max(1, (s.W - 1 + CW - 1) / CW);
#define PSKEL_LIFE_BITS(KV, RV)                                                                     \
    if (k == KV && r == RV) {                                                                       \
        const int nty = std::max(1, (s.H - 1 + (RV - 2 * KV) - 1) / (RV - 2 * KV));                 \
        const int ntiles = ntx * nty;                                                               \
        hipLaunchKernelGGL((k_life_bits<KV, RV>), dim3((ntiles + 3) / 4), dim3(256), 0, st, s, ntx, \
                           ntiles);                                                                 \
        return hipGetLastError();                                                                   \
    }
#define PSKEL_LIFE_BITS_K(KV)                                                                        \
    PSKEL_LIFE_BITS(KV, 24) PSKEL_LIFE_BITS(KV, 32) PSKEL_LIFE_BITS(KV, 40) PSKEL_LIFE_BITS(KV, 48) \
    PSKEL_LIFE_BITS(KV, 64)
    PSKEL_LIFE_BITS_K(2)
    PSKEL_LIFE_BITS_K(3)
    PSKEL_LIFE_BITS_K(4)
    PSKEL_LIFE_BITS_K(5)
    PSKEL_LIFE_BITS_K(6)
    PSKEL_LIFE_BITS_K(7)
    PSKEL_LIFE_BITS_K(8)
#undef PSKEL_LIFE_BITS_K
#undef PSKEL_LIFE_BITS
    return hipErrorInvalidValue;
}

hipError_t launch_life_plain(const Sweep2D& s, const Points& p, int bx, int by, hipStream_t st)
{
    const dim3 block(bx, by);
    const dim3 grid((s.W + bx - 1) / bx, (s.H + by - 1) / by);
    if (p.np == 8) hipLaunchKernelGGL((k_life_plain<8>), grid, block, 0, st, s, p);
    else hipLaunchKernelGGL((k_life_plain<0>), grid, block, 0, st, s, p);
    return hipGetLastError();
}

hipError_t launch_life_tuned(const Sweep2D& s, const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    const int ry = cfg.tile_y > 0 ? cfg.tile_y : 4;
    const int nwy = cfg.tile_z > 0 ? cfg.tile_z : 4;
    if (nwy < 1 || nwy > 4) return hipErrorInvalidValue;
    const dim3 block(kWave, nwy);
#define PSKEL_LIFE_CASE(RYV)                                                                     \
    case RYV: {                                                                                  \
        const dim3 grid((s.W + 1023) / 1024, (s.H + nwy * RYV - 1) / (nwy * RYV));               \
        hipLaunchKernelGGL((k_life_swar<RYV>), grid, block, 0, st, s);                           \
        break;                                                                                   \
    }
    switch (ry) {
        PSKEL_LIFE_CASE(1)
        PSKEL_LIFE_CASE(2)
        PSKEL_LIFE_CASE(4)
        PSKEL_LIFE_CASE(8)
        PSKEL_LIFE_CASE(16)
    default: return hipErrorInvalidValue;
    }
#undef PSKEL_LIFE_CASE
    return hipGetLastError();
}

}  // namespace pskel
