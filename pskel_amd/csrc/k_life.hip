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
//   k_life_ttile  K generations per launch on a u64-word LDS tile.
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

// ---------------------------------------------------------------------------
// temporal tile: the tile is a grid of u64 words (8 cells each), core CXW words
// wide and CY rows tall; the loaded region adds one word (8 cells >= K) on each
// side and K rows above and below.  Level j lives in LDS buffer j&1; ring cells
// are never recomputed and hold the ring of the buffer their level maps to.
// Per level each thread evaluates whole words: out = rule(counts) merged with
// the previous word where the cell is not an interior cell.
// ---------------------------------------------------------------------------
template <int K, int CXW, int CY>
__global__ __launch_bounds__(256) void k_life_ttile(Sweep2D s)
{
    constexpr int LXW = CXW + 2;  // words per LDS row
    constexpr int LY = CY + 2 * K;
    constexpr int NT = 256;
    __shared__ uint64_t A[LY * LXW];
    __shared__ uint64_t B[LY * LXW];
    __shared__ uint64_t M[LY * LXW];  // 0xff bytes where the cell is interior
    const int tid = threadIdx.x;
    const int w0 = blockIdx.x * (CXW * 8);  // core origin (cells)
    const int h0 = blockIdx.y * CY;
    const int gx0 = w0 - 8, gy0 = h0 - K;
    const int H = s.H, Wd = s.W, r = s.radius;
    const size_t pitch = s.pitch;
    const uint8_t* __restrict__ src = (const uint8_t*)s.src;
    const uint8_t* __restrict__ oth = (const uint8_t*)s.ring_other;

    for (int q = tid; q < LXW * LY; q += NT) {
        const int y = q / LXW, xw = q - y * LXW;
        const int gh = gy0 + y, gw = gx0 + xw * 8;
        uint64_t a = 0, b = 0, m = 0;
        if (gh >= 0 && gh < H && gw >= 0 && gw < (int)pitch) {
            const uint2 va = *reinterpret_cast<const uint2*>(src + (size_t)gh * pitch + gw);
            a = (uint64_t)va.x | ((uint64_t)va.y << 32);
            const bool interior_row = gh >= r && gh < H - r;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int c = gw + k;
                if (interior_row && c >= r && c < Wd - r) m |= (uint64_t)0xff << (8 * k);
            }
            if (m != ~0ull) {
                const uint2 vb = *reinterpret_cast<const uint2*>(oth + (size_t)gh * pitch + gw);
                b = (uint64_t)vb.x | ((uint64_t)vb.y << 32);
            }
        }
        A[q] = a;
        B[q] = b & ~m;
        M[q] = m;
    }
    __syncthreads();

#pragma unroll
    for (int j = 1; j <= K; ++j) {
        const uint64_t* __restrict__ P = (j & 1) ? A : B;
        uint64_t* __restrict__ Q = (j & 1) ? B : A;
        // every word of rows [j, LY - j); words beyond the loaded region read as
        // dead.  Wrong values enter only at the region edge and move one cell
        // per level, so after K <= 4 levels they stay inside the 8-cell halo.
        const int ny = LY - 2 * j;
        for (int q = tid; q < LXW * ny; q += NT) {
            const int y = j + q / LXW, x = q % LXW;
            const int i0 = y * LXW + x;
            const bool has_l = x > 0, has_r = x < LXW - 1;
            Seg up{P[i0 - LXW], 0, 0, 0}, mid{P[i0], 0, 0, 0}, dn{P[i0 + LXW], 0, 0, 0};
            up.l = has_l ? (uint32_t)(P[i0 - LXW - 1] >> 56) : 0u;
            up.r = has_r ? (uint32_t)(P[i0 - LXW + 1] & 0xffu) : 0u;
            mid.l = has_l ? (uint32_t)(P[i0 - 1] >> 56) : 0u;
            mid.r = has_r ? (uint32_t)(P[i0 + 1] & 0xffu) : 0u;
            dn.l = has_l ? (uint32_t)(P[i0 + LXW - 1] >> 56) : 0u;
            dn.r = has_r ? (uint32_t)(P[i0 + LXW + 1] & 0xffu) : 0u;
            const uint64_t ul = (up.lo << 8) | up.l, ur = (up.lo >> 8) | ((uint64_t)up.r << 56);
            const uint64_t ml = (mid.lo << 8) | mid.l, mr = (mid.lo >> 8) | ((uint64_t)mid.r << 56);
            const uint64_t dl = (dn.lo << 8) | dn.l, dr = (dn.lo >> 8) | ((uint64_t)dn.r << 56);
            const uint64_t cnt = ((ul + up.lo) + ur) + ((dl + dn.lo) + dr) + (ml + mr);
            const uint64_t nw = life_rule(cnt, mid.lo);
            const uint64_t m = M[i0];
            Q[i0] = (nw & m) | (Q[i0] & ~m);
        }
        __syncthreads();
    }

    const uint64_t* __restrict__ R = (K & 1) ? B : A;
    uint8_t* __restrict__ dst = (uint8_t*)s.dst;
    for (int q = tid; q < CXW * CY; q += NT) {
        const int y = q / CXW, xw = q % CXW;
        const int gh = h0 + y, gw = w0 + xw * 8;
        if (gh < r || gh >= H - r || gw >= Wd) continue;
        const int i0 = (y + K) * LXW + xw + 1;
        const uint64_t m = M[i0];
        if (m == 0) continue;
        const uint64_t v = R[i0];
        uint8_t* p = dst + (size_t)gh * pitch + gw;
        if (m == ~0ull) {
            uint2 x;
            x.x = (uint32_t)v;
            x.y = (uint32_t)(v >> 32);
            *reinterpret_cast<uint2*>(p) = x;
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((m >> (8 * k)) & 0xffu) p[k] = (uint8_t)(v >> (8 * k));
        }
    }
}

}  // namespace

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

hipError_t launch_life_ttile(const Sweep2D& s, int k, const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    constexpr int CXW = 30;  // 240 cells per tile row, 32 words loaded
    const int cy = cfg.tile_y > 0 ? cfg.tile_y : 32;
    const dim3 block(256);
#define PSKEL_LIFE_TT(KV, CYV)                                                                   \
    if (k == KV && cy == CYV) {                                                                  \
        const dim3 grid((s.W + CXW * 8 - 1) / (CXW * 8), (s.H + CYV - 1) / CYV);                 \
        hipLaunchKernelGGL((k_life_ttile<KV, CXW, CYV>), grid, block, 0, st, s);                 \
        return hipGetLastError();                                                                \
    }
    PSKEL_LIFE_TT(2, 16)
    PSKEL_LIFE_TT(2, 32)
    PSKEL_LIFE_TT(2, 64)
    PSKEL_LIFE_TT(3, 16)
    PSKEL_LIFE_TT(3, 32)
    PSKEL_LIFE_TT(3, 64)
    PSKEL_LIFE_TT(4, 16)
    PSKEL_LIFE_TT(4, 32)
    PSKEL_LIFE_TT(4, 64)
#undef PSKEL_LIFE_TT
    return hipErrorInvalidValue;
}

}  // namespace pskel
