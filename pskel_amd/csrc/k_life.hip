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

// ---------------------------------------------------------------------------
// bit-sliced temporal tile.  A wave owns a tile of 64 lanes x 32 columns and R
// rows; each lane keeps its 32 columns of every row as one 32-bit word (bit b
// = column x + b), packed from the u8 cells on load.  The wave runs K
// generations in registers:
//   * per row, the left and right neighbour words come from the same word
//     shifted by one, with the carry bit from the adjacent lane (a shuffle);
//   * horizontal 3-sums H (of the row above and below) and 2-sums M (of the
//     row itself, centre excluded) as two bit planes, each computed once per
//     row and generation;
//   * the 8-neighbour count H(y-1) + H(y+1) + M(y) by full adders, with a
//     flag for counts of 4 or more; B3/S23 is then
//         alive' = twos & ~flag & (ones | alive).
// Cells outside the grid read as dead; a tile's outer lanes and its K top and
// bottom rows are a halo whose errors move one cell per generation, so lanes
// 1..62 and rows [K, R-K) are exact for K <= 31.  The grid's ring (rows 0 and
// H-1, columns 0 and W-1) is never computed: after each generation a ring cell
// takes the ring of the buffer its level maps to (odd levels: the other
// buffer's, even: the current one's), restored only in tiles that hold ring
// cells.  No LDS and no barriers: waves run independently.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack_bits32(const uint4& a, const uint4& b)
{
    // byte i of x (0 or 1) lands on bit 21+i of x * 0x00204081 with no carries
    auto nib = [](uint32_t x) { return ((x * 0x00204081u) >> 21) & 0xFu; };
    return nib(a.x) | nib(a.y) << 4 | nib(a.z) << 8 | nib(a.w) << 12 | nib(b.x) << 16 | nib(b.y) << 20 |
           nib(b.z) << 24 | nib(b.w) << 28;
}

__device__ __forceinline__ uint32_t unpack_nibble(uint32_t w, int q)
{
    // bit i of the nibble lands on bit 8i of nibble * 0x00204081 with no carries
    return (((w >> (4 * q)) & 0xFu) * 0x00204081u) & 0x01010101u;
}

template <int K, int R>
__global__ __launch_bounds__(256) void k_life_bits(Sweep2D s, int ntx, int ntiles)
{
    static_assert(K >= 1 && K <= 31 && R > 2 * K, "tile taller than its halo");
    constexpr int CW = 62 * 32;  // core columns
    constexpr int CR = R - 2 * K;  // core rows
    constexpr int NB = (R + 31) / 32;
    const int lane = threadIdx.x % kWave;
    const int t = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (t >= ntiles) return;  // whole waves; nothing below synchronises
    const int tx = t % ntx, ty = t / ntx;
    const int H = s.H, W = s.W;
    const int x0 = tx * CW - 32, y0 = ty * CR - K;
    const int xl = x0 + lane * 32;  // this lane's first column (a multiple of 32)
    const size_t pitch = s.pitch;
    const uint8_t* __restrict__ src = (const uint8_t*)s.src;
    const uint8_t* __restrict__ oth = (const uint8_t*)s.ring_other;
    uint8_t* __restrict__ dst = (uint8_t*)s.dst;
    const bool lane_in = xl >= 0 && xl < (int)pitch;  // rows are padded to whole 1 KiB: 32 bytes stay inside

    // ring cells this tile holds (uniform across the wave)
    const bool edge_rows = y0 <= 0 || y0 + R >= H;
    const bool edge_cols = x0 <= 0 || x0 + 64 * 32 >= W;
    const uint32_t lm = xl == 0 ? 1u : 0u;  // column 0 at bit 0 of this lane
    const int bw = W - 1 - xl;
    const uint32_t rm = (bw >= 0 && bw < 32) ? (1u << bw) : 0u;  // column W-1
    uint32_t la[NB], lb[NB], ra[NB], rb[NB];  // ring-column bits per row: A (current) and B (other buffer)
    uint32_t at = 0, bt = 0, ab = 0, bb = 0;  // ring rows 0 and H-1: A and B words
#pragma unroll
    for (int q = 0; q < NB; ++q) la[q] = lb[q] = ra[q] = rb[q] = 0;

    // The tile's rows through one buffer resource from its first row in the
    // grid; rows outside the grid and lanes outside the padded row take an
    // offset past it and load zeros.  Rows go in groups of four: eight loads
    // in flight, and a scheduling barrier after each group so the compiler
    // does not hoist every group's loads (and their registers) to the top.
    uint32_t c[R];
    const int ylo = max(y0, 0), yhi = min(y0 + R, H);
    const auto rs = plane_rsrc(src + (size_t)ylo * pitch, yhi > ylo ? (int)((size_t)(yhi - ylo) * pitch) : 0);
    constexpr int BR = 4;
    static_assert(R % BR == 0, "whole groups of rows");
#pragma unroll
    for (int yb = 0; yb < R; yb += BR) {
        u32x4_t v[BR][2];
#pragma unroll
        for (int b = 0; b < BR; ++b) {
            const int gh = y0 + yb + b;
            const uint32_t o = (gh >= ylo && gh < yhi && lane_in) ? (uint32_t)((size_t)(gh - ylo) * pitch + xl)
                                                                  : kNoAccess;
            v[b][0] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)o, 0, 0);
            v[b][1] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(o + 16), 0, 0);
        }
#pragma unroll
        for (int b = 0; b < BR; ++b) {
            const uint4 x = make_uint4(v[b][0][0], v[b][0][1], v[b][0][2], v[b][0][3]);
            const uint4 e = make_uint4(v[b][1][0], v[b][1][1], v[b][1][2], v[b][1][3]);
            const int y = yb + b, gh = y0 + y;
            const uint32_t w = pack_bits32(x, e);
            c[y] = w;
            // ring values of both buffers, in the tiles that hold ring cells
            const bool row_in = gh >= 0 && gh < H;
            if (edge_rows && row_in && (gh == 0 || gh == H - 1)) {
                uint32_t wb = 0;
                if (lane_in) {
                    const uint4* p = reinterpret_cast<const uint4*>(oth + (size_t)gh * pitch + xl);
                    wb = pack_bits32(p[0], p[1]);
                }
                if (gh == 0) {
                    at = w;
                    bt = wb;
                }
                if (gh == H - 1) {  // also row 0 when H == 1
                    ab = w;
                    bb = wb;
                }
            }
            if (edge_cols && row_in) {
                const uint32_t bit = 1u << (y % 32);
                if (lm) {
                    la[y / 32] |= (w & 1u) ? bit : 0u;
                    lb[y / 32] |= oth[(size_t)gh * pitch] ? bit : 0u;
                }
                if (rm) {
                    ra[y / 32] |= (w & rm) ? bit : 0u;
                    rb[y / 32] |= oth[(size_t)gh * pitch + (W - 1)] ? bit : 0u;
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }

    // horizontal sums of a row: H = left + centre + right, M = left + right
    auto hm = [&](uint32_t x, uint32_t& h0, uint32_t& h1, uint32_t& m0, uint32_t& m1) {
        const uint32_t up = __shfl_up(x, 1, kWave);    // lane-1's word (lane 0: its own, a halo)
        const uint32_t dn = __shfl_down(x, 1, kWave);  // lane+1's word (lane 63: its own, a halo)
        const uint32_t l = (x << 1) | (up >> 31), r = (x >> 1) | (dn << 31);
        h0 = l ^ x ^ r;
        h1 = (l & x) | (r & (l ^ x));
        m0 = l ^ r;
        m1 = l & r;
    };
    // Generation j can be exact only on rows [j, R-j) (the halo's error
    // moves in one row per generation), from rows [j-1, R-j] of the level
    // before: the rest is skipped (uniform branches).
#pragma unroll 1
    for (int j = 1; j <= K; ++j) {
        const bool odd = j & 1;
        uint32_t hp0 = 0, hp1 = 0, hc0 = 0, hc1 = 0, mc0 = 0, mc1 = 0;
        if (j == 1) hm(c[0], hc0, hc1, mc0, mc1);
#pragma unroll
        for (int y = 0; y < R; ++y) {
            uint32_t hn0 = 0, hn1 = 0, mn0 = 0, mn1 = 0;
            if (y + 1 < R && y + 1 >= j - 1 && y + 1 <= R - j) hm(c[y + 1], hn0, hn1, mn0, mn1);
            if (y < j || y >= R - j) {
                hp0 = hc0;
                hp1 = hc1;
                hc0 = hn0;
                hc1 = hn1;
                mc0 = mn0;
                mc1 = mn1;
                continue;
            }
            const uint32_t s0 = hp0 ^ hn0 ^ mc0, c0 = (hp0 & hn0) | (mc0 & (hp0 ^ hn0));
            const uint32_t tt = hp1 ^ hn1 ^ mc1, tc = (hp1 & hn1) | (mc1 & (hp1 ^ hn1));
            const uint32_t s1 = tt ^ c0, f = tc | (tt & c0);
            uint32_t nw = s1 & ~f & (s0 | c[y]);
            const int gh = y0 + y;
            if (edge_rows && gh == 0) nw = odd ? bt : at;
            else if (edge_rows && gh == H - 1) nw = odd ? bb : ab;
            else if (edge_cols && gh > 0 && gh < H - 1) {
                const uint32_t lbit = ((odd ? lb : la)[y / 32] >> (y % 32)) & 1u;
                const uint32_t rbit = ((odd ? rb : ra)[y / 32] >> (y % 32)) & 1u;
                nw = (nw & ~(lm | rm)) | (lbit ? lm : 0u) | (rbit ? rm : 0u);
            }
            c[y] = nw;
            hp0 = hc0;
            hp1 = hc1;
            hc0 = hn0;
            hc1 = hn1;
            mc0 = mn0;
            mc1 = mn1;
        }
    }

    // core rows and lanes: interior cells only (the ring and the padding are never written)
    if (lane < 1 || lane > 62 || !lane_in) return;
    const bool full = xl >= 1 && xl + 32 <= W - 1;
#pragma unroll
    for (int y = K; y < R - K; ++y) {
        const int gh = y0 + y;
        if (gh < 1 || gh > H - 2) continue;
        uint8_t* p = dst + (size_t)gh * pitch + xl;
        const uint32_t w = c[y];
        if (full) {
            uint4 a, b;
            a.x = unpack_nibble(w, 0);
            a.y = unpack_nibble(w, 1);
            a.z = unpack_nibble(w, 2);
            a.w = unpack_nibble(w, 3);
            b.x = unpack_nibble(w, 4);
            b.y = unpack_nibble(w, 5);
            b.z = unpack_nibble(w, 6);
            b.w = unpack_nibble(w, 7);
            reinterpret_cast<uint4*>(p)[0] = a;
            reinterpret_cast<uint4*>(p)[1] = b;
        } else {
            for (int i = 0; i < 32; ++i) {
                const int col = xl + i;
                if (col >= 1 && col < W - 1) p[i] = (uint8_t)((w >> i) & 1u);
            }
        }
    }
}

}  // namespace

hipError_t launch_life_bits(const Sweep2D& s, int k, const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    // rows per lane: cfg.tile_y rounded up to 32, 40, 48 or 64, and at least 2k + 16
    int r = std::max(cfg.tile_y > 0 ? cfg.tile_y : 32, 2 * k + 16);
    r = r <= 32 ? 32 : (r <= 40 ? 40 : (r <= 48 ? 48 : 64));
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
#define PSKEL_LIFE_BITS_K(KV) \
    PSKEL_LIFE_BITS(KV, 32) PSKEL_LIFE_BITS(KV, 40) PSKEL_LIFE_BITS(KV, 48) PSKEL_LIFE_BITS(KV, 64)
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
