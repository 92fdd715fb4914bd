// Register-tiled 2D sweep bodies shared by the standalone tuned kernels and the
// batched persistent kernel.  A block of (64, NWY) threads computes one tile.
#pragma once

#include "k_common.hpp"

namespace pskel {

// ---------------------------------------------------------------------------
// point evaluation over a register window (radius 1 patterns)
// ---------------------------------------------------------------------------
// Window access: row i (0..2 = h-1..h+1), column j in -1..4 relative to the
// lane's float4 (-1: left halo, 4: right halo).
struct Win {
    const float4* v;  // 3 rows
    const float* lf;  // 3 left halos
    const float* rt;  // 3 right halos
    __device__ __forceinline__ float at(int i, int j) const
    {
        if (j < 0) return lf[i];
        if (j > 3) return rt[i];
        const float4 x = v[i];
        return j == 0 ? x.x : j == 1 ? x.y : j == 2 ? x.z : x.w;
    }
};

template <int PAT, class PT>
__device__ __forceinline__ float eval_pat(const Win& W, int j, const PT& p)
{
    if constexpr (PAT == PAT_J5) {
        float acc = W.at(0, j) + W.at(2, j);
        acc = acc + W.at(1, j - 1);
        acc = acc + W.at(1, j + 1);
        return p.coeff * acc;
    } else {  // PAT_BOX9, row-major weighted sum
        float acc = p.wt[0] * W.at(0, j - 1);
        acc = acc + p.wt[1] * W.at(0, j);
        acc = acc + p.wt[2] * W.at(0, j + 1);
        acc = acc + p.wt[3] * W.at(1, j - 1);
        acc = acc + p.wt[4] * W.at(1, j);
        acc = acc + p.wt[5] * W.at(1, j + 1);
        acc = acc + p.wt[6] * W.at(2, j - 1);
        acc = acc + p.wt[7] * W.at(2, j);
        acc = acc + p.wt[8] * W.at(2, j + 1);
        return acc;
    }
}

// ---------------------------------------------------------------------------
// tuned tile: 64 lanes x float4 = 256 columns per wave, RY rows per lane;
// block = (64, NWY), the NWY waves stack along h; tile (tx, ty) covers columns
// [256 tx, 256 tx + 256) and rows [NWY RY ty, NWY RY (ty + 1)).
// ---------------------------------------------------------------------------
template <int PAT, int RY, class PT>
__device__ __forceinline__ void tile_f32(const Sweep2D& s, const PT& p, int tx, int ty)
{
    const int lane = threadIdx.x;
    const int w0 = tx * 256;
    const int w = w0 + lane * 4;
    const int hb = (ty * (int)blockDim.y + (int)threadIdx.y) * RY;  // first output row
    const int H = s.H, Wd = s.W, r = s.radius;
    const float* __restrict__ src = (const float*)s.src;
    float* __restrict__ dst = (float*)s.dst;
    const size_t pitch = s.pitch;
    const bool col_live = w < Wd;

    float4 v[RY + 2];
    float lfl[RY + 2], rtl[RY + 2];
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) {
        const int h = hb - 1 + i;
        v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        lfl[i] = 0.f;
        rtl[i] = 0.f;
        if (h >= 0 && h < H) {
            const float* row = src + (size_t)h * pitch;
            if (col_live) v[i] = *reinterpret_cast<const float4*>(row + w);
            if (lane == 0 && w0 > 0) lfl[i] = row[w0 - 1];
            if (lane == kWave - 1 && w0 + 256 < Wd) rtl[i] = row[w0 + 256];
        }
    }
    float lf[RY + 2], rt[RY + 2];
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) {
        lf[i] = from_lower_lane(v[i].w, lfl[i]);  // lane 0: the cell loaded for it
        rt[i] = from_upper_lane(v[i].x, rtl[i]);  // lane 63: likewise
    }
    if (!col_live) return;
    const bool full = (w >= r) && (w + 3 < Wd - r);
#pragma unroll
    for (int i = 1; i <= RY; ++i) {
        const int h = hb - 1 + i;
        if (h < r || h >= H - r) continue;
        const Win win{&v[i - 1], &lf[i - 1], &rt[i - 1]};
        float o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = eval_pat<PAT, PT>(win, j, p);
        float* row = dst + (size_t)h * pitch + w;
        if (full) {
            *reinterpret_cast<float4*>(row) = make_float4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (w + j >= r && w + j < Wd - r) row[j] = o[j];
        }
    }
}

// ---------------------------------------------------------------------------
// SWAR helpers.  Byte k of a little-endian word is cell w0 + k.
// ---------------------------------------------------------------------------
constexpr uint64_t kOnes = 0x0101010101010101ull;
constexpr uint64_t kHigh = 0x8080808080808080ull;
constexpr uint64_t kLow7 = 0x7f7f7f7f7f7f7f7full;

// 0x01 in every byte equal to k (byte values <= 11, so no carry leaves a byte).
__device__ __forceinline__ uint64_t bytes_eq(uint64_t x, uint64_t k)
{
    const uint64_t y = x ^ (k * kOnes);
    const uint64_t nz = (y + kLow7) & kHigh;
    return ((~nz) & kHigh) >> 7;
}

__device__ __forceinline__ uint64_t life_rule(uint64_t cnt, uint64_t cur)
{
    return bytes_eq(cnt, 3) | (bytes_eq(cnt, 2) & cur);
}

// A 16-cell lane segment: lo = cells 0..7, hi = cells 8..15, plus the byte to
// the left (cell -1) and to the right (cell 16).
struct Seg {
    uint64_t lo, hi;
    uint32_t l, r;
};

// Horizontal sums for a row segment: with_c = left + centre + right,
// no_c = left + right (bytewise, <= 3).
__device__ __forceinline__ void hsum(const Seg& s, uint64_t& lo3, uint64_t& hi3, uint64_t& lo2,
                                     uint64_t& hi2)
{
    const uint64_t lo_l = (s.lo << 8) | (uint64_t)s.l;
    const uint64_t hi_l = (s.hi << 8) | (s.lo >> 56);
    const uint64_t lo_r = (s.lo >> 8) | (s.hi << 56);
    const uint64_t hi_r = (s.hi >> 8) | ((uint64_t)s.r << 56);
    lo2 = lo_l + lo_r;
    hi2 = hi_l + hi_r;
    lo3 = lo2 + s.lo;
    hi3 = hi2 + s.hi;
}

// ---------------------------------------------------------------------------
// SWAR kernel: wave = 64 lanes x 16 cells = 1024 columns; block = (64, NWY).
// ---------------------------------------------------------------------------
template <int RY>
__device__ __forceinline__ void tile_life(const Sweep2D& s, int tx, int ty)
{
    const int lane = threadIdx.x;
    const int w0 = tx * 1024;
    const int w = w0 + lane * 16;
    const int hb = (ty * (int)blockDim.y + (int)threadIdx.y) * RY;
    const int H = s.H, Wd = s.W, r = s.radius;
    const size_t pitch = s.pitch;
    const uint8_t* __restrict__ src = (const uint8_t*)s.src;
    uint8_t* __restrict__ dst = (uint8_t*)s.dst;
    const bool col_live = w < Wd;

    Seg g[RY + 2];
    uint32_t el[RY + 2], er[RY + 2];
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) {
        const int h = hb - 1 + i;
        g[i].lo = 0;
        g[i].hi = 0;
        el[i] = 0;
        er[i] = 0;
        if (h >= 0 && h < H) {
            const uint8_t* row = src + (size_t)h * pitch;
            if (col_live) {
                const uint4 x = *reinterpret_cast<const uint4*>(row + w);
                g[i].lo = (uint64_t)x.x | ((uint64_t)x.y << 32);
                g[i].hi = (uint64_t)x.z | ((uint64_t)x.w << 32);
            }
            if (lane == 0 && w0 > 0) el[i] = row[w0 - 1];
            if (lane == kWave - 1 && w0 + 1024 < Wd) er[i] = row[w0 + 1024];
        }
    }
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) {
        g[i].l = from_lower_lane((uint32_t)(g[i].hi >> 56), el[i]);
        g[i].r = from_upper_lane((uint32_t)(g[i].lo & 0xffu), er[i]);
    }
    if (!col_live) return;

    uint64_t s3lo[RY + 2], s3hi[RY + 2], s2lo[RY + 2], s2hi[RY + 2];
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) hsum(g[i], s3lo[i], s3hi[i], s2lo[i], s2hi[i]);

    const bool full = (w >= r) && (w + 15 < Wd - r);
#pragma unroll
    for (int i = 1; i <= RY; ++i) {
        const int h = hb - 1 + i;
        if (h < r || h >= H - r) continue;
        const uint64_t clo = s3lo[i - 1] + s3lo[i + 1] + s2lo[i];
        const uint64_t chi = s3hi[i - 1] + s3hi[i + 1] + s2hi[i];
        const uint64_t nlo = life_rule(clo, g[i].lo);
        const uint64_t nhi = life_rule(chi, g[i].hi);
        uint8_t* row = dst + (size_t)h * pitch + w;
        if (full) {
            uint4 x;
            x.x = (uint32_t)nlo;
            x.y = (uint32_t)(nlo >> 32);
            x.z = (uint32_t)nhi;
            x.w = (uint32_t)(nhi >> 32);
            *reinterpret_cast<uint4*>(row) = x;
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (w + j >= r && w + j < Wd - r)
                    row[j] = (uint8_t)((j < 8 ? (nlo >> (8 * j)) : (nhi >> (8 * (j - 8)))) & 0xffu);
            }
        }
    }
}


// ---------------------------------------------------------------------------
// temporal tile: core CX x CY cells, loaded region (CX + 8) x (CY + 2K) with a
// 4-column aligned halo (K <= 4), kgen <= K sweeps in an LDS ping-pong (A, B:
// ttile_lds floats each); 256 threads, tid = 0..255, all of which must call it
// (it synchronises the block).  Tile (bx, by) has its core at columns
// [CX bx, CX bx + CX) and rows [CY by, CY by + CY).  Cells outside the
// interior are never computed: at level j an LDS buffer holds the ring of the
// global buffer that level j lives in (level parity alternates between src and
// ring_other), so halo reads near the grid edge see exactly what a single
// sweep would see.  Level j is exact on the core widened by K - j cells, so
// fewer levels than K still leave the core exact.
// ---------------------------------------------------------------------------
template <int K, int CX, int CY>
__host__ __device__ constexpr int ttile_lds()
{
    return (CY + 2 * K) * (CX + 8);
}

template <int PAT, int K, int CX, int CY, class PT>
__device__ __forceinline__ void ttile_body(const Sweep2D& s, const PT& p, int bx, int by, int kgen, int tid,
                                           float* __restrict__ A, float* __restrict__ B)
{
    constexpr int LX = CX + 8;
    constexpr int LY = CY + 2 * K;
    constexpr int NT = 256;
    const int w0 = bx * CX;                    // core origin
    const int h0 = by * CY;
    const int gx0 = w0 - 4, gy0 = h0 - K;      // loaded region origin
    const int H = s.H, Wd = s.W, r = s.radius;
    const size_t pitch = s.pitch;
    const float* __restrict__ src = (const float*)s.src;
    const float* __restrict__ oth = (const float*)s.ring_other;

    // load: float4 granules, LX/4 per row
    constexpr int G = LX / 4;
    for (int q = tid; q < G * LY; q += NT) {
        const int y = q / G, x4 = (q - y * G) * 4;
        const int gh = gy0 + y, gw = gx0 + x4;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
        if (gh >= 0 && gh < H && gw >= 0 && gw < (int)pitch) {
            a = *reinterpret_cast<const float4*>(src + (size_t)gh * pitch + gw);
            const bool ring_row = gh < r || gh >= H - r;
            const bool ring_col = gw < r || gw + 3 >= Wd - r;
            if (ring_row || ring_col) b = *reinterpret_cast<const float4*>(oth + (size_t)gh * pitch + gw);
        }
        *reinterpret_cast<float4*>(&A[y * LX + x4]) = a;
        *reinterpret_cast<float4*>(&B[y * LX + x4]) = b;
    }
    __syncthreads();

    // Level j: each thread walks one column x of the level's region down one
    // of NS strips of rows, keeping the row above and its own row (x-1, x, x+1)
    // in registers, so a cell costs three LDS reads (the row below).  Sums keep
    // the mask order.
    constexpr int NS = NT / 128;  // strips per column
    static_assert(LX <= 128, "one thread per column of the loaded region");
    // a strip is two whole waves, so its row range is uniform across the wave
    const int col = tid % 128, strip = __builtin_amdgcn_readfirstlane(tid / 128);
#pragma unroll
    for (int j = 1; j <= K; ++j) {
        if (j > kgen) break;  // uniform
        const float* __restrict__ P = (j & 1) ? A : B;
        float* __restrict__ Q = (j & 1) ? B : A;
        const int xlo = 4 - (K - j), xhi = 4 + CX + (K - j);
        const int ylo = j, yhi = LY - j;
        const int nx = xhi - xlo, ny = yhi - ylo;
        const int rows = (ny + NS - 1) / NS;
        const int x = xlo + col;
        const int gw = gx0 + x;
        // only interior cells are written: a ring column's lane and the ring
        // rows of the strip compute nothing (their LDS cells keep the ring)
        if (col < nx && gw >= r && gw < Wd - r) {
            const int ys = ylo + strip * rows;
            const int ya = max(ys, r - gy0), yb = min(min(ys + rows, yhi), H - r - gy0);
            const float* c = P + (ya - 1) * LX + x;
            float um = c[-1], u0 = c[0], up = c[1];                 // row y - 1
            float cm = c[LX - 1], c0 = c[LX], cp = c[LX + 1];       // row y
#pragma unroll 3
            for (int y = ya; y < yb; ++y) {
                c += LX;
                const float dm = c[LX - 1], d0 = c[LX], dp = c[LX + 1];  // row y + 1
                float acc;
                if constexpr (PAT == PAT_J5) {
                    acc = u0 + d0;
                    acc = acc + cm;
                    acc = acc + cp;
                    acc = p.coeff * acc;
                } else {
                    acc = p.wt[0] * um;
                    acc = acc + p.wt[1] * u0;
                    acc = acc + p.wt[2] * up;
                    acc = acc + p.wt[3] * cm;
                    acc = acc + p.wt[4] * c0;
                    acc = acc + p.wt[5] * cp;
                    acc = acc + p.wt[6] * dm;
                    acc = acc + p.wt[7] * d0;
                    acc = acc + p.wt[8] * dp;
                }
                Q[y * LX + x] = acc;
                um = cm; u0 = c0; up = cp;
                cm = dm; c0 = d0; cp = dp;
            }
        }
        __syncthreads();
    }

    const float* __restrict__ R = (kgen & 1) ? B : A;
    float* __restrict__ dst = (float*)s.dst;
    for (int q = tid; q < CX * CY; q += NT) {
        const int y = q / CX, x = q % CX;
        const int gh = h0 + y, gw = w0 + x;
        if (gh < r || gh >= H - r || gw < r || gw >= Wd - r) continue;
        dst[(size_t)gh * pitch + gw] = R[(y + K) * LX + (x + 4)];
    }
    // the caller may reuse A and B for its next tile
    __syncthreads();
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

// life_bits_tile: one wave (lane 0..63) evolves bit tile t of a grid tiled for
// K generations (tiles along w: ntx); kgen <= K generations run (the core rows
// [K, R-K) are exact after any kgen <= K) and level kgen is written to dst.
template <int K, int R>
__device__ __forceinline__ void life_bits_tile(const Sweep2D& s, int ntx, int t, int lane, int kgen)
{
    static_assert(K >= 1 && K <= 31 && R > 2 * K, "tile taller than its halo");
    constexpr int CW = 62 * 32;  // core columns
    constexpr int CR = R - 2 * K;  // core rows
    constexpr int NB = (R + 31) / 32;
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
        const uint32_t up = from_lower_lane(x, x);  // lane-1's word (lane 0: its own, a halo)
        const uint32_t dn = from_upper_lane(x, x);  // lane+1's word (lane 63: its own, a halo)
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
    for (int j = 1; j <= kgen; ++j) {
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


}  // namespace pskel
