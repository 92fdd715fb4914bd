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
        const float up = __shfl_up(v[i].w, 1, kWave);
        const float dn = __shfl_down(v[i].x, 1, kWave);
        lf[i] = lane == 0 ? lfl[i] : up;
        rt[i] = lane == kWave - 1 ? rtl[i] : dn;
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
        const uint32_t up = __shfl_up((uint32_t)(g[i].hi >> 56), 1, kWave);
        const uint32_t dn = __shfl_down((uint32_t)(g[i].lo & 0xffu), 1, kWave);
        g[i].l = lane == 0 ? el[i] : up;
        g[i].r = lane == kWave - 1 ? er[i] : dn;
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


}  // namespace pskel
