// DEPTH fused sweeps per launch (temporal depth 2 or 3) of the 7-point
// scaled sum with COPY faces, f32, bf16 or f16 storage.
//
// A block owns an extended (w, d) tile and marches along h.  At plane h it
// computes the first sweep s1(h) on every cell of the tile from the input
// planes h-1, h, h+1, then each later sweep l one plane behind the one
// before, s_l(h-l+1) from s_(l-1)(h-l), s_(l-1)(h-l+1), s_(l-1)(h-l+2), on the
// cells whose neighbours all lie in the tile, and on grid faces, which copy.
// The last sweep's plane h-DEPTH+1 is stored.  Every cell of every sweep
// evaluates
//     coeff * (((((h- + h+) + w-) + w+) + d-) + d+)
// in f32 with no contraction, so a fused launch equals DEPTH single sweeps bit
// for bit.  Each segment of h planes recomputes the earlier sweeps on the
// DEPTH-1 planes before and after it.
//
// Tile.  The block's waves form NWW rows of NWD waves.  A wave holds RW rows
// of 64 lanes x VD cells, so the tile is TW = NWW*RW rows by TD = NWD*64*VD
// cells.  Along w the two rows just outside the tile are loaded as a halo
// (by the top and bottom row of waves), s1 is exact on all TW rows, each later
// sweep on one row fewer at each edge, and the tile outputs the TW-2(DEPTH-1)
// rows inside; neighbouring tiles overlap by 2(DEPTH-1) rows.  Along d no
// halo is loaded: tiles overlap by one lane vector on each side, which covers
// the DEPTH-1 cells each edge loses (DEPTH <= VD).  Each output cell is
// written by exactly one tile.
//
// Exchange.  Planes along h sit in registers.  A wave's interior rows see
// their w neighbours in its own registers; only its top and bottom rows go
// through LDS, with the halo rows, and only the first and last cell of each
// row cross to the neighbouring wave along d.  Lane neighbours along d come
// from lane shuffles.  LDS is double-buffered: one barrier per plane, and
// consecutive lanes touch consecutive 8- or 16-byte words (no bank conflicts).
//
// 16-bit storage.  The register rings hold storage words, two cells per 32-bit
// register, and decode on use; results are rounded two at a time into such a
// register, and faces are merged in by a per-register bit mask.
// A first-sweep value is rounded to storage (to nearest, or stochastically
// with the word of its own sweep index, plane and cell) before the second
// sweep reads it, so the pair equals two stored sweeps bit for bit; faces
// keep the input word through both sweeps.
//
// Global memory goes through buffer instructions whose range check drops what
// must not happen: a plane outside the segment's reach gets a resource of 0
// bytes (loads return 0, stores are discarded), and a row or vector that must
// not be read or written an offset past the plane.  So every step issues the
// same loads and stores, the compiler counts them exactly, and the wait for
// plane h+1 never drains the planes prefetched after it.
#pragma once

#include <atomic>
#include <type_traits>

#include "k_common.hpp"

namespace pskel {

struct March2Shape {
    int vd, rw, nwd, nww, pf;
    int depth;     // fused sweeps per launch (2 or 3)
    int hseg;      // h planes per block; 0: chosen to fill the blocks the device holds at once
    int ntd, ntw;  // tiles along d and w
};

inline int ceil_div_i(int a, int b) { return (a + b - 1) / b; }

// cfg.tile_x = cells per lane (VD), tile_y = rows per wave (RW), tile_z = h
// planes per block (0: chosen from the device's resident blocks), block_x =
// threads per block (64 * NWD * NWW), block_y = waves along d (NWD) + 16 *
// (planes prefetched beyond h+1, minus 1); depth = sweeps of the launch
inline March2Shape march2_shape(const Sweep3D& s, const pskel_launch_cfg_t& cfg, int depth = 2)
{
    March2Shape m;
    m.depth = depth;
    // default at depth 2: 18-row tiles of 2 x 6 waves (16 rows advanced per
    // tile, so a 512-wide grid takes 32 tiles and 8 segments fill the 256 CUs
    // of an MI355X); at depth 3: 16-row tiles of 2 x 4 waves of 4 rows (10-25%
    // faster than the 12-row tiles on the grids measured, 3% slower at 512^3,
    // where those fill the CUs exactly)
    m.vd = cfg.tile_x > 0 ? cfg.tile_x : 4;
    m.rw = cfg.tile_y > 0 ? cfg.tile_y : (depth == 3 ? 4 : 3);
    const int nw = cfg.block_x > 0 ? cfg.block_x / kWave : (depth == 3 ? 8 : 12);
    m.nwd = cfg.block_y > 0 ? cfg.block_y % 16 : 2;
    m.pf = cfg.block_y > 0 ? 1 + cfg.block_y / 16 : 1;
    m.nww = m.nwd > 0 ? nw / m.nwd : 0;
    m.hseg = cfg.tile_z > 0 ? cfg.tile_z : 0;
    const int TD = m.nwd * kWave * m.vd, TW = m.nww * m.rw;
    m.ntd = s.D <= TD ? 1 : 1 + ceil_div_i(s.D - TD, TD - 2 * m.vd);
    m.ntw = s.W <= TW ? 1 : 1 + ceil_div_i(s.W - TW, TW - 2 * (depth - 1));
    return m;
}

// the instantiated shapes (vd, rw, nwd, nww, pf): at depth 2 eleven for f32 and
// five for 16-bit storage (rows of 512, 256 and 128 cells along d); at depth 3
// two for f32, with tiles of 512 cells: 16 rows of 2 x 4 waves of 4 rows, and
// 12 rows of 2 x 6 waves of 2 rows (the 12-wave shape of 3 rows spills at
// depth 3); each also at depth 2 for the pair that closes a run
inline bool march2_supported(const March2Shape& m, int storage)
{
    auto key = [](int vd, int rw, int nwd, int nww, int pf) {
        return ((((vd * 16 + rw) * 16 + nwd) * 16) + nww) * 16 + pf;
    };
    const int k = key(m.vd, m.rw, m.nwd, m.nww, m.pf);
    if (m.depth == 3)  // 8 waves: three register rings of 4 rows x 4 cells fit 2 waves per SIMD
        return storage == PSKEL_DTYPE_F32 && (k == key(4, 4, 2, 4, 1) || k == key(4, 2, 2, 6, 1));
    if (m.depth != 2) return false;
    if (storage != PSKEL_DTYPE_F32)
        return k == key(4, 3, 2, 6, 1) || k == key(4, 2, 2, 8, 1) || k == key(4, 3, 1, 12, 1) ||
               k == key(2, 3, 1, 12, 1) || k == key(4, 2, 2, 6, 1);
    return k == key(2, 4, 4, 4, 1) || k == key(2, 4, 4, 2, 1) || k == key(4, 2, 2, 8, 1) ||
           k == key(4, 2, 2, 4, 1) || k == key(2, 4, 4, 4, 2) || k == key(2, 4, 2, 4, 1) ||
           k == key(4, 4, 2, 4, 1) || k == key(4, 3, 2, 6, 1) || k == key(4, 3, 1, 12, 1) ||
           k == key(2, 3, 1, 12, 1) || k == key(4, 2, 2, 6, 1);
}

hipError_t march3d_t2_f32(const Sweep3D& s, const Points& p, const March2Shape& m, hipStream_t st);
hipError_t march3d_t2_bf16(const Sweep3D& s, const Points& p, const March2Shape& m, hipStream_t st);
hipError_t march3d_t2_f16(const Sweep3D& s, const Points& p, const March2Shape& m, hipStream_t st);
// the grid of those launches
size_t march3d_t2_f32_blocks(const Sweep3D& s, const March2Shape& m);
size_t march3d_t2_bf16_blocks(const Sweep3D& s, const March2Shape& m);
size_t march3d_t2_f16_blocks(const Sweep3D& s, const March2Shape& m);

#ifdef PSKEL_MARCH_INSTANTIATE
namespace {

template <int VD>
struct LVec;  // one lane's vector as an LDS word
template <>
struct LVec<2> {
    using T = float2;
};
template <>
struct LVec<4> {
    using T = float4;
};

// VD consecutive elements of T (4, 8 or 16 bytes) through a buffer resource
template <typename T, int VD>
__device__ __forceinline__ void buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, T (&v)[VD])
{
    constexpr int B = VD * (int)sizeof(T);
    static_assert(B == 4 || B == 8 || B == 16, "lane vector of 4, 8 or 16 bytes");
    if constexpr (B == 4) {
        const uint32_t x = __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
        __builtin_memcpy(v, &x, 4);
    } else if constexpr (B == 8) {
        const u32x2_t x = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
        __builtin_memcpy(v, &x, 8);
    } else {
        const u32x4_t x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
        __builtin_memcpy(v, &x, 16);
    }
}

// Cache policy of the output stores: non-temporal (gfx950 `nt`, bit 1 of the
// buffer intrinsics' aux operand).  No block of the launch reads the output
// grid; measured 1.5% faster for f32 and 0.7% for bf16 at 512^3.
constexpr int kStoreNT = 2;

template <typename T, int VD>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, const T (&v)[VD])
{
    constexpr int B = VD * (int)sizeof(T);
    static_assert(B == 4 || B == 8 || B == 16, "lane vector of 4, 8 or 16 bytes");
    if constexpr (B == 4) {
        uint32_t x;
        __builtin_memcpy(&x, v, 4);
        __builtin_amdgcn_raw_buffer_store_b32(x, r, (int)off, 0, kStoreNT);
    } else if constexpr (B == 8) {
        u32x2_t x;
        __builtin_memcpy(&x, v, 8);
        __builtin_amdgcn_raw_buffer_store_b64(x, r, (int)off, 0, kStoreNT);
    } else {
        u32x4_t x;
        __builtin_memcpy(&x, v, 16);
        __builtin_amdgcn_raw_buffer_store_b128(x, r, (int)off, 0, kStoreNT);
    }
}

// x when A, else y (references of one type; the other is unused)
template <bool A, typename X, typename Y>
__device__ __forceinline__ auto& either(X& x, Y& y)
{
    if constexpr (A) return x;
    else return y;
}

template <int VD>
__device__ __forceinline__ void lds_put(typename LVec<VD>::T* p, const float (&v)[VD])
{
    typename LVec<VD>::T x;
    __builtin_memcpy(&x, v, sizeof(x));
    *p = x;
}

template <int VD>
__device__ __forceinline__ void lds_get(const typename LVec<VD>::T* p, float (&v)[VD])
{
    const typename LVec<VD>::T x = *p;
    __builtin_memcpy(v, &x, sizeof(x));
}

typedef float f32x2_t __attribute__((ext_vector_type(2)));

// A lane's VD cells as storage words: f32 one cell per word; bf16 and f16 two
// cells per 32-bit word, cell 2k in the low half of word k and 2k+1 in the
// high half (the order of memory).
template <int ST, int VD>
struct LaneWords {
    static constexpr int CPW = ST == PSKEL_DTYPE_F32 ? 1 : 2;  // cells per word
    static constexpr int N = VD / CPW;
    using W = std::conditional_t<ST == PSKEL_DTYPE_F32, float, uint32_t>;

    // cell j as stored (u - offset for 16-bit storage)
    __device__ __forceinline__ static float raw(const W (&w)[N], int j)
    {
        if constexpr (ST == PSKEL_DTYPE_F32) {
            return w[j];
        } else if constexpr (ST == PSKEL_DTYPE_BF16) {
            return __uint_as_float((j & 1) ? (w[j >> 1] & 0xffff0000u) : (w[j >> 1] << 16));
        } else {
            return __half2float(__ushort_as_half((uint16_t)((j & 1) ? (w[j >> 1] >> 16) : w[j >> 1])));
        }
    }
    // v rounded to storage, to nearest even (as Stor<ST>::enc_raw)
    __device__ __forceinline__ static void put(W (&w)[N], const float (&v)[VD])
    {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if constexpr (ST == PSKEL_DTYPE_F32) {
                w[k] = v[k];
            } else {
                // f16: each value rounded to f32 first (as_rounded_f32)
                const f32x2_t f = ST == PSKEL_DTYPE_F16 ? f32x2_t{as_rounded_f32(v[2 * k]), as_rounded_f32(v[2 * k + 1])}
                                                        : f32x2_t{v[2 * k], v[2 * k + 1]};
                if constexpr (ST == PSKEL_DTYPE_BF16) w[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2_t));
                else w[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(f, f16x2_t));
            }
        }
    }
    // v rounded stochastically with the random word r[j] of each cell (as
    // Stor<ST>::enc_sr_raw); 16-bit only
    __device__ __forceinline__ static void put_sr(W (&w)[N], const float (&v)[VD], const uint32_t (&r)[VD])
    {
        static_assert(ST != PSKEL_DTYPE_F32, "stochastic rounding is for 16-bit storage");
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const float a = v[2 * k], b = v[2 * k + 1];
            // the first conversion replaces the low half of a's own register,
            // the second the high half
            if constexpr (ST == PSKEL_DTYPE_BF16) {
                bf16x2_t o = __builtin_bit_cast(bf16x2_t, a);
                o = __builtin_amdgcn_cvt_sr_bf16_f32(o, a, r[2 * k], false);
                o = __builtin_amdgcn_cvt_sr_bf16_f32(o, b, r[2 * k + 1], true);
                w[k] = __builtin_bit_cast(uint32_t, o);
            } else {
                f16x2_t o = __builtin_bit_cast(f16x2_t, a);
                o = __builtin_amdgcn_cvt_sr_f16_f32(o, a, r[2 * k], false);
                o = __builtin_amdgcn_cvt_sr_f16_f32(o, b, r[2 * k + 1], true);
                w[k] = __builtin_bit_cast(uint32_t, o);
            }
        }
    }
    // the bits of `keep` where m is set, of `w` elsewhere
    __device__ __forceinline__ static W merge(uint32_t m, W keep, W w)
    {
        const uint32_t x = (__builtin_bit_cast(uint32_t, keep) & m) | (__builtin_bit_cast(uint32_t, w) & ~m);
        return __builtin_bit_cast(W, x);
    }
    // the mask of cell j within its word
    static constexpr uint32_t cell_mask(int j) { return CPW == 1 ? ~0u : 0xffffu << (16 * (j % CPW)); }
};

template <int ST, int VD, int RW, int NWD, int NWW, int PF, int DEPTH, bool RES, bool SR>
__global__ __launch_bounds__(NWD * NWW * kWave) void k3d_march2(Sweep3D s, float coeff, int hseg, int ntd, int ntw)
{
    using T = typename Stor<ST>::T;  // storage element
    using LW = LaneWords<ST, VD>;
    using Word = typename LW::W;  // register word: one f32 cell or two 16-bit cells
    constexpr int NV = LW::N;  // words per lane vector
    constexpr int NW = NWD * NWW;
    constexpr int TD = NWD * kWave * VD;
    constexpr int TW = NWW * RW;
    constexpr int NL = NWD * kWave;  // LDS words per row
    constexpr int NS = DEPTH - 1;  // sweeps kept in register rings (all but the last)
    static_assert(RW >= 2 && NWW >= 2 && (VD == 2 || VD == 4), "fused tile shape");
    static_assert(DEPTH >= 2 && DEPTH <= VD, "tiles overlap by one lane vector along d");
    using LV = typename LVec<VD>::T;
    // input plane h: each row of waves' top and bottom row, then the halo rows above and below the tile
    __shared__ LV lin[2][2 * NWW + 2][NL];
    __shared__ LV ls[NS][2][2 * NWW][NL];         // s_l(h-l): each row of waves' top and bottom row
    __shared__ float lein[2][NWW][NWD][RW][2];    // first and last cell of each wave's rows, input plane h
    __shared__ float les[NS][2][NWW][NWD][RW][2];  // the same for s_l(h-l)
    __shared__ double red[NW];

    const int lane = threadIdx.x % kWave;
    // uniform across the wave, so rows, tile edges and LDS slots are scalars
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int wd = wave % NWD, ww = wave / NWD;
    const uint32_t lid = xcd_remap(blockIdx.x, gridDim.x);

    const int W = s.W, H = s.H, D = s.D, r = s.radius;
    const int pitch = (int)s.pitch;
    const size_t plane = (size_t)W * s.pitch;
    // 16-bit storage: the stored values x = u - off are summed as they are and
    // the result is fma(coeff, sum, c0) (scaled_sum_offset_term)
    const float c0 = s.c0;
    // stochastic rounding: the keys of the launch's sweeps
    uint32_t keys[DEPTH];
#pragma unroll
    for (int l = 0; l < DEPTH; ++l) keys[l] = SR ? sweep_key(s.seed, s.sweep + l) : 0u;
    double rr = 0.0;

    // one segment: planes [h0, h1) of tile (td, tw)
    auto segment = [&](int td, int tw, int h0, int h1) __attribute__((always_inline)) {
        // the extended tile and the cells it outputs
        const int ed0 = td * (TD - 2 * VD);
        const int od0 = td == 0 ? 0 : ed0 + VD;
        const int od1 = ed0 + TD >= D ? D : ed0 + TD - VD;
        const int ew0 = tw * (TW - 2 * NS);
        const int ow0 = tw == 0 ? 0 : ew0 + NS;
        const int ow1 = ew0 + TW >= W ? W : ew0 + TW - NS;

        const int col = wd * kWave + lane;  // LDS word of this lane in a row
        const int d = ed0 + col * VD;
        const int wr = ew0 + ww * RW;
        const int hlo = max(h0 - DEPTH, 0), hhi = min(h1 + DEPTH, H);  // input planes read
        // sweep l (1-based) runs on planes [h0 - (DEPTH - l), h1 + (DEPTH - l)) within the grid;
        // step h computes its plane h - l + 1, and the last sweep's plane h1 - 1 at step h1 + DEPTH - 2
        const int ha = max(h0 - NS, 0);
        const int hend = h1 + DEPTH - 2;

        const bool dlive = d < D;
        const bool dout = dlive && d >= od0 && d < od1;
        const bool dint = d >= r && d + VD - 1 < D - r;
        // cells of this lane on a d face, as a bit mask per register word
        uint32_t dmask[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) dmask[k] = 0u;
#pragma unroll
        for (int j = 0; j < VD; ++j)
            if (!dint && (d + j < r || d + j >= D - r)) dmask[j / LW::CPW] |= LW::cell_mask(j);
        bool row_int[RW], row_out[RW];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int w = wr + i;
            row_int[i] = w >= r && w < W - r;
            row_out[i] = dout && w < W && w >= ow0 && w < ow1;
        }
        // every cell of this wave interior along w and d (wave-uniform): on an
        // interior plane its results need no face merge
        const int dw0 = ed0 + wd * kWave * VD;
        const bool wave_int = dw0 >= r && dw0 + kWave * VD <= D - r && wr >= r && wr + RW <= W - r;
        // the halo row this row of waves loads (top: the row above the tile, bottom: the row below)
        const bool top = ww == 0, bottom = ww == NWW - 1;
        const int hrow = top ? ew0 - 1 : ew0 + TW;
        const int hslot = top ? 2 * NWW : 2 * NWW + 1;

        // byte offsets of this lane's vectors inside a plane (plane bytes < 2^31)
        const int plane_bytes = (int)(plane * sizeof(T));
        uint32_t ld_off[RW], st_off[RW];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const uint32_t o = (uint32_t)(((size_t)(wr + i) * pitch + d) * sizeof(T));
            ld_off[i] = (dlive && wr + i < W) ? o : kNoAccess;
            st_off[i] = row_out[i] ? o : kNoAccess;
        }
        const uint32_t hl_off = ((top || bottom) && dlive && hrow >= 0 && hrow < W)
                                    ? (uint32_t)(((size_t)hrow * pitch + d) * sizeof(T))
                                    : kNoAccess;

        // stochastic rounding: the in-plane words A of this lane's cells,
        // hashed once and shared by every plane and both sweeps
        uint32_t ra[RW][VD];
        if constexpr (SR) {
            const uint32_t ckey = sr_cell_key(s.seed);
#pragma unroll
            for (int i = 0; i < RW; ++i)
#pragma unroll
                for (int j = 0; j < VD; ++j) ra[i][j] = sr_cell_bits(ckey, (uint32_t)(wr + i) * D + d + j);
        }

        constexpr int NQ = PF + 3;  // ring slots: planes h-1 .. h+PF+1
        constexpr bool F32 = ST == PSKEL_DTYPE_F32;
        Word in[NQ][RW][NV];        // input planes as loaded (storage words)
        Word hal[NQ][NV];           // input planes of the halo row
        Word sw[NS][NQ][RW][NV];    // sweep l+1's planes, storage words
        // the rings as floats: 16-bit storage decodes each plane once (input
        // planes h-1, h, h+1; sweep l+1's planes h-l-2, h-l-1, h-l); f32 words
        // are the floats, so the same registers serve (no copies to carry)
        float fin_own[F32 ? 1 : NQ][RW][VD];
        float fs_own[F32 ? 1 : NS][NQ][RW][VD];
        auto& fin = either<F32>(in, fin_own);
        auto& fs = either<F32>(sw, fs_own);

        auto load = [&](int p, Word (&v)[RW][NV], Word (&hv)[NV]) {
            const bool ok = p >= hlo && p < hhi;  // uniform across the block
            const auto rs = plane_rsrc((const T*)s.src + (size_t)(ok ? p : 0) * plane, ok ? plane_bytes : 0);
#pragma unroll
            for (int i = 0; i < RW; ++i) buf_load<Word, NV>(rs, ld_off[i], v[i]);
            buf_load<Word, NV>(rs, hl_off, hv);
        };
        auto decode = [&](const Word (&w)[RW][NV], float (&f)[RW][VD]) {
            if constexpr (!F32) {
#pragma unroll
                for (int i = 0; i < RW; ++i)
#pragma unroll
                    for (int j = 0; j < VD; ++j) f[i][j] = LW::raw(w[i], j);
            }
        };
        // a plane's boundary rows and edge cells into LDS
        auto stage = [&](LV (*L)[NL], float (*E)[NWD][RW][2], const float (&v)[RW][VD]) {
            lds_put<VD>(&L[2 * ww][col], v[0]);
            lds_put<VD>(&L[2 * ww + 1][col], v[RW - 1]);
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < RW; ++i) E[ww][wd][i][0] = v[i][0];
            }
            if (lane == kWave - 1) {
#pragma unroll
                for (int i = 0; i < RW; ++i) E[ww][wd][i][1] = v[i][VD - 1];
            }
        };
        // one sweep at plane p on this lane's cells: lo/mid/hi hold planes p-1,
        // p, p+1 of its own rows; L/E plane p's boundary rows and edge cells.
        // With HALO, L carries the halo rows (the input plane); without, the
        // tile's edge rows take a stand-in neighbour (those cells are never output).
        // out: storage words; a face keeps its word `midw`, any other cell rounds
        // its result (stochastically: with the word ra ^ B(sweep, plane), where
        // `key` is the sweep's key)
        auto sweep = [&](auto halo, int p, const float (&lo)[RW][VD], const float (&mid)[RW][VD],
                         const float (&hi)[RW][VD], const Word (&midw)[RW][NV], const LV (*L)[NL],
                         const float (*E)[NWD][RW][2], uint32_t key, Word (&out)[RW][NV]) {
            constexpr bool HALO = decltype(halo)::value;
            const bool pint = p >= r && p < H - r;
            const uint32_t rb = SR ? sr_plane_bits(key, s.plane0 + (uint32_t)p) : 0u;  // uniform
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                float wm[VD], wp[VD];
                if (i > 0) {
#pragma unroll
                    for (int j = 0; j < VD; ++j) wm[j] = mid[i - 1][j];
                } else {
                    const int slot = !top ? 2 * ww - 1 : (HALO ? 2 * NWW : 2 * ww);
                    lds_get<VD>(&L[slot][col], wm);
                }
                if (i < RW - 1) {
#pragma unroll
                    for (int j = 0; j < VD; ++j) wp[j] = mid[i + 1][j];
                } else {
                    const int slot = !bottom ? 2 * ww + 2 : (HALO ? 2 * NWW + 1 : 2 * ww + 1);
                    lds_get<VD>(&L[slot][col], wp);
                }
                // cells d-1 and d+VD: the neighbouring lane's, or across the wave edge
                const float left = from_lower_lane(mid[i][VD - 1], E[ww][wd > 0 ? wd - 1 : 0][i][1]);
                const float right = from_upper_lane(mid[i][0], E[ww][wd < NWD - 1 ? wd + 1 : NWD - 1][i][0]);
                float val[VD];
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    const float dm = j > 0 ? mid[i][j - 1] : left;
                    const float dp = j < VD - 1 ? mid[i][j + 1] : right;
                    float acc = lo[i][j] + hi[i][j];
                    acc = acc + wm[j];
                    acc = acc + wp[j];
                    acc = acc + dm;
                    acc = acc + dp;
                    if constexpr (ST == PSKEL_DTYPE_F32) val[j] = coeff * acc;
                    else val[j] = __builtin_fmaf(coeff, acc, c0);
                }
                if constexpr (SR) {
                    uint32_t rw[VD];
#pragma unroll
                    for (int j = 0; j < VD; ++j) rw[j] = ra[i][j] ^ rb;
                    LW::put_sr(out[i], val, rw);
                } else {
                    LW::put(out[i], val);
                }
            }
            // faces keep their word: the whole row on a face plane or row
            // (uniform), else the lane's d-face cells.  A wave with no face
            // cell on an interior plane skips this (a uniform branch).
            if (!(pint && wave_int)) {
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    const uint32_t frow = (!pint || !row_int[i]) ? ~0u : 0u;
#pragma unroll
                    for (int k = 0; k < NV; ++k) out[i][k] = LW::merge(frow | dmask[k], midw[i][k], out[i][k]);
                }
            }
        };

#pragma unroll
        for (int l = 0; l < NS; ++l)
#pragma unroll
            for (int i = 0; i < RW; ++i) {
#pragma unroll
                for (int k = 0; k < NV; ++k) sw[l][0][i][k] = sw[l][1][i][k] = Word(0);
#pragma unroll
                for (int j = 0; j < VD; ++j) fs[l][0][i][j] = fs[l][1][i][j] = 0.f;
            }
        // prologue: each load followed by discarded stores, the order of the
        // steady state, so the wait counts the compiler derives at the loop head
        // hold on every path into it
        {
            const auto none = plane_rsrc(s.dst, 0);
#pragma unroll
            for (int k = 0; k < PF + 2; ++k) {
                load(ha - 1 + k, in[k], hal[k]);
#pragma unroll
                for (int i = 0; i < RW; ++i) buf_store<Word, NV>(none, kNoAccess, sw[0][0][i]);
            }
        }
        // planes ha-1 and ha, decoded (the steady state decodes plane h+1)
        decode(in[0], fin[0]);
        decode(in[1], fin[1]);

        int b = 0;
        auto clear = [&](Word (&w)[RW][NV], float (&f)[RW][VD]) {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
#pragma unroll
                for (int k = 0; k < NV; ++k) w[i][k] = Word(0);
#pragma unroll
                for (int j = 0; j < VD; ++j) f[i][j] = 0.f;
            }
        };
        // step h at ring phase S: input slot (S + k) % NQ holds plane h - 1 + k,
        // slot (S + k) % NQ of sweep l+1's ring (l = 0 .. NS-1) its plane h - l - 2 + k
        auto step = [&](auto phase, int h) __attribute__((always_inline)) {
            constexpr int S = decltype(phase)::value;
            constexpr int I0 = S % NQ, I1 = (S + 1) % NQ, I2 = (S + 2) % NQ, IF = (S + NQ - 1) % NQ;
            load(h + PF + 1, in[IF], hal[IF]);
            // plane h and s_(l+1)(h-l-1) into LDS
            stage(lin[b], lein[b], fin[I1]);
            if (top || bottom) {
                float fh[VD];
#pragma unroll
                for (int j = 0; j < VD; ++j) fh[j] = LW::raw(hal[I1], j);
                lds_put<VD>(&lin[b][hslot][col], fh);
            }
#pragma unroll
            for (int l = 0; l < NS; ++l) stage(ls[l][b], les[l][b], fs[l][I1]);
            __syncthreads();
            decode(in[I2], fin[I2]);  // plane h+1
            // sweep 1 at plane h (h >= ha: the loop starts there)
            if (h < min(h1 + NS, H)) {
                sweep(std::true_type{}, h, fin[I0], fin[I1], fin[I2], in[I1], lin[b], lein[b], keys[0], sw[0][I2]);
                decode(sw[0][I2], fs[0][I2]);
            } else {
                clear(sw[0][I2], fs[0][I2]);
            }
            // sweeps 2 .. DEPTH-1 at planes h-1 .. h-NS+1 (uniform bounds)
            unroll_seq([&](auto lc) __attribute__((always_inline)) {
                constexpr int l = decltype(lc)::value + 1;
                if constexpr (l < NS) {
                    const int q = h - l;
                    if (q >= max(h0 - (NS - l), 0) && q < min(h1 + (NS - l), H)) {
                        sweep(std::false_type{}, q, fs[l - 1][I0], fs[l - 1][I1], fs[l - 1][I2], sw[l - 1][I1],
                              ls[l - 1][b], les[l - 1][b], keys[l], sw[l][I2]);
                        decode(sw[l][I2], fs[l][I2]);
                    } else {
                        clear(sw[l][I2], fs[l][I2]);
                    }
                }
            }, std::make_integer_sequence<int, NS>{});
            // the last sweep at plane h - NS, stored
            const int pl = h - NS;
            const bool out2 = pl >= h0 && pl < h1;  // uniform across the block
            Word o[RW][NV];
            if (out2) {
                sweep(std::false_type{}, pl, fs[NS - 1][I0], fs[NS - 1][I1], fs[NS - 1][I2], sw[NS - 1][I1],
                      ls[NS - 1][b], les[NS - 1][b], keys[NS], o);
                if constexpr (RES) {
#pragma unroll
                    for (int i = 0; i < RW; ++i) {
                        if (!row_out[i]) continue;
#pragma unroll
                        for (int j = 0; j < VD; ++j) {
                            if (d + j < D) {
                                const double dv = (double)LW::raw(o[i], j) - (double)fs[NS - 1][I1][i][j];
                                rr += dv * dv;
                            }
                        }
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < RW; ++i)
#pragma unroll
                    for (int k = 0; k < NV; ++k) o[i][k] = Word(0);
            }
            const auto rs = plane_rsrc((T*)s.dst + (size_t)(out2 ? pl : 0) * plane, out2 ? plane_bytes : 0);
#pragma unroll
            for (int i = 0; i < RW; ++i) buf_store<Word, NV>(rs, st_off[i], o[i]);
            b ^= 1;
        };
        // whole rounds of NQ steps: a step past hend loads and stores through
        // 0-byte resources and computes nothing that is kept
        for (int h = ha; h <= hend; h += NQ) {
            unroll_seq([&](auto ph) __attribute__((always_inline)) { step(ph, h + decltype(ph)::value); },
                       std::make_integer_sequence<int, NQ>{});
        }
    };  // segment

    // The block's work as a range [a, e) of (tile, plane) pairs in tile-major
    // order (index tile * H + plane), marched as one segment per tile it
    // touches.
    // hseg > 0: one segment of hseg planes; tiles fastest, then segments, so
    //   blocks that run at once on one XCD hold neighbouring tiles at the same
    //   planes and the rows they share are read from HBM once.
    // hseg < 0: the grid's tile-planes split evenly, -hseg per block, over
    //   one block per resident slot: no second round of blocks.  A block's
    //   range may end in one tile and go on in the next.  Neighbouring tiles'
    //   blocks are consecutive in the XCD order and march at plane offsets
    //   below the range length.
    const int tiles = ntd * ntw;
    long long a, e;
    if (hseg > 0) {
        const int tile = (int)(lid % tiles), sg = (int)(lid / tiles);
        a = (long long)tile * H + (long long)sg * hseg;
        e = (long long)tile * H + min(sg * hseg + hseg, H);
    } else {
        const long long per = -(long long)hseg;
        a = (long long)lid * per;
        e = min(a + per, (long long)tiles * H);
    }
#pragma unroll 1
    while (a < e) {
        const int t = (int)(a / H), p0 = (int)(a % H);
        const int p1 = (int)min((long long)H, (long long)p0 + (e - a));
        segment(t % ntd, t / ntd, p0, p1);
        a += p1 - p0;
    }

    if constexpr (RES) {
        const double t = block_sum<NW>(rr, red);
        if (threadIdx.x == 0) s.partials[lid] = t;
    }
}

// Blocks and h planes per block.  cfg.tile_z fixes the planes (segments of
// that many planes, *hseg > 0).  By default two splits are scored by the
// useful fraction of the device's block slots (CUs x resident blocks per CU),
// net of the 2*DEPTH planes a segment reads twice:
// * n equal segments per tile (*hseg > 0): (tiles * n / (rounds * slots)) *
//   (hseg / (hseg + 2*DEPTH));
// * the tiles' planes split evenly over at most one block per slot, at least
//   4*DEPTH planes each (*hseg = -planes per block): (blocks / slots) *
//   (per / (per + 2*DEPTH * tiles a range touches)).
// The even split is taken when it scores 2% more.
template <int ST, int VD, int RW, int NWD, int NWW, int PF, int DEPTH>
size_t march2_grid(const Sweep3D& s, const March2Shape& m, int* hseg)
{
    const int64_t tiles = (int64_t)m.ntd * m.ntw;
    int q = m.hseg > 0 ? std::min(m.hseg, s.H) : 0;
    if (q <= 0) {
        // resident blocks per CU of this kernel: the least over its variants
        // (queried once; grids may launch from several host threads)
        static std::atomic<int> occ_cache{0};
        int occ = occ_cache.load(std::memory_order_relaxed);
        if (occ == 0) {
            constexpr int NT = NWD * NWW * kWave;
            int o[4] = {1, 1, 1, 1};
            bool ok = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &o[0], k3d_march2<ST, VD, RW, NWD, NWW, PF, DEPTH, true, false>, NT, 0) == hipSuccess &&
                      hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &o[1], k3d_march2<ST, VD, RW, NWD, NWW, PF, DEPTH, false, false>, NT, 0) == hipSuccess;
            if constexpr (ST != PSKEL_DTYPE_F32)
                ok = ok &&
                     hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &o[2], k3d_march2<ST, VD, RW, NWD, NWW, PF, DEPTH, true, true>, NT, 0) == hipSuccess &&
                     hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &o[3], k3d_march2<ST, VD, RW, NWD, NWW, PF, DEPTH, false, true>, NT, 0) == hipSuccess;
            occ = ok ? std::max(1, std::min(std::min(o[0], o[1]), std::min(o[2], o[3]))) : 1;
            occ_cache.store(occ, std::memory_order_relaxed);
        }
        static std::atomic<int> cus_of[64] = {};  // CUs per device, queried once
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
        int cus = cus_of[dev].load(std::memory_order_relaxed);
        if (cus <= 0) {
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
                cus = 1;
            cus_of[dev].store(cus, std::memory_order_relaxed);
        }
        const int64_t slots = (int64_t)cus * occ;
        double best = -1.0;
        for (int n = 1; n <= std::min(s.H, 256); ++n) {
            const int hs = (s.H + n - 1) / n;
            const int64_t nb = tiles * ((s.H + hs - 1) / hs);
            const int64_t rounds = (nb + slots - 1) / slots;
            const double u = (double)nb / (double)(rounds * slots) * hs / (hs + 2.0 * DEPTH);
            if (u > best + 1e-9) {
                best = u;
                q = hs;
            }
        }
        const int64_t total = tiles * (int64_t)s.H;
        const int64_t minlen = 4 * DEPTH;
        const int64_t nb2 = std::max<int64_t>(1, std::min<int64_t>(slots, (total + minlen - 1) / minlen));
        const int64_t per = (total + nb2 - 1) / nb2;
        const double touched = 1.0 + (double)(per - 1) / (double)s.H;  // tiles a range spans, on average
        const double u2 = (double)nb2 / (double)slots * (double)per / ((double)per + 2.0 * DEPTH * touched);
        if (u2 > best * 1.02 && per <= 0x7fffffff) {
            *hseg = -(int)per;
            return (size_t)((total + per - 1) / per);
        }
    }
    *hseg = q;
    return (size_t)(tiles * ((s.H + q - 1) / q));
}

template <int ST, int VD, int RW, int NWD, int NWW, int PF, int DEPTH>
hipError_t march2_launch(const Sweep3D& s, const Points& p, const March2Shape& m, hipStream_t st)
{
    int hseg = 0;
    const dim3 grid((unsigned)march2_grid<ST, VD, RW, NWD, NWW, PF, DEPTH>(s, m, &hseg)),
        block(NWD * NWW * kWave);
    const bool sr = ST != PSKEL_DTYPE_F32 && s.rounding == PSKEL_ROUND_STOCHASTIC;
#define PSKEL_MARCH2_GO(RESV, SRV)                                                                     \
    hipLaunchKernelGGL((k3d_march2<ST, VD, RW, NWD, NWW, PF, DEPTH, RESV, SRV>), grid, block, 0, st, s, \
                       p.coeff, hseg, m.ntd, m.ntw)
    if constexpr (ST == PSKEL_DTYPE_F32) {
        if (s.partials) PSKEL_MARCH2_GO(true, false);
        else PSKEL_MARCH2_GO(false, false);
    } else {
        if (s.partials && sr) PSKEL_MARCH2_GO(true, true);
        else if (s.partials) PSKEL_MARCH2_GO(true, false);
        else if (sr) PSKEL_MARCH2_GO(false, true);
        else PSKEL_MARCH2_GO(false, false);
    }
#undef PSKEL_MARCH2_GO
    return hipGetLastError();
}

// instantiated shapes (vd, rw, nwd, nww, pf, depth), as march2_supported lists them
#define PSKEL_MARCH2_SHAPES_F32(X) \
    X(2, 4, 4, 4, 1, 2)            \
    X(2, 4, 4, 2, 1, 2)            \
    X(4, 2, 2, 8, 1, 2)            \
    X(4, 2, 2, 4, 1, 2)            \
    X(2, 4, 4, 4, 2, 2)            \
    X(2, 4, 2, 4, 1, 2)            \
    X(4, 4, 2, 4, 1, 2)            \
    X(4, 3, 2, 6, 1, 2)            \
    X(4, 3, 1, 12, 1, 2)           \
    X(2, 3, 1, 12, 1, 2)           \
    X(4, 4, 2, 4, 1, 3)            \
    X(4, 2, 2, 6, 1, 2)            \
    X(4, 2, 2, 6, 1, 3)
#define PSKEL_MARCH2_SHAPES_16(X) \
    X(4, 3, 2, 6, 1, 2)           \
    X(4, 2, 2, 8, 1, 2)           \
    X(4, 3, 1, 12, 1, 2)          \
    X(2, 3, 1, 12, 1, 2)          \
    X(4, 2, 2, 6, 1, 2)

template <int ST>
size_t march2_blocks(const Sweep3D& s, const March2Shape& m)
{
    int hseg = 0;
#define PSKEL_MARCH2(VD, RW, NWD, NWW, PF, DEPTH)                                                         \
    if (m.vd == VD && m.rw == RW && m.nwd == NWD && m.nww == NWW && m.pf == PF && m.depth == DEPTH) \
        return march2_grid<ST, VD, RW, NWD, NWW, PF, DEPTH>(s, m, &hseg);
    if constexpr (ST == PSKEL_DTYPE_F32) {
        PSKEL_MARCH2_SHAPES_F32(PSKEL_MARCH2)
    } else {
        PSKEL_MARCH2_SHAPES_16(PSKEL_MARCH2)
    }
#undef PSKEL_MARCH2
    return 0;
}

template <int ST>
hipError_t march2_dispatch(const Sweep3D& s, const Points& p, const March2Shape& m, hipStream_t st)
{
    if (!march2_supported(m, ST)) return hipErrorInvalidValue;
#define PSKEL_MARCH2(VD, RW, NWD, NWW, PF, DEPTH)                                                         \
    if (m.vd == VD && m.rw == RW && m.nwd == NWD && m.nww == NWW && m.pf == PF && m.depth == DEPTH) \
        return march2_launch<ST, VD, RW, NWD, NWW, PF, DEPTH>(s, p, m, st);
    if constexpr (ST == PSKEL_DTYPE_F32) {
        PSKEL_MARCH2_SHAPES_F32(PSKEL_MARCH2)
    } else {
        PSKEL_MARCH2_SHAPES_16(PSKEL_MARCH2)
    }
#undef PSKEL_MARCH2
    return hipErrorInvalidValue;
}

}  // namespace
#endif  // PSKEL_MARCH_INSTANTIATE

}  // namespace pskel
