// K fused sweeps per launch (2 <= K <= 4) of a radius-1 2D stencil, the
// 5-point scaled sum (PAT_J5) or the weighted 3x3 box (PAT_BOX9), f32, with
// the sweeps pipelined in registers instead of an LDS tile.
//
// A wave owns a tile of 64 lanes x VD cells along w and marches down a
// segment of rows along h.  At step i it receives input row yi and computes
// sweep l on row yi - l, from sweep l-1's rows yi-l-1, yi-l, yi-l+1, which
// sit in a ring of three rows per sweep; the last sweep's row is stored.
// Neighbours along w come from the lane's own cells and, for its first and
// last cell, from the lanes beside it (DPP row shifts, no LDS).  Every cell
// evaluates the same expression in the same order as the other 2D kernels,
// with no contraction, so a launch equals K single sweeps bit for bit.
//
// Tiles.  Along w a tile's first and last lane vector are a halo: sweep l is
// exact on the tile minus l cells at each side, so with K <= VD the 62 inner
// lanes (62 * VD cells) are exact after K sweeps, and neighbouring tiles
// overlap by two lane vectors.  Along h a segment of rows [a, e) reads input
// rows [a - K, e + K); its sweeps run only on the rows the output needs.
//
// The ring.  2D sweeps never write the grid's ring (rows 0 and H-1, columns 0
// and W-1): sweep l reads, at a ring cell, the ring of the buffer sweep l
// lives in (even l: src, odd l: ring_other).  A computed ring row is replaced
// by that buffer's row, loaded once per segment that reaches it.  A ring
// column cell takes, at sweep 1, ring_other's value (loaded with each input
// row by the lane that holds it) and at sweep l >= 2 sweep l-2's value, which
// its ring still holds.  Cells past the grid (x < 0, x >= W, rows outside
// [0, H)) hold zeros or stale values that only ever reach ring cells, which
// are replaced.
//
// Memory.  Input rows come through buffer loads PD rows ahead; rows
// outside the grid get an offset past the buffer (the range check returns
// zeros).  Every step issues the same loads and one store (out of range
// while the last sweep has no output row yet), so the wait counts the
// compiler derives hold on every path and the prefetch is never drained.
// Waves are independent: no barriers, no LDS.
#pragma once

#include <type_traits>

#include "k_common.hpp"

namespace pskel {

constexpr int kMarch2DWaves = 4;  // waves per block
constexpr int kMarch2DVd = 4;     // cells per lane along w

// cells one tile writes along w
__host__ __device__ constexpr int march2d_out_cells() { return (kWave - 2) * kMarch2DVd; }

// Launch geometry: tiles along w, rows per segment, segments, wave units.
struct March2DShape {
    int ntw = 0, seg = 0, nseg = 0, units = 0;
};

// seg <= 0 picks segments for about 3072 waves (3 per SIMD), at least 4 rows
// each: the march runs 3 waves per SIMD by its VGPRs, and on Jacobi-2D 2048^2
// segments of 6 rows (3069 waves) beat 8 rows by 10% despite the longer warm-up
inline March2DShape march2d_shape(int W, int H, int seg)
{
    March2DShape m;
    const int rows = H - 2;  // interior rows
    if (W < 3 || rows < 1) return m;
    m.ntw = (W + march2d_out_cells() - 1) / march2d_out_cells();
    if (seg <= 0) {
        const int nseg = std::max(1, 3072 / m.ntw);
        seg = std::max(4, (rows + nseg - 1) / nseg);
    }
    m.seg = seg;
    m.nseg = (rows + seg - 1) / seg;
    m.units = m.ntw * m.nseg;
    return m;
}

// grids whose buffer (plus the rows a segment reads past it) a 32-bit buffer
// offset covers, at least 2 lane vectors wide
inline bool march2d_fits(int W, int H, size_t pitch)
{
    return W > 2 * kMarch2DVd && (uint64_t)(H + 8) * pitch * 4 < 0x7fffffffull;
}

#ifdef PSKEL_MARCH2D_INSTANTIATE
namespace {

template <int N>
using IC = std::integral_constant<int, N>;

// f(IC<0>), f(IC<1>), ... while f returns true
template <class F, int... Is>
__device__ __forceinline__ bool m2d_unroll(F&& f, std::integer_sequence<int, Is...>)
{
    return (f(IC<Is>()) && ...);
}

template <int VD>
__device__ __forceinline__ void m2d_load(const __amdgpu_buffer_rsrc_t& r, uint32_t off, float (&v)[VD])
{
    static_assert(VD == 4, "four cells per lane");
    const u32x4_t x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    v[0] = __uint_as_float(x[0]);
    v[1] = __uint_as_float(x[1]);
    v[2] = __uint_as_float(x[2]);
    v[3] = __uint_as_float(x[3]);
}

// one output row of a sweep: u, c, d are the rows above, at and below, with
// their left (cell x0-1) and right (cell x0+VD) neighbours from the lanes beside;
// p holds the scaled sum's coeff and the 3x3 weights wt
template <int PAT, int VD, class PT>
__device__ __forceinline__ void m2d_row(const PT& p, const float (&u)[VD], float uL, float uR,
                                        const float (&c)[VD], float cL, float cR, const float (&d)[VD], float dL,
                                        float dR, float (&o)[VD])
{
#pragma unroll
    for (int j = 0; j < VD; ++j) {
        const float cm = j > 0 ? c[j - 1] : cL, cp = j < VD - 1 ? c[j + 1] : cR;
        float acc;
        if constexpr (PAT == PAT_J5) {
            acc = u[j] + d[j];
            acc = acc + cm;
            acc = acc + cp;
            acc = p.coeff * acc;
        } else {
            const float um = j > 0 ? u[j - 1] : uL, up = j < VD - 1 ? u[j + 1] : uR;
            const float dm = j > 0 ? d[j - 1] : dL, dp = j < VD - 1 ? d[j + 1] : dR;
            acc = p.wt[0] * um;
            acc = acc + p.wt[1] * u[j];
            acc = acc + p.wt[2] * up;
            acc = acc + p.wt[3] * cm;
            acc = acc + p.wt[4] * c[j];
            acc = acc + p.wt[5] * cp;
            acc = acc + p.wt[6] * dm;
            acc = acc + p.wt[7] * d[j];
            acc = acc + p.wt[8] * dp;
        }
        o[j] = acc;
    }
}

// One wave's unit g (tile g % ntw, segment g / ntw, g < ntw * segments):
// K sweeps of the segment's rows.  Waves are independent (no barrier), so a
// caller may run any set of units in any block.
template <int PAT, int K, int PD, class PT>
__device__ __forceinline__ void march2d_wave(const Sweep2D& s, const PT& p, int ntw, int seg, int g, int lane)
{
    constexpr int VD = kMarch2DVd;
    constexpr int NL = K - 1;     // sweeps kept in rings (all but the last)
    constexpr int NI = PD + 3;    // input slots: rows yi-2 .. yi + PD
    static_assert(K >= 2 && K <= VD, "the halo lane vector covers K cells");
    static_assert(PD % 3 == 0, "the step loop unrolls over the input slots, a multiple of the three-row rings");
    const int t = g % ntw, sg = g / ntw;
    const int H = s.H, W = s.W;
    const int a = 1 + sg * seg, e = min(a + seg, H - 1);
    const int N = e - a + 2 * K;  // steps: input rows a-K .. e+K-1
    const int x0 = t * march2d_out_cells() - VD;
    const int xs = x0 + lane * VD;  // this lane's first cell
    const uint32_t pb = (uint32_t)s.pitch * 4u;
    const int bytes = (int)((uint32_t)H * pb);
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(s.src, bytes);
    const __amdgpu_buffer_rsrc_t ro = plane_rsrc(s.ring_other, bytes);
    // byte offset of this lane's vector in row y (rows past H fall out of range)
    auto roff = [&](int y) -> uint32_t { return (y >= 0 && xs >= 0) ? (uint32_t)y * pb + (uint32_t)xs * 4u : kNoAccess; };

    // ring columns: a tile that holds column 0 or W-1 (uniform), and the cell
    // of this lane that is one (-1: none; a lane holds at most one, W > VD)
    const bool edge = x0 <= 0 || (x0 <= W - 1 && W - 1 < x0 + kWave * VD);
    int ie = -1;
    if (xs <= 0 && 0 < xs + VD) ie = -xs;
    else if (xs <= W - 1 && W - 1 < xs + VD) ie = W - 1 - xs;
    bool pm[VD];
#pragma unroll
    for (int j = 0; j < VD; ++j) pm[j] = edge && ie == j;
    auto eoff = [&](int y) -> uint32_t {
        return (edge && ie >= 0 && y >= 0) ? (uint32_t)y * pb + (uint32_t)(xs + ie) * 4u : kNoAccess;
    };

    // ring rows 0 and H-1 of both buffers, when a sweep of this segment reaches them
    float tsrc[VD], toth[VD], bsrc[VD], both[VD];
    const bool top = a - K + 1 <= 0, bot = e + K - 2 >= H - 1;
    m2d_load<VD>(rs, top ? roff(0) : kNoAccess, tsrc);
    m2d_load<VD>(ro, top ? roff(0) : kNoAccess, toth);
    m2d_load<VD>(rs, bot ? roff(H - 1) : kNoAccess, bsrc);
    m2d_load<VD>(ro, bot ? roff(H - 1) : kNoAccess, both);

    float in[NI][VD];         // input rows, slot = step % NI
    float oe[NI];             // ring_other's value at this lane's ring-column cell, same slots
    float inL[3], inR[3];     // neighbours of input rows, slot = step % 3
    float lv[NL][3][VD];      // sweeps 1..K-1, slot = step % 3
    float lvL[NL][3], lvR[NL][3];
    const int y0 = a - K;
    // stores: lanes 1..62 whose vector starts inside the grid write all VD
    // cells (a ring-column cell gets the ring value it holds, a cell past W
    // the row's padding, which stays inside the row: the pitch is a multiple
    // of VD cells >= W); every step issues one store, out of range before the
    // last sweep's first output row
    const __amdgpu_buffer_rsrc_t rd = plane_rsrc(s.dst, bytes);
    const bool st_lane = lane >= 1 && lane <= kWave - 2 && xs < W;
    // prologue: each load followed by a discarded store, the order of the
    // steady state, so the wait counts at the loop head hold on every path
    const u32x4_t none = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int q = 0; q < PD; ++q) {
        m2d_load<VD>(rs, roff(y0 + q), in[q]);
        oe[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ro, (int)eoff(y0 + q), 0, 0));
        __builtin_amdgcn_raw_buffer_store_b128(none, rd, (int)kNoAccess, 0, 0);
    }

    // a computed row y of sweep l (1 <= l <= K), patched where it holds ring
    // cells; `prev2` is sweep l-2's row y (or the input row, l = 2)
    auto patch = [&](int l, int y, float (&v)[VD], const float (&prev2)[VD], float oth_edge) {
        if (y == 0 || y == H - 1) {  // uniform
            const bool odd = l & 1;
#pragma unroll
            for (int j = 0; j < VD; ++j)
                v[j] = y == 0 ? (odd ? toth[j] : tsrc[j]) : (odd ? both[j] : bsrc[j]);
        } else if (edge) {
#pragma unroll
            for (int j = 0; j < VD; ++j) v[j] = pm[j] ? (l == 1 ? oth_edge : prev2[j]) : v[j];
        }
    };

#pragma unroll 1
    for (int base = 0; base < N; base += NI) {
        // NI steps with static ring slots
        auto run = [&](auto P) -> bool {
            constexpr int ph = decltype(P)::value;
            constexpr int s0 = ph % 3, s1 = (ph + 2) % 3, s2 = (ph + 1) % 3;
            constexpr int i0 = ph, i1 = (ph + NI - 1) % NI, i2 = (ph + NI - 2) % NI;
            const int i = base + ph;
            if (i >= N) return false;
            const int yi = y0 + i;
            // prefetch row yi + PD into the slot of row yi - 3 (no longer read)
            m2d_load<VD>(rs, roff(yi + PD), in[(ph + PD) % NI]);
            oe[(ph + PD) % NI] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ro, (int)eoff(yi + PD), 0, 0));
            inL[s0] = from_lower_lane(in[i0][VD - 1], 0.f);
            inR[s0] = from_upper_lane(in[i0][0], 0.f);
            if (i >= 2) {  // sweep 1 on row yi - 1
                float o[VD];
                m2d_row<PAT, VD>(p, in[i2], inL[s2], inR[s2], in[i1], inL[s1], inR[s1], in[i0], inL[s0], inR[s0], o);
                patch(1, yi - 1, o, in[i1], oe[i1]);
#pragma unroll
                for (int j = 0; j < VD; ++j) lv[0][s0][j] = o[j];
                lvL[0][s0] = from_lower_lane(o[VD - 1], 0.f);
                lvR[0][s0] = from_upper_lane(o[0], 0.f);
            }
#pragma unroll
            for (int l = 2; l <= NL; ++l) {  // sweep l on row yi - l
                if (i >= 2 * l) {
                    float o[VD];
                    m2d_row<PAT, VD>(p, lv[l - 2][s2], lvL[l - 2][s2], lvR[l - 2][s2], lv[l - 2][s1], lvL[l - 2][s1],
                                     lvR[l - 2][s1], lv[l - 2][s0], lvL[l - 2][s0], lvR[l - 2][s0], o);
                    // sweep l-2's row yi - l: the input row (l = 2), else ring l-3 two steps ago
                    if (l == 2) patch(l, yi - l, o, in[i2], 0.f);
                    else patch(l, yi - l, o, lv[l >= 3 ? l - 3 : 0][s2], 0.f);
#pragma unroll
                    for (int j = 0; j < VD; ++j) lv[l - 1][s0][j] = o[j];
                    lvL[l - 1][s0] = from_lower_lane(o[VD - 1], 0.f);
                    lvR[l - 1][s0] = from_upper_lane(o[0], 0.f);
                }
            }
            // sweep K on row yi - K (an interior row from step 2K on), stored
            u32x4_t ov = none;
            if (i >= 2 * K) {
                float o[VD];
                constexpr int m = NL - 1;
                m2d_row<PAT, VD>(p, lv[m][s2], lvL[m][s2], lvR[m][s2], lv[m][s1], lvL[m][s1], lvR[m][s1], lv[m][s0],
                                 lvL[m][s0], lvR[m][s0], o);
                // ring-column cells: sweep K-2's value (the input row, K = 2)
                if (K == 2) patch(K, yi - K, o, in[i2], 0.f);
                else patch(K, yi - K, o, lv[K >= 3 ? K - 3 : 0][s2], 0.f);
#pragma unroll
                for (int j = 0; j < VD; ++j) ov[j] = __float_as_uint(o[j]);
            }
            const uint32_t so = (st_lane && i >= 2 * K) ? (uint32_t)(yi - K) * pb + (uint32_t)xs * 4u : kNoAccess;
            __builtin_amdgcn_raw_buffer_store_b128(ov, rd, (int)so, 0, 0);
            return true;
        };
        if (!m2d_unroll(run, std::make_integer_sequence<int, NI>())) break;
    }
}

template <int PAT, int K, int PD>
__global__ __launch_bounds__(kMarch2DWaves * kWave) void k2d_march(Sweep2D s, Points p, int ntw, int seg, int units)
{
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int g = (int)xcd_remap(blockIdx.x, gridDim.x) * kMarch2DWaves + wave;
    if (g >= units) return;  // the whole wave
    march2d_wave<PAT, K, PD>(s, p, ntw, seg, g, lane);
}

}  // namespace
#endif  // PSKEL_MARCH2D_INSTANTIATE

}  // namespace pskel
