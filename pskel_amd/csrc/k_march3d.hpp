// 2.5D marching kernel for the 7-point scaled sum with COPY faces, shared by
// the per-storage translation units (k_march3d_{f32,bf16,f16}.hip).
//
// A block owns a (w, d) tile and walks a segment of h planes, keeping the
// decoded planes h-1/h/h+1 in registers and staging plane h (with its one-cell
// w/d halo) in a double-buffered LDS tile: one barrier per plane.  Each lane
// holds VD consecutive d cells of RW consecutive w rows.  The raw plane h+2 is
// in flight while plane h is computed.  Interior cells take a fast path with
// no face tests.  Every cell evaluates
//     coeff * (((((h- + h+) + w-) + w+) + d-) + d+)
// in f32 with no contraction, the order of the Jacobi mask
// (-1,0,0) (1,0,0) (0,-1,0) (0,1,0) (0,0,-1) (0,0,1).  16-bit storage sums
// the stored values x = u - off in the same order and stores
// fma(coeff, sum, c0), with c0 = (6 * coeff - 1) * off (scaled_sum_offset_term).
//
// Two forms of the register queue along h, picked by element size (both
// measured on MI355X, 512^3):
//   * f32: the queue shifts by one plane per step (register moves), and the
//     loads of rows, halos and planes outside the grid are skipped;
//   * 16-bit: the h loop is unrolled by 3 so the queue slots are renamed
//     instead of moved, and every lane issues the same loads every plane with
//     addresses clamped into the grid, so the count of loads in flight is
//     static and a wait for plane h+1 never drains plane h+2.  A 16-bit lane
//     moves few bytes per plane and is bound by instruction issue, which the
//     renaming relieves; the f32 kernel is bound by HBM, where the extra
//     clamped loads cost more than the moves.
#pragma once

#include <utility>

#include "k_common.hpp"

namespace pskel {

constexpr int kMarchWaves = 4;

struct MarchShape {
    int vd, rw, hseg, ntd, ntw, nseg;
};

inline MarchShape march_shape(const Sweep3D& s, int storage, const pskel_launch_cfg_t& cfg)
{
    MarchShape m;
    const int esz = storage == PSKEL_DTYPE_F32 ? 4 : 2;
    m.vd = cfg.tile_x > 0 ? cfg.tile_x : (esz == 4 ? 1 : 2);
    m.rw = cfg.tile_y > 0 ? cfg.tile_y : 4;
    m.hseg = cfg.tile_z > 0 ? cfg.tile_z : 64;
    if (m.hseg > s.H) m.hseg = s.H;
    m.ntd = (s.D + kWave * m.vd - 1) / (kWave * m.vd);
    m.ntw = (s.W + kMarchWaves * m.rw - 1) / (kMarchWaves * m.rw);
    m.nseg = (s.H + m.hseg - 1) / m.hseg;
    return m;
}

// LDS bytes of one marching block: two planes of (TW + 2) x (TD + 2 VD) floats.
constexpr int march_lds_bytes(int vd, int rw)
{
    return 2 * (kMarchWaves * rw + 2) * (kWave * vd + 2 * vd) * 4;
}

hipError_t march3d_f32(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st);
hipError_t march3d_bf16(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st);
hipError_t march3d_f16(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st);

#ifdef PSKEL_MARCH_INSTANTIATE
namespace {

// f(integral_constant<int, S>) for S = 0, 1, ..., in order.
template <typename F, int... S>
__device__ __forceinline__ void unroll_seq(F&& f, std::integer_sequence<int, S...>)
{
    (f(std::integral_constant<int, S>{}), ...);
}

template <int ST, int VD, int RW, bool RES, bool SR>
__global__ __launch_bounds__(256) void k3d_march(Sweep3D s, float coeff, int hseg, int ntd,
                                                 int ntw, int nseg)
{
    using Tr = Stor<ST>;
    using T = typename Tr::T;
    constexpr bool REN = sizeof(T) == 2;  // renamed (unrolled) queue for 16-bit storage
    constexpr int NW = kMarchWaves;
    constexpr int TD = kWave * VD;
    constexpr int TW = NW * RW;
    constexpr int LROW = TD + 2 * VD;  // [VD-1] left halo, [VD, VD+TD) tile, [VD+TD] right halo
    __shared__ float lds[2][TW + 2][LROW];
    __shared__ double red[NW];

    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const uint32_t nb = (uint32_t)(ntd * ntw * nseg);
    const uint32_t lid = xcd_remap(blockIdx.x, nb);
    const int td = (int)(lid % ntd);
    const int tw = (int)((lid / ntd) % ntw);
    const int sg = (int)(lid / ((uint32_t)ntd * ntw));

    const int W = s.W, H = s.H, D = s.D, r = s.radius;
    const int pitch = (int)s.pitch;
    const size_t plane = (size_t)W * s.pitch;
    // 16-bit storage: the stored values (u - off) are summed as they are and
    // the offset term folds into one fused multiply-add (scaled_sum_offset_term)
    const float c0 = s.c0;
    auto encode = [&](float acc, uint32_t rw) __attribute__((always_inline)) {
        const float v = ST == PSKEL_DTYPE_F32 ? coeff * acc : __builtin_fmaf(coeff, acc, c0);
        return SR ? Tr::enc_sr_raw(v, rw) : Tr::enc_raw(v);
    };

    const int d0 = td * TD, w0 = tw * TW;
    const int d = d0 + lane * VD;
    const int wr = w0 + wave * RW;
    const int h0 = sg * hseg;
    const int h1 = min(h0 + hseg, H);
    const bool dlive = d < D;
    // every cell of this lane's vector is an interior d (no face, inside the grid)
    const bool dint = d >= r && d + VD - 1 < D - r;

    // halo roles
    const bool hal_wlo = wave == 0 && w0 > 0;              // loads row w0-1
    const bool hal_whi = wave == NW - 1 && w0 + TW < W;     // loads row w0+TW
    const bool hal_dlo = lane == 0 && d0 > 0;               // loads d0-1 for its rows
    const bool hal_dhi = lane == kWave - 1 && d0 + TD < D;  // loads d0+TD for its rows
    const int hw_off = (hal_wlo ? -(wave * RW + 1) : RW * (NW - wave)) * pitch;  // w-halo row - own row
    const int hd_off = hal_dlo ? -1 - lane * VD : TD - lane * VD;                 // d-halo cell - own cell

    bool row_live[RW], row_int[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        const int w = wr + i;
        row_live[i] = dlive && w < W;
        row_int[i] = w >= r && w < W - r;
    }

    // own cell (h0, wr, d)
    const T* __restrict__ pin = (const T*)s.src + (size_t)h0 * plane + (size_t)wr * pitch + d;
    T* __restrict__ pout = (T*)s.dst + (size_t)h0 * plane + (size_t)wr * pitch + d;

    // Queue: qr/qw/qd slot k holds a raw plane (own rows, w-halo row vector of
    // this wave's role, d-halo cell per row of this lane's role); pl holds the
    // decoded planes h-1, h, h+1.
    T qr[3][RW][VD];
    T qw[3][VD];
    T qd[3][RW];
    float pl[3][RW][VD];

    // Clamped addresses for the 16-bit form: what a clamped load returns only
    // ever reaches cells that are not computed -- rows past W and lanes past D
    // are never stored, halo values of lanes/waves without that role are never
    // read, and planes past the segment (or the grid) only feed faces.  Rows
    // are padded to whole tiles, so every lane's d stays inside the pitch.
    int row_off[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i) row_off[i] = (min(wr + i, W - 1) - wr) * pitch;
    const int hw_addr = (hal_wlo || hal_whi) ? hw_off : row_off[0];
    const int hlast = min(h1, H - 1);  // last plane a block reads

    auto load_plane = [&](const T* p, T (&raw)[RW][VD], T (&hw)[VD], T (&hd)[RW]) {
        if constexpr (REN) {
#pragma unroll
            for (int i = 0; i < RW; ++i) load_vec<T, VD>(p + row_off[i], raw[i]);
#pragma unroll
            for (int i = 0; i < RW; ++i) hd[i] = p[row_off[i] + ((hal_dlo || hal_dhi) ? hd_off : 0)];
            load_vec<T, VD>(p + hw_addr, hw);
        } else {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                if (row_live[i]) {
                    load_vec<T, VD>(p + i * pitch, raw[i]);
                } else {
#pragma unroll
                    for (int j = 0; j < VD; ++j) raw[i][j] = T(0);
                }
                hd[i] = T(0);
                if ((hal_dlo || hal_dhi) && wr + i < W) hd[i] = p[i * pitch + hd_off];
            }
            if (dlive && (hal_wlo || hal_whi)) {
                load_vec<T, VD>(p + hw_off, hw);
            } else {
#pragma unroll
                for (int j = 0; j < VD; ++j) hw[j] = T(0);
            }
        }
    };
    auto zero_plane = [&](T (&raw)[RW][VD], T (&hw)[VD], T (&hd)[RW]) {
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            hd[i] = T(0);
#pragma unroll
            for (int j = 0; j < VD; ++j) raw[i][j] = T(0);
        }
#pragma unroll
        for (int j = 0; j < VD; ++j) hw[j] = T(0);
    };
    // plane h + k into slot `k` (k = 1, 2), as far as the segment needs it
    auto load_ahead = [&](int h, int k, T (&raw)[RW][VD], T (&hw)[VD], T (&hd)[RW]) {
        if constexpr (REN) {
            load_plane(pin + (size_t)(min(h + k, hlast) - h0) * plane, raw, hw, hd);
        } else {
            if (h + k < H && h + k <= h1) load_plane(pin + (size_t)(h + k - h0) * plane, raw, hw, hd);
            else zero_plane(raw, hw, hd);
        }
    };
    auto decode = [&](float (&v)[RW][VD], const T (&raw)[RW][VD]) {
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
            for (int j = 0; j < VD; ++j) v[i][j] = Tr::dec_raw(raw[i][j]);
    };

    // pl[0] = plane h0-1 (only read at h0 > 0), pl[1] = plane h0, slots 0/1 = planes h0, h0+1
    if (REN || h0 > 0) {
        T tr[RW][VD], thw[VD], thd[RW];
        load_plane(pin - (h0 > 0 ? plane : 0), tr, thw, thd);
        decode(pl[0], tr);
    } else {
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
            for (int j = 0; j < VD; ++j) pl[0][i][j] = 0.f;
    }
    load_plane(pin, qr[0], qw[0], qd[0]);
    load_ahead(h0, 1, qr[1], qw[1], qd[1]);
    decode(pl[1], qr[0]);

    // stochastic rounding: the in-plane words of this lane's cells, hashed once
    const uint32_t key = SR ? sweep_key(s.seed, s.sweep) : 0u;
    uint32_t ra[RW][VD];
    if constexpr (SR) {
        const uint32_t ckey = sr_cell_key(s.seed);
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
            for (int j = 0; j < VD; ++j) ra[i][j] = sr_cell_bits(ckey, (uint32_t)(wr + i) * D + d + j);
    }
    double rr = 0.0;
    int b = 0;

    // one plane at queue phase S: slot (S + k) % 3 holds plane h + k, and
    // pl[(S + k) % 3] the decoded plane h - 1 + k
    auto step = [&](auto phase, int h) __attribute__((always_inline)) {
        constexpr int S = decltype(phase)::value;
        constexpr int CUR = S % 3, NXT = (S + 1) % 3, FUT = (S + 2) % 3;
        float (&prv)[RW][VD] = pl[S % 3];
        float (&cur)[RW][VD] = pl[(S + 1) % 3];
        float (&nxt)[RW][VD] = pl[(S + 2) % 3];
        T* pout_h = pout + (size_t)(h - h0) * plane;
        const uint32_t rbh = SR ? sr_plane_bits(key, s.plane0 + h) : 0u;  // uniform: one hash per plane
        load_ahead(h, 2, qr[FUT], qw[FUT], qd[FUT]);
        decode(nxt, qr[NXT]);
        // stage plane h
        float (*L)[LROW] = lds[b];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
#pragma unroll
            for (int j = 0; j < VD; ++j) L[1 + wave * RW + i][VD + lane * VD + j] = cur[i][j];
            if (hal_dlo) L[1 + wave * RW + i][VD - 1] = Tr::dec_raw(qd[CUR][i]);
            if (hal_dhi) L[1 + wave * RW + i][VD + TD] = Tr::dec_raw(qd[CUR][i]);
        }
        if (hal_wlo || hal_whi) {
            const int row = hal_wlo ? 0 : TW + 1;
#pragma unroll
            for (int j = 0; j < VD; ++j) L[row][VD + lane * VD + j] = Tr::dec_raw(qw[CUR][j]);
        }
        __syncthreads();

        const bool hint = h >= r && h < H - r;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            if (!row_live[i]) continue;
            const int lr = 1 + wave * RW + i;
            T o[VD];
            if (hint && row_int[i] && dint) {
                // interior fast path: no per-cell face tests
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    const float wm = (i > 0) ? cur[i - 1][j] : L[lr - 1][VD + lane * VD + j];
                    const float wp = (i < RW - 1) ? cur[i + 1][j] : L[lr + 1][VD + lane * VD + j];
                    const float dm = (j > 0) ? cur[i][j - 1] : L[lr][VD + lane * VD + j - 1];
                    const float dp = (j < VD - 1) ? cur[i][j + 1] : L[lr][VD + lane * VD + j + 1];
                    float acc = prv[i][j] + nxt[i][j];
                    acc = acc + wm;
                    acc = acc + wp;
                    acc = acc + dm;
                    acc = acc + dp;
                    o[j] = encode(acc, ra[i][j] ^ rbh);
                }
            } else {
                const bool wface = !hint || !row_int[i];
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    const int dd = d + j;
                    if (wface || dd < r || dd >= D - r) {
                        o[j] = qr[CUR][i][j];
                    } else {
                        const float wm = (i > 0) ? cur[i - 1][j] : L[lr - 1][VD + lane * VD + j];
                        const float wp = (i < RW - 1) ? cur[i + 1][j] : L[lr + 1][VD + lane * VD + j];
                        const float dm = (j > 0) ? cur[i][j - 1] : L[lr][VD + lane * VD + j - 1];
                        const float dp = (j < VD - 1) ? cur[i][j + 1] : L[lr][VD + lane * VD + j + 1];
                        float acc = prv[i][j] + nxt[i][j];
                        acc = acc + wm;
                        acc = acc + wp;
                        acc = acc + dm;
                        acc = acc + dp;
                        o[j] = encode(acc, ra[i][j] ^ rbh);
                    }
                }
            }
            if constexpr (RES) {
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    const double dv = (double)Tr::dec_raw(o[j]) - (double)Tr::dec_raw(qr[CUR][i][j]);
                    rr += dv * dv;
                }
            }
            store_vec<T, VD>(pout_h + i * pitch, o);
        }
        b ^= 1;
    };

    if constexpr (REN) {
        for (int h = h0; h < h1; h += 3) {
            // the tail test is uniform across the block (h1 is)
            unroll_seq([&](auto ph) __attribute__((always_inline)) {
                if (h + decltype(ph)::value < h1) step(ph, h + decltype(ph)::value);
            }, std::make_integer_sequence<int, 3>{});
        }
    } else {
        for (int h = h0; h < h1; ++h) {
            step(std::integral_constant<int, 0>{}, h);
            // shift the queue by one plane
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                qd[0][i] = qd[1][i];
                qd[1][i] = qd[2][i];
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    pl[0][i][j] = pl[1][i][j];
                    pl[1][i][j] = pl[2][i][j];
                    qr[0][i][j] = qr[1][i][j];
                    qr[1][i][j] = qr[2][i][j];
                }
            }
#pragma unroll
            for (int j = 0; j < VD; ++j) {
                qw[0][j] = qw[1][j];
                qw[1][j] = qw[2][j];
            }
        }
    }

    if constexpr (RES) {
        const double t = block_sum<NW>(rr, red);
        if (threadIdx.x == 0) s.partials[lid] = t;
    }
}

template <int ST, int VD, int RW>
hipError_t march_launch(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st)
{
    if constexpr (march_lds_bytes(VD, RW) > 64 * 1024) {
        return hipErrorInvalidValue;
    } else {
        const dim3 grid((unsigned)m.ntd * m.ntw * m.nseg), block(256);
        const bool sr = ST != PSKEL_DTYPE_F32 && s.rounding == PSKEL_ROUND_STOCHASTIC;
#define PSKEL_MARCH(RESV, SRV)                                                                     \
    hipLaunchKernelGGL((k3d_march<ST, VD, RW, RESV, SRV>), grid, block, 0, st, s, p.coeff, m.hseg, \
                       m.ntd, m.ntw, m.nseg)
        if constexpr (ST == PSKEL_DTYPE_F32) {
            if (s.partials) PSKEL_MARCH(true, false);
            else PSKEL_MARCH(false, false);
        } else {
            if (s.partials && sr) PSKEL_MARCH(true, true);
            else if (s.partials) PSKEL_MARCH(true, false);
            else if (sr) PSKEL_MARCH(false, true);
            else PSKEL_MARCH(false, false);
        }
#undef PSKEL_MARCH
        return hipGetLastError();
    }
}

template <int ST, int VD>
hipError_t march_rw(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st)
{
    switch (m.rw) {
    case 1: return march_launch<ST, VD, 1>(s, p, m, st);
    case 2: return march_launch<ST, VD, 2>(s, p, m, st);
    case 4: return march_launch<ST, VD, 4>(s, p, m, st);
    case 8: return march_launch<ST, VD, 8>(s, p, m, st);
    default: return hipErrorInvalidValue;
    }
}

template <int ST>
hipError_t march_vd(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st)
{
    switch (m.vd) {
    case 1: return march_rw<ST, 1>(s, p, m, st);
    case 2: return march_rw<ST, 2>(s, p, m, st);
    case 4: return march_rw<ST, 4>(s, p, m, st);
    case 8:
        if constexpr (ST != PSKEL_DTYPE_F32) return march_rw<ST, 8>(s, p, m, st);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

}  // namespace
#endif  // PSKEL_MARCH_INSTANTIATE

}  // namespace pskel
