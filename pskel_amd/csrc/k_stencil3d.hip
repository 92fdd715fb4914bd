// 3D stencil sweeps for gfx950, with optional fused FP64 residual.
//
// Reference contract: Stencil3D visits every cell and leaves the faces to the
// user kernel (include/PSkelStencil.hpp:2095-2113); Mask3D::get returns the halo
// value outside the grid (include/PSkelMask.hpp:186-199).  Layout of Array3D is
// (h*W + w)*D + d, so d is contiguous (include/PSkelArray.hpp:642); here rows of
// d are pitched.  The reference's 3D GPU kernel maps lanes to w, which is not
// contiguous (include/PSkelStencil.hpp:685-701); every kernel here puts lanes
// on d.
//
//   k3d_plain  one lane per cell, block (bx along d, by along w), grid z = h.
//   k3d_march  2.5D marching: a block owns a (w, d) tile, walks a segment of h
//              planes keeping h-1/h/h+1 in registers and staging plane h (with
//              its one-cell w/d halo) in a double-buffered LDS tile, one barrier
//              per plane.  Vector loads of VD cells per lane along d.
//
// The residual ||u_new - u_old||_2 is accumulated in FP64 per block (fixed
// reduction order) and reduced by a second single-block pass; FP64 is used for
// nothing else on the device.
#include "k_common.hpp"

namespace pskel {

namespace {

// ---------------------------------------------------------------------------
// plain
// ---------------------------------------------------------------------------
template <int ST, int KIND, int NP>
__global__ __launch_bounds__(1024) void k3d_plain(Sweep3D s, Points p)
{
    using Tr = Stor<ST>;
    using T = typename Tr::T;
    __shared__ double red[16];
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = blockIdx.y * blockDim.y + threadIdx.y;
    const int h = blockIdx.z;
    const int W = s.W, H = s.H, D = s.D, r = s.radius;
    const size_t pitch = s.pitch;
    const T* __restrict__ in = (const T*)s.src;
    T* __restrict__ out = (T*)s.dst;
    const float off = s.store_offset;
    double rr = 0.0;
    if (d < D && w < W) {
        const size_t idx = ((size_t)h * W + w) * pitch + d;
        const bool edge = h < r || h >= H - r || w < r || w >= W - r || d < r || d >= D - r;
        T res;
        if (edge && p.boundary == PSKEL_BOUNDARY_COPY) {
            res = in[idx];
        } else {
            const int np = NP > 0 ? NP : p.np;
            float acc = 0.f;
#pragma unroll
            for (int i = 0; i < (NP > 0 ? NP : kMaxPoints); ++i) {
                if (NP == 0 && i >= np) break;
                const int hh = h + p.dh[i], ww = w + p.dw[i], dd = d + p.dd[i];
                float v;
                if (!edge || ((unsigned)hh < (unsigned)H && (unsigned)ww < (unsigned)W &&
                              (unsigned)dd < (unsigned)D))
                    v = Tr::dec(in[((size_t)hh * W + ww) * pitch + dd], off);
                else
                    v = p.halo;
                if constexpr (KIND == PSKEL_KIND_SCALED_SUM) acc = (i == 0) ? v : acc + v;
                else acc = (i == 0) ? p.wt[i] * v : acc + p.wt[i] * v;
            }
            if constexpr (KIND == PSKEL_KIND_SCALED_SUM) acc = p.coeff * acc;
            res = Tr::enc(acc, off);
        }
        out[idx] = res;
        if (s.partials) {
            const double dv = (double)Tr::dec_raw(res) - (double)Tr::dec_raw(in[idx]);
            rr = dv * dv;
        }
    }
    if (s.partials) {
        const int nthreads = blockDim.x * blockDim.y;
        const int tid = threadIdx.y * blockDim.x + threadIdx.x;
        // block sizes are multiples of 64 (checked on the host)
        const int lane = tid % kWave, wv = tid / kWave;
        rr = wave_sum(rr);
        if (lane == 0) red[wv] = rr;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int i = 0; i < nthreads / kWave; ++i) t += red[i];
            const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
            s.partials[b] = t;
        }
    }
}

template <int ST, int KIND>
hipError_t plain3d_np(const Sweep3D& s, const Points& p, dim3 grid, dim3 block, hipStream_t st)
{
    switch (p.np) {
    case 6: hipLaunchKernelGGL((k3d_plain<ST, KIND, 6>), grid, block, 0, st, s, p); break;
    case 7: hipLaunchKernelGGL((k3d_plain<ST, KIND, 7>), grid, block, 0, st, s, p); break;
    default: hipLaunchKernelGGL((k3d_plain<ST, KIND, 0>), grid, block, 0, st, s, p); break;
    }
    return hipGetLastError();
}

template <int ST>
hipError_t plain3d_kind(const Sweep3D& s, const Points& p, dim3 grid, dim3 block, hipStream_t st)
{
    if (p.kind == PSKEL_KIND_SCALED_SUM) return plain3d_np<ST, PSKEL_KIND_SCALED_SUM>(s, p, grid, block, st);
    return plain3d_np<ST, PSKEL_KIND_WEIGHTED_SUM>(s, p, grid, block, st);
}

// ---------------------------------------------------------------------------
// marching kernel (7-point scaled sum, COPY faces)
// ---------------------------------------------------------------------------
constexpr int kMarchWaves = 4;

template <int ST, int VD, int RW, bool RES>
__global__ __launch_bounds__(256) void k3d_march(Sweep3D s, float coeff, int hseg, int ntd,
                                                 int ntw, int nseg)
{
    using Tr = Stor<ST>;
    using T = typename Tr::T;
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
    const float off = s.store_offset;

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

    // own cell (h, wr, d) of the plane being loaded / stored
    const T* __restrict__ pin = (const T*)s.src + (size_t)h0 * plane + (size_t)wr * pitch + d;
    T* __restrict__ pout = (T*)s.dst + (size_t)h0 * plane + (size_t)wr * pitch + d;

    // register queue along h: cur (h) and nxt (h+1) are resident while the loads
    // of plane h+2 are in flight, so a plane's load latency overlaps a whole step
    float prv[RW][VD], cur[RW][VD], nxt[RW][VD];
    T craw[RW][VD], nraw[RW][VD], fraw[RW][VD];
    T hw_c[VD], hw_n[VD], hw_f[VD];   // w-halo row vector (this wave's role)
    T hd_c[RW], hd_n[RW], hd_f[RW];   // d-halo cell per row (lane role)

    auto load_plane = [&](const T* p, T (&raw)[RW][VD], T (&hw)[VD], T (&hd)[RW]) {
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
    };
    auto decode = [&](float (&v)[RW][VD], const T (&raw)[RW][VD]) {
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
            for (int j = 0; j < VD; ++j) v[i][j] = Tr::dec(raw[i][j], off);
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

    if (h0 > 0) {
        T tr[RW][VD], thw[VD], thd[RW];
        load_plane(pin - plane, tr, thw, thd);
        decode(prv, tr);
    } else {
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
            for (int j = 0; j < VD; ++j) prv[i][j] = 0.f;
    }
    load_plane(pin, craw, hw_c, hd_c);
    if (h0 + 1 < H) load_plane(pin + plane, nraw, hw_n, hd_n);
    else zero_plane(nraw, hw_n, hd_n);
    decode(cur, craw);

    double rr = 0.0;
    int b = 0;
    for (int h = h0; h < h1; ++h) {
        if (h + 2 < H && h + 1 < h1) load_plane(pin + 2 * plane, fraw, hw_f, hd_f);
        else zero_plane(fraw, hw_f, hd_f);
        decode(nxt, nraw);
        // stage plane h
        float (*L)[LROW] = lds[b];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
#pragma unroll
            for (int j = 0; j < VD; ++j) L[1 + wave * RW + i][VD + lane * VD + j] = cur[i][j];
            if (hal_dlo) L[1 + wave * RW + i][VD - 1] = Tr::dec(hd_c[i], off);
            if (hal_dhi) L[1 + wave * RW + i][VD + TD] = Tr::dec(hd_c[i], off);
        }
        if (hal_wlo || hal_whi) {
            const int row = hal_wlo ? 0 : TW + 1;
#pragma unroll
            for (int j = 0; j < VD; ++j) L[row][VD + lane * VD + j] = Tr::dec(hw_c[j], off);
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
                    o[j] = Tr::enc(coeff * acc, off);
                }
            } else {
                const bool wface = !hint || !row_int[i];
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    const int dd = d + j;
                    if (wface || dd < r || dd >= D - r) {
                        o[j] = craw[i][j];
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
                        o[j] = Tr::enc(coeff * acc, off);
                    }
                }
            }
            if constexpr (RES) {
#pragma unroll
                for (int j = 0; j < VD; ++j) {
                    const double dv = (double)Tr::dec_raw(o[j]) - (double)Tr::dec_raw(craw[i][j]);
                    rr += dv * dv;
                }
            }
            store_vec<T, VD>(pout + i * pitch, o);
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            hd_c[i] = hd_n[i];
            hd_n[i] = hd_f[i];
#pragma unroll
            for (int j = 0; j < VD; ++j) {
                prv[i][j] = cur[i][j];
                cur[i][j] = nxt[i][j];
                craw[i][j] = nraw[i][j];
                nraw[i][j] = fraw[i][j];
            }
        }
#pragma unroll
        for (int j = 0; j < VD; ++j) {
            hw_c[j] = hw_n[j];
            hw_n[j] = hw_f[j];
        }
        b ^= 1;
        pin += plane;
        pout += plane;
    }

    if constexpr (RES) {
        const double t = block_sum<NW>(rr, red);
        if (threadIdx.x == 0) s.partials[lid] = t;
    }
}

struct MarchShape {
    int vd, rw, hseg, ntd, ntw, nseg;
};

MarchShape march_shape(const Sweep3D& s, int storage, const pskel_launch_cfg_t& cfg)
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

template <int ST, int VD, int RW>
hipError_t march_launch(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st)
{
    if constexpr (march_lds_bytes(VD, RW) > 64 * 1024) {
        return hipErrorInvalidValue;
    } else {
        const dim3 grid((unsigned)m.ntd * m.ntw * m.nseg), block(256);
        if (s.partials)
            hipLaunchKernelGGL((k3d_march<ST, VD, RW, true>), grid, block, 0, st, s, p.coeff, m.hseg,
                               m.ntd, m.ntw, m.nseg);
        else
            hipLaunchKernelGGL((k3d_march<ST, VD, RW, false>), grid, block, 0, st, s, p.coeff, m.hseg,
                               m.ntd, m.ntw, m.nseg);
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

// ---------------------------------------------------------------------------
// second-stage reduction: one block, fixed order
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_reduce_partials(const double* __restrict__ part, size_t n,
                                                          double* __restrict__ out)
{
    __shared__ double red[16];
    double a = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) a += part[i];
    const double t = block_sum<16>(a, red);
    if (threadIdx.x == 0) *out = sqrt(t);
}

}  // namespace

size_t plain3d_blocks(const Sweep3D& s, int bx, int by)
{
    return (size_t)((s.D + bx - 1) / bx) * ((s.W + by - 1) / by) * s.H;
}

hipError_t launch_3d_plain(const Sweep3D& s, const Points& p, int storage, int bx, int by,
                           hipStream_t st)
{
    const dim3 block(bx, by);
    const dim3 grid((s.D + bx - 1) / bx, (s.W + by - 1) / by, s.H);
    switch (storage) {
    case PSKEL_DTYPE_F32: return plain3d_kind<PSKEL_DTYPE_F32>(s, p, grid, block, st);
    case PSKEL_DTYPE_BF16: return plain3d_kind<PSKEL_DTYPE_BF16>(s, p, grid, block, st);
    case PSKEL_DTYPE_F16: return plain3d_kind<PSKEL_DTYPE_F16>(s, p, grid, block, st);
    default: return hipErrorInvalidValue;
    }
}

size_t tuned3d_blocks(const Sweep3D& s, int storage, const pskel_launch_cfg_t& cfg)
{
    const MarchShape m = march_shape(s, storage, cfg);
    return (size_t)m.ntd * m.ntw * m.nseg;
}

hipError_t launch_3d_tuned(const Sweep3D& s, const Points& p, int storage,
                           const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    const MarchShape m = march_shape(s, storage, cfg);
    switch (storage) {
    case PSKEL_DTYPE_F32: return march_vd<PSKEL_DTYPE_F32>(s, p, m, st);
    case PSKEL_DTYPE_BF16: return march_vd<PSKEL_DTYPE_BF16>(s, p, m, st);
    case PSKEL_DTYPE_F16: return march_vd<PSKEL_DTYPE_F16>(s, p, m, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reduce_partials(const double* partials, size_t n, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, partials, n, out);
    return hipGetLastError();
}

}  // namespace pskel
