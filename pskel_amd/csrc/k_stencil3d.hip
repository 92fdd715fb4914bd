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

template <int ST, int VD, int RW>
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
    const size_t pitch = s.pitch;
    const size_t plane = (size_t)W * pitch;
    const float off = s.store_offset;
    const T* __restrict__ in = (const T*)s.src;
    T* __restrict__ out = (T*)s.dst;

    const int d0 = td * TD, w0 = tw * TW;
    const int d = d0 + lane * VD;
    const int wr = w0 + wave * RW;
    const int h0 = sg * hseg;
    const int h1 = min(h0 + hseg, H);
    const bool dlive = d < D;

    // halo roles
    const bool hal_wlo = wave == 0 && w0 > 0;              // loads row w0-1
    const bool hal_whi = wave == NW - 1 && w0 + TW < W;     // loads row w0+TW
    const bool hal_dlo = lane == 0 && d0 > 0;               // loads d0-1 for its rows
    const bool hal_dhi = lane == kWave - 1 && d0 + TD < D;  // loads d0+TD for its rows

    float prv[RW][VD], cur[RW][VD], nxt[RW][VD];
    T craw[RW][VD], nraw[RW][VD];
    float hw_c[VD], hw_n[VD];  // w-halo row vector (this wave's role), current / next plane
    float hd_c[RW], hd_n[RW];  // d-halo scalar per row (lane role), current / next plane

    auto load_plane = [&](int h, float (&v)[RW][VD], T (&raw)[RW][VD], float (&hw)[VD],
                          float (&hd)[RW]) {
        const T* base = in + (size_t)h * plane;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int w = wr + i;
            if (dlive && w < W) {
                load_vec<T, VD>(base + (size_t)w * pitch + d, raw[i]);
            } else {
#pragma unroll
                for (int j = 0; j < VD; ++j) raw[i][j] = T(0);
            }
#pragma unroll
            for (int j = 0; j < VD; ++j) v[i][j] = Tr::dec(raw[i][j], off);
            hd[i] = 0.f;
            if (w < W) {
                if (hal_dlo) hd[i] = Tr::dec(base[(size_t)w * pitch + d0 - 1], off);
                if (hal_dhi) hd[i] = Tr::dec(base[(size_t)w * pitch + d0 + TD], off);
            }
        }
#pragma unroll
        for (int j = 0; j < VD; ++j) hw[j] = 0.f;
        if (dlive && (hal_wlo || hal_whi)) {
            const int w = hal_wlo ? w0 - 1 : w0 + TW;
            T t[VD];
            load_vec<T, VD>(base + (size_t)w * pitch + d, t);
#pragma unroll
            for (int j = 0; j < VD; ++j) hw[j] = Tr::dec(t[j], off);
        }
    };

    if (h0 > 0) {
        T tr[RW][VD];
        float thw[VD], thd[RW];
        load_plane(h0 - 1, prv, tr, thw, thd);
    } else {
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
            for (int j = 0; j < VD; ++j) prv[i][j] = 0.f;
    }
    load_plane(h0, cur, craw, hw_c, hd_c);

    double rr = 0.0;
    int b = 0;
    for (int h = h0; h < h1; ++h) {
        if (h + 1 < H) {
            load_plane(h + 1, nxt, nraw, hw_n, hd_n);
        } else {
#pragma unroll
            for (int i = 0; i < RW; ++i)
#pragma unroll
                for (int j = 0; j < VD; ++j) { nxt[i][j] = 0.f; nraw[i][j] = T(0); }
        }
        // stage plane h
#pragma unroll
        for (int i = 0; i < RW; ++i) {
#pragma unroll
            for (int j = 0; j < VD; ++j) lds[b][1 + wave * RW + i][VD + lane * VD + j] = cur[i][j];
            if (hal_dlo) lds[b][1 + wave * RW + i][VD - 1] = hd_c[i];
            if (hal_dhi) lds[b][1 + wave * RW + i][VD + TD] = hd_c[i];
        }
        if (hal_wlo || hal_whi) {
            const int row = hal_wlo ? 0 : TW + 1;
#pragma unroll
            for (int j = 0; j < VD; ++j) lds[b][row][VD + lane * VD + j] = hw_c[j];
        }
        __syncthreads();

        const bool hface = h < r || h >= H - r;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int w = wr + i;
            if (!dlive || w >= W) continue;
            const bool wface = hface || w < r || w >= W - r;
            const int lr = 1 + wave * RW + i;
            T o[VD];
#pragma unroll
            for (int j = 0; j < VD; ++j) {
                const int dd = d + j;
                if (wface || dd < r || dd >= D - r || dd >= D) {
                    o[j] = craw[i][j];
                } else {
                    const float wm = (i > 0) ? cur[i - 1][j] : lds[b][lr - 1][VD + lane * VD + j];
                    const float wp = (i < RW - 1) ? cur[i + 1][j] : lds[b][lr + 1][VD + lane * VD + j];
                    const float dm = (j > 0) ? cur[i][j - 1] : lds[b][lr][VD + lane * VD + j - 1];
                    const float dp = (j < VD - 1) ? cur[i][j + 1] : lds[b][lr][VD + lane * VD + j + 1];
                    float acc = prv[i][j] + nxt[i][j];
                    acc = acc + wm;
                    acc = acc + wp;
                    acc = acc + dm;
                    acc = acc + dp;
                    o[j] = Tr::enc(coeff * acc, off);
                }
                if (s.partials) {
                    const double dv = (double)Tr::dec_raw(o[j]) - (double)Tr::dec_raw(craw[i][j]);
                    rr += dv * dv;
                }
            }
            store_vec<T, VD>(out + (size_t)h * plane + (size_t)w * pitch + d, o);
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            hd_c[i] = hd_n[i];
#pragma unroll
            for (int j = 0; j < VD; ++j) {
                prv[i][j] = cur[i][j];
                cur[i][j] = nxt[i][j];
                craw[i][j] = nraw[i][j];
            }
        }
#pragma unroll
        for (int j = 0; j < VD; ++j) hw_c[j] = hw_n[j];
        b ^= 1;
    }

    if (s.partials) {
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
        hipLaunchKernelGGL((k3d_march<ST, VD, RW>), grid, block, 0, st, s, p.coeff, m.hseg, m.ntd,
                           m.ntw, m.nseg);
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
