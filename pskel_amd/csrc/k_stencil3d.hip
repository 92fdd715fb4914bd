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
//   k3d_march2 two fused sweeps per launch of the 7-point scaled sum, defined in
//              k_march3d_t2.hpp and dispatched from launch_3d_ttile below.
//
// The residual ||u_new - u_old||_2 is accumulated in FP64 per block (fixed
// reduction order) and reduced by a second single-block pass; FP64 is used for
// nothing else on the device.
#include "k_march3d.hpp"
#include "k_march3d_t2.hpp"

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
    // 16-bit scaled sums add the stored values (u - off) and fold the offset
    // into one fused multiply-add (scaled_sum_offset_term); weighted sums
    // decode u = x + off, as f32 storage holds u itself
    constexpr bool DEV = ST != PSKEL_DTYPE_F32 && KIND == PSKEL_KIND_SCALED_SUM;
    const float halo = DEV ? p.halo - off : p.halo;
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
            if (!edge) {
                // interior: every point in the grid, at the cell's index plus the
                // point's element offset (uniform: computed on the scalar unit)
#pragma unroll
                for (int i = 0; i < (NP > 0 ? NP : kMaxPoints); ++i) {
                    if (NP == 0 && i >= np) break;
                    const long long po = ((long long)p.dh[i] * W + p.dw[i]) * (long long)pitch + p.dd[i];
                    const T x = in[(long long)idx + po];
                    const float v = DEV ? Tr::dec_raw(x) : Tr::dec(x, off);
                    if constexpr (KIND == PSKEL_KIND_SCALED_SUM) acc = (i == 0) ? v : acc + v;
                    else acc = (i == 0) ? p.wt[i] * v : acc + p.wt[i] * v;
                }
            } else {
#pragma unroll
                for (int i = 0; i < (NP > 0 ? NP : kMaxPoints); ++i) {
                    if (NP == 0 && i >= np) break;
                    const int hh = h + p.dh[i], ww = w + p.dw[i], dd = d + p.dd[i];
                    float v;
                    if ((unsigned)hh < (unsigned)H && (unsigned)ww < (unsigned)W && (unsigned)dd < (unsigned)D) {
                        const T x = in[((size_t)hh * W + ww) * pitch + dd];
                        v = DEV ? Tr::dec_raw(x) : Tr::dec(x, off);
                    } else {
                        v = halo;
                    }
                    if constexpr (KIND == PSKEL_KIND_SCALED_SUM) acc = (i == 0) ? v : acc + v;
                    else acc = (i == 0) ? p.wt[i] * v : acc + p.wt[i] * v;
                }
            }
            if constexpr (DEV) acc = __builtin_fmaf(p.coeff, acc, s.c0);
            else if constexpr (KIND == PSKEL_KIND_SCALED_SUM) acc = p.coeff * acc;
            if (s.rounding == PSKEL_ROUND_STOCHASTIC) {
                const uint32_t key = sweep_key(s.seed, s.sweep);
                const uint32_t rw = sr_cell_bits(sr_cell_key(s.seed), (uint32_t)w * D + d) ^
                                    sr_plane_bits(key, s.plane0 + h);
                res = DEV ? Tr::enc_sr_raw(acc, rw) : Tr::enc_sr(acc, off, rw);
            } else {
                res = DEV ? Tr::enc_raw(acc) : Tr::enc(acc, off);
            }
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
    case PSKEL_DTYPE_F32: return march3d_f32(s, p, m, st);
    case PSKEL_DTYPE_BF16: return march3d_bf16(s, p, m, st);
    case PSKEL_DTYPE_F16: return march3d_f16(s, p, m, st);
    default: return hipErrorInvalidValue;
    }
}

size_t ttile3d_blocks(const Sweep3D& s, int storage, const pskel_launch_cfg_t& cfg, int depth)
{
    const March2Shape m = march2_shape(s, cfg, depth);
    switch (storage) {
    case PSKEL_DTYPE_F32: return march3d_t2_f32_blocks(s, m);
    case PSKEL_DTYPE_BF16: return march3d_t2_bf16_blocks(s, m);
    case PSKEL_DTYPE_F16: return march3d_t2_f16_blocks(s, m);
    default: return 0;
    }
}

bool ttile3d_shape_ok(const pskel_launch_cfg_t& cfg, int storage, int depth)
{
    if (cfg.block_x > 0 && cfg.block_x % kWave != 0) return false;
    Sweep3D s{};
    s.W = s.H = s.D = 1;
    const March2Shape m = march2_shape(s, cfg, depth);
    const int threads = cfg.block_x > 0 ? cfg.block_x : (depth == 3 ? 8 : 12) * kWave;
    return m.nwd > 0 && m.nww > 0 && m.nwd * m.nww * kWave == threads && march2_supported(m, storage);
}

hipError_t launch_3d_ttile(const Sweep3D& s, const Points& p, int storage, const pskel_launch_cfg_t& cfg, int depth,
                           hipStream_t st)
{
    const March2Shape m = march2_shape(s, cfg, depth);
    switch (storage) {
    case PSKEL_DTYPE_F32: return march3d_t2_f32(s, p, m, st);
    case PSKEL_DTYPE_BF16: return march3d_t2_bf16(s, p, m, st);
    case PSKEL_DTYPE_F16: return march3d_t2_f16(s, p, m, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reduce_partials(const double* partials, size_t n, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, partials, n, out);
    return hipGetLastError();
}

}  // namespace pskel
