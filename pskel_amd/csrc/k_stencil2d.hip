// 2D float stencil sweeps for gfx950.
//
// Reference contract: stencilTilingCU (include/PSkelStencil.hpp:572-652) calls
// the user kernel for every interior cell w in [r, W-r), h in [r, H-r); the ring
// is never written.  Every kernel here computes the same expression in the same
// order (no contraction: built with -ffp-contract=off), so all variants are
// bit-identical to each other and to runIterativeCPU.
//
//   k2d_plain  one lane per cell, the caller's (bx, by) block: the runCUDA shape.
//   k2d_tuned  wave64 x float4 row segments, RY rows per lane, neighbours along w
//              exchanged across lanes, all row loads issued before any math.
//   k2d_ttile  overlapped temporal tile: K sweeps in LDS per launch, K-cell halo
//              recomputed redundantly (the rewrite of PSkelStencilTiling).
//   k2d_march  K sweeps per launch pipelined in registers down a segment of
//              rows, DPP neighbour exchange, no LDS (k_march2d.hpp).
#include "k_tiles2d.hpp"
#define PSKEL_MARCH2D_INSTANTIATE
#include "k_march2d.hpp"

namespace pskel {

namespace {

constexpr int KIND_SCALED = PSKEL_KIND_SCALED_SUM;
constexpr int KIND_WEIGHTED = PSKEL_KIND_WEIGHTED_SUM;

// ---------------------------------------------------------------------------
// plain
// ---------------------------------------------------------------------------
template <int KIND, int NP>
__global__ __launch_bounds__(1024) void k2d_plain(Sweep2D s, Points p)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y * blockDim.y + threadIdx.y;
    const int r = s.radius;
    if (w < r || w >= s.W - r || h < r || h >= s.H - r) return;
    const float* __restrict__ c = (const float*)s.src + (size_t)h * s.pitch + w;
    const int np = NP > 0 ? NP : p.np;
    const ptrdiff_t pitch = (ptrdiff_t)s.pitch;
    float acc;
    if constexpr (KIND == KIND_SCALED) {
        acc = c[p.dh[0] * pitch + p.dw[0]];
#pragma unroll
        for (int i = 1; i < (NP > 0 ? NP : kMaxPoints); ++i) {
            if (NP == 0 && i >= np) break;
            acc = acc + c[p.dh[i] * pitch + p.dw[i]];
        }
        acc = p.coeff * acc;
    } else {
        acc = p.wt[0] * c[p.dh[0] * pitch + p.dw[0]];
#pragma unroll
        for (int i = 1; i < (NP > 0 ? NP : kMaxPoints); ++i) {
            if (NP == 0 && i >= np) break;
            acc = acc + p.wt[i] * c[p.dh[i] * pitch + p.dw[i]];
        }
    }
    ((float*)s.dst)[(size_t)h * s.pitch + w] = acc;
}

template <int KIND>
hipError_t plain_dispatch(const Sweep2D& s, const Points& p, dim3 grid, dim3 block, hipStream_t st)
{
    switch (p.np) {
    case 4: hipLaunchKernelGGL((k2d_plain<KIND, 4>), grid, block, 0, st, s, p); break;
    case 5: hipLaunchKernelGGL((k2d_plain<KIND, 5>), grid, block, 0, st, s, p); break;
    case 8: hipLaunchKernelGGL((k2d_plain<KIND, 8>), grid, block, 0, st, s, p); break;
    case 9: hipLaunchKernelGGL((k2d_plain<KIND, 9>), grid, block, 0, st, s, p); break;
    default: hipLaunchKernelGGL((k2d_plain<KIND, 0>), grid, block, 0, st, s, p); break;
    }
    return hipGetLastError();
}

template <int PAT, int RY>
__global__ __launch_bounds__(256) void k2d_tuned(Sweep2D s, Points p)
{
    tile_f32<PAT, RY>(s, p, blockIdx.x, blockIdx.y);
}

template <int PAT>
hipError_t tuned_dispatch(const Sweep2D& s, const Points& p, int ry, int nwy, hipStream_t st)
{
    const dim3 block(kWave, nwy);
#define PSKEL_TUNED_CASE(RYV)                                                                    \
    case RYV: {                                                                                  \
        const dim3 grid((s.W + 255) / 256, (s.H + nwy * RYV - 1) / (nwy * RYV));                 \
        hipLaunchKernelGGL((k2d_tuned<PAT, RYV>), grid, block, 0, st, s, p);                     \
        break;                                                                                   \
    }
    switch (ry) {
        PSKEL_TUNED_CASE(1)
        PSKEL_TUNED_CASE(2)
        PSKEL_TUNED_CASE(4)
        PSKEL_TUNED_CASE(8)
        PSKEL_TUNED_CASE(16)
    default: return hipErrorInvalidValue;
    }
#undef PSKEL_TUNED_CASE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// temporal tile: K sweeps per launch in LDS (the body is ttile_body in
// k_tiles2d.hpp, shared with the batched kernel)
// ---------------------------------------------------------------------------
template <int PAT, int K, int CX, int CY>
__global__ __launch_bounds__(256) void k2d_ttile(Sweep2D s, Points p)
{
    __shared__ float A[ttile_lds<K, CX, CY>()];
    __shared__ float B[ttile_lds<K, CX, CY>()];
    ttile_body<PAT, K, CX, CY>(s, p, (int)blockIdx.x, (int)blockIdx.y, K, (int)threadIdx.x, A, B);
}

template <int PAT, int K>
hipError_t ttile_launch_k(const Sweep2D& s, const Points& p, int cy, hipStream_t st)
{
    constexpr int CX = 120;
    const dim3 block(256);
    switch (cy) {
    case 8: {
        const dim3 grid((s.W + CX - 1) / CX, (s.H + 7) / 8);
        hipLaunchKernelGGL((k2d_ttile<PAT, K, CX, 8>), grid, block, 0, st, s, p);
        break;
    }
    case 16: {
        const dim3 grid((s.W + CX - 1) / CX, (s.H + 15) / 16);
        hipLaunchKernelGGL((k2d_ttile<PAT, K, CX, 16>), grid, block, 0, st, s, p);
        break;
    }
    case 32: {
        const dim3 grid((s.W + CX - 1) / CX, (s.H + 31) / 32);
        hipLaunchKernelGGL((k2d_ttile<PAT, K, CX, 32>), grid, block, 0, st, s, p);
        break;
    }
    case 64: {
        const dim3 grid((s.W + CX - 1) / CX, (s.H + 63) / 64);
        hipLaunchKernelGGL((k2d_ttile<PAT, K, CX, 64>), grid, block, 0, st, s, p);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int PAT>
hipError_t ttile_dispatch(const Sweep2D& s, const Points& p, int k, int cy, hipStream_t st)
{
    switch (k) {
    case 2: return ttile_launch_k<PAT, 2>(s, p, cy, st);
    case 3: return ttile_launch_k<PAT, 3>(s, p, cy, st);
    case 4: return ttile_launch_k<PAT, 4>(s, p, cy, st);
    default: return hipErrorInvalidValue;
    }
}

template <int PAT, int PD>
hipError_t march2d_dispatch(const Sweep2D& s, const Points& p, int k, int seg, hipStream_t st)
{
    if (!march2d_fits(s.W, s.H, s.pitch)) return hipErrorInvalidValue;
    const March2DShape m = march2d_shape(s.W, s.H, seg);
    if (m.units == 0) return hipSuccess;  // no interior cells: a sweep writes nothing
    const dim3 grid((unsigned)((m.units + kMarch2DWaves - 1) / kMarch2DWaves)), block(kMarch2DWaves * kWave);
    switch (k) {
    case 2: hipLaunchKernelGGL((k2d_march<PAT, 2, PD>), grid, block, 0, st, s, p, m.ntw, m.seg, m.units); break;
    case 3: hipLaunchKernelGGL((k2d_march<PAT, 3, PD>), grid, block, 0, st, s, p, m.ntw, m.seg, m.units); break;
    case 4: hipLaunchKernelGGL((k2d_march<PAT, 4, PD>), grid, block, 0, st, s, p, m.ntw, m.seg, m.units); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

bool march2d_ok(int W, int H, size_t pitch) { return march2d_fits(W, H, pitch); }

hipError_t launch_2d_march_f32(const Sweep2D& s, const Points& p, int pattern, int k,
                               const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    if (cfg.tile_x != kMarch2DVd || cfg.tile_y != 0 || cfg.tile_z < 0) return hipErrorInvalidValue;
    // input rows loaded 3 ahead: deeper prefetch measured no faster (the
    // 9-point march is bound by VALU issue, the 5-point one by its short waves)
    if (pattern == PAT_J5) return march2d_dispatch<PAT_J5, 3>(s, p, k, cfg.tile_z, st);
    if (pattern == PAT_BOX9) return march2d_dispatch<PAT_BOX9, 3>(s, p, k, cfg.tile_z, st);
    return hipErrorInvalidValue;
}

hipError_t launch_2d_plain_f32(const Sweep2D& s, const Points& p, int bx, int by, hipStream_t st)
{
    const dim3 block(bx, by);
    const dim3 grid((s.W + bx - 1) / bx, (s.H + by - 1) / by);
    if (p.kind == PSKEL_KIND_SCALED_SUM) return plain_dispatch<KIND_SCALED>(s, p, grid, block, st);
    return plain_dispatch<KIND_WEIGHTED>(s, p, grid, block, st);
}

hipError_t launch_2d_tuned_f32(const Sweep2D& s, const Points& p, int pattern,
                               const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    const int ry = cfg.tile_y > 0 ? cfg.tile_y : 4;
    const int nwy = cfg.tile_z > 0 ? cfg.tile_z : 4;
    if (nwy < 1 || nwy > 4) return hipErrorInvalidValue;
    if (pattern == PAT_J5) return tuned_dispatch<PAT_J5>(s, p, ry, nwy, st);
    if (pattern == PAT_BOX9) return tuned_dispatch<PAT_BOX9>(s, p, ry, nwy, st);
    return hipErrorInvalidValue;
}

hipError_t launch_2d_ttile_f32(const Sweep2D& s, const Points& p, int pattern, int k,
                               const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    const int cy = cfg.tile_y > 0 ? cfg.tile_y : 32;
    if (pattern == PAT_J5) return ttile_dispatch<PAT_J5>(s, p, k, cy, st);
    if (pattern == PAT_BOX9) return ttile_dispatch<PAT_BOX9>(s, p, k, cy, st);
    return hipErrorInvalidValue;
}

}  // namespace pskel
