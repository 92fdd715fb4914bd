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
#include "k_tiles2d.hpp"

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
// temporal tile: core CX x CY cells, loaded region (CX + 8) x (CY + 2K) with a
// 4-column aligned halo (K <= 4), K sweeps in LDS ping-pong.
// Cells outside the interior are never computed: at level j an LDS buffer holds
// the ring of the global buffer that level j lives in (level parity alternates
// between src and ring_other), so halo reads near the grid edge see exactly what
// a single sweep would see.
// ---------------------------------------------------------------------------
template <int PAT, int K, int CX, int CY>
__global__ __launch_bounds__(256) void k2d_ttile(Sweep2D s, Points p)
{
    constexpr int LX = CX + 8;
    constexpr int LY = CY + 2 * K;
    constexpr int NT = 256;
    __shared__ float A[LY * LX];
    __shared__ float B[LY * LX];
    const int tid = threadIdx.x;
    const int w0 = blockIdx.x * CX;            // core origin
    const int h0 = blockIdx.y * CY;
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
    const int col = tid % 128, strip = tid / 128;
#pragma unroll
    for (int j = 1; j <= K; ++j) {
        const float* __restrict__ P = (j & 1) ? A : B;
        float* __restrict__ Q = (j & 1) ? B : A;
        const int xlo = 4 - (K - j), xhi = 4 + CX + (K - j);
        const int ylo = j, yhi = LY - j;
        const int nx = xhi - xlo, ny = yhi - ylo;
        const int rows = (ny + NS - 1) / NS;
        if (col < nx) {
            const int x = xlo + col;
            const int gw = gx0 + x;
            const bool col_int = gw >= r && gw < Wd - r;
            const int ya = ylo + strip * rows, yb = min(ya + rows, yhi);
            const float* c = P + (ya - 1) * LX + x;
            float um = c[-1], u0 = c[0], up = c[1];                 // row y - 1
            float cm = c[LX - 1], c0 = c[LX], cp = c[LX + 1];       // row y
#pragma unroll 2
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
                const int gh = gy0 + y;
                if (col_int && gh >= r && gh < H - r) Q[y * LX + x] = acc;
                um = cm; u0 = c0; up = cp;
                cm = dm; c0 = d0; cp = dp;
            }
        }
        __syncthreads();
    }

    const float* __restrict__ R = (K & 1) ? B : A;
    float* __restrict__ dst = (float*)s.dst;
    for (int q = tid; q < CX * CY; q += NT) {
        const int y = q / CX, x = q % CX;
        const int gh = h0 + y, gw = w0 + x;
        if (gh < r || gh >= H - r || gw < r || gw >= Wd - r) continue;
        dst[(size_t)gh * pitch + gw] = R[(y + K) * LX + (x + 4)];
    }
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

}  // namespace

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
