// Internal declarations shared by the host layer and the kernel translation units.
// Nothing here is part of the public C ABI (include/pskel_amd.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pskel_amd.h"

namespace pskel {

// ---------------------------------------------------------------------------
// Error plumbing: every C entry point returns a status and records a message.
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
pskel_status_t fail(pskel_status_t st, const std::string& msg);
pskel_status_t hip_fail(hipError_t e, const char* what);

// ---------------------------------------------------------------------------
// Host-side data movement on several threads (at most 16).
// ---------------------------------------------------------------------------
// f(a, b) over disjoint ranges covering [0, n), pieces of at least `grain`.
template <class F>
void par_range(size_t n, size_t grain, F f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(hw, 16), (n + grain - 1) / grain);
    if (nt <= 1) {
        if (n) f(size_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (size_t t = 0; t < nt; ++t) {
        const size_t a = t * per, b = std::min(n, a + per);
        if (a >= b) break;
        th.emplace_back(f, a, b);
    }
    for (auto& x : th) x.join();
}

inline void par_copy(void* dst, const void* src, size_t bytes)
{
    par_range(bytes, size_t(2) << 20,
              [=](size_t a, size_t b) { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
}

#define PSKEL_HIP(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return ::pskel::hip_fail(e_, #call);             \
    } while (0)

// ---------------------------------------------------------------------------
// Stencil descriptor as the kernels see it (kernel argument, by value).
// ---------------------------------------------------------------------------
constexpr int kMaxPoints = PSKEL_MAX_POINTS;

struct Points {
    int32_t np;
    int32_t radius;
    int32_t kind;
    int32_t boundary;
    float coeff;
    float halo;
    int32_t dh[kMaxPoints];
    int32_t dw[kMaxPoints];
    int32_t dd[kMaxPoints];
    float wt[kMaxPoints];
};

// Point-set patterns the tuned kernels are specialised for.  The tuned kernels
// accumulate in exactly this order, so a descriptor only qualifies when its
// points are listed in this order (life: any order, the sum is integer).
enum Pattern : int {
    PAT_GENERIC = 0,
    PAT_J5 = 1,   // scaled sum (-1,0) (1,0) (0,-1) (0,1)
    PAT_BOX9 = 2, // weighted sum over the 3x3 box, row-major
    PAT_LIFE8 = 3,// life over the 8 Moore neighbours
    PAT_J7 = 4,   // 3D scaled sum (-1,0,0) (1,0,0) (0,-1,0) (0,1,0) (0,0,-1) (0,0,1)
    PAT_CLOUD = 5,// 2D CLOUDSIM (wind advection + conduction, two field arguments)
};

int classify(const pskel_stencil_desc_t& d, int ndim);
int desc_radius(const pskel_stencil_desc_t& d, int ndim);
void fill_points(Points& p, const pskel_stencil_desc_t& d, int ndim);

// ---------------------------------------------------------------------------
// Grid: device-resident ping-pong pair plus an optional spare buffer.
// 2D storage is row-pitched: cell (h, w) at h*pitch + w.
// 3D storage pitches the contiguous d axis: (h*W + w)*pitch + d.
// pitch (in elements) keeps every row 256-byte aligned and >= the tuned tile.
// ---------------------------------------------------------------------------
struct Grid {
    int ndim = 2;
    int storage = PSKEL_DTYPE_F32;
    int elem = 4;
    size_t W = 0, H = 0, D = 1;
    size_t pitch = 0;   // elements per contiguous row
    size_t nrows = 0;   // 2D: H, 3D: H*W
    size_t bytes = 0;   // per buffer
    float store_offset = 0.0f;
    bool auto_offset = false;  // store_offset is the mean of the last INPUT upload
    int device = 0;
    int rounding = PSKEL_ROUND_NEAREST;
    uint32_t seed = 0;
    uint64_t sweeps_done = 0;  // sweeps since the last INPUT upload (stochastic rounding stream)
    uint32_t plane_origin = 0; // global h of plane 0 (a slab of a host-staged grid), for the rounding words

    void* buf[3] = {nullptr, nullptr, nullptr}; // [cur, other, spare]
    int parity = 0;       // which uploaded buffer's ring `cur` carries (0 = input)
    int spare_parity = -1;// ring carried by the spare (-1: not allocated / unknown)

    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    double* d_partials = nullptr;   // per-block FP64 residual partials
    size_t partials_cap = 0;
    double* d_res = nullptr;        // reduced residuals, one per residual sweep
    size_t res_cap = 0;

    pskel_field_t* fields[4] = {nullptr, nullptr, nullptr, nullptr};  // attached field arguments
    bool odd_2d_launches = false;  // 2D launches fuse an odd number of sweeps (no spare buffer)
};

void advance_state(Grid& g, size_t k);

// ---------------------------------------------------------------------------
// Launch plumbing shared by the kernel TUs.
// ---------------------------------------------------------------------------
struct Sweep2D {
    const void* src;
    void* dst;
    const void* ring_other; // buffer carrying the other parity's ring (temporal tiles)
    int W, H;
    size_t pitch;
    int radius;
};

struct Sweep3D {
    const void* src;
    void* dst;
    int W, H, D;
    size_t pitch;
    int radius;
    float store_offset;
    double* partials;   // non-null: accumulate sum((new-old)^2) per block
    int rounding;       // pskel_rounding_t of 16-bit stores
    uint32_t seed;
    uint64_t sweep;     // index of the sweep this launch performs
    uint32_t plane0;    // global h of plane 0 (stochastic rounding words use the global plane)
    float c0;           // 16-bit scaled sums: scaled_sum_offset_term of the stencil and offset
};

// 16-bit storage holds x = u - off.  A scaled sum of n stored cells is then
// coeff * sum(x_i + off) - off = coeff * sum(x_i) + (coeff * n - 1) * off, so
// the kernels sum the stored values as they are and add this term once, in
// one fused multiply-add: fma(coeff, sum(x_i), c0).  It is evaluated in double
// and rounded once to float, on the host.
inline float scaled_sum_offset_term(float coeff, int npoints, float off)
{
    return (float)(((double)coeff * (double)npoints - 1.0) * (double)off);
}

// CLOUDSIM field arguments as its kernels see them (reversed device layout:
// the field value of cell (h, w) at h*fpitch + w).
struct CloudArgs {
    const float* wx;
    const float* wy;
    size_t fpitch;
    float dt;
};

hipError_t launch_cloud_plain(const Sweep2D& s, const CloudArgs& a, int bx, int by, hipStream_t st);
hipError_t launch_cloud_tuned(const Sweep2D& s, const CloudArgs& a, const pskel_launch_cfg_t& cfg, hipStream_t st);
bool field_info(const pskel_field_t* f, int* ndim, size_t* W, size_t* H, size_t* D, const float** data,
                size_t* pitch);

// Planned launch shape for a residual-reduction: number of blocks a 3D sweep uses.
size_t sweep3d_blocks(const Grid& g, int pattern, const pskel_launch_cfg_t& cfg);

// 2D float (f32) sweeps; return hipSuccess or the launch error.
hipError_t launch_2d_plain_f32(const Sweep2D& s, const Points& p, int bx, int by, hipStream_t st);
hipError_t launch_2d_tuned_f32(const Sweep2D& s, const Points& p, int pattern,
                               const pskel_launch_cfg_t& cfg, hipStream_t st);
hipError_t launch_2d_ttile_f32(const Sweep2D& s, const Points& p, int pattern, int k,
                               const pskel_launch_cfg_t& cfg, hipStream_t st);
// k fused sweeps (2..4) marched in registers: cfg.tile_x = 4 cells per lane,
// cfg.tile_z = rows per segment (0: automatic)
hipError_t launch_2d_march_f32(const Sweep2D& s, const Points& p, int pattern, int k,
                               const pskel_launch_cfg_t& cfg, hipStream_t st);
// a grid the register march takes: wider than 8 cells, buffer under 2 GiB
bool march2d_ok(int W, int H, size_t pitch);

// 2D life (u8) sweeps.
hipError_t launch_life_plain(const Sweep2D& s, const Points& p, int bx, int by, hipStream_t st);
hipError_t launch_life_tuned(const Sweep2D& s, const pskel_launch_cfg_t& cfg, hipStream_t st);
hipError_t launch_life_bits(const Sweep2D& s, int k, const pskel_launch_cfg_t& cfg, hipStream_t st);

// 3D sweeps for every storage type.
hipError_t launch_3d_plain(const Sweep3D& s, const Points& p, int storage, int bx, int by,
                           hipStream_t st);
hipError_t launch_3d_tuned(const Sweep3D& s, const Points& p, int storage,
                           const pskel_launch_cfg_t& cfg, hipStream_t st);
size_t plain3d_blocks(const Sweep3D& s, int bx, int by);
size_t tuned3d_blocks(const Sweep3D& s, int storage, const pskel_launch_cfg_t& cfg);
// two fused sweeps per launch (7-point scaled sum, any storage); cfg's tile
// fields describe the fused kernel's tile
// `depth` fused sweeps (2 or 3) of the 7-point Jacobi per launch
hipError_t launch_3d_ttile(const Sweep3D& s, const Points& p, int storage, const pskel_launch_cfg_t& cfg, int depth,
                           hipStream_t st);
size_t ttile3d_blocks(const Sweep3D& s, int storage, const pskel_launch_cfg_t& cfg, int depth);
// an instantiated fused tile shape at that depth
bool ttile3d_shape_ok(const pskel_launch_cfg_t& cfg, int storage, int depth);

// Deterministic second-stage reduction: out = sqrt(sum(partials[0..n))).
hipError_t launch_reduce_partials(const double* partials, size_t n, double* out, hipStream_t st);

// Utilities.
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
hipError_t launch_fill_zero(void* dst, size_t bytes, hipStream_t st);
hipError_t launch_copy_ring2d(const Sweep2D& s, int elem, hipStream_t st);

}  // namespace pskel

// The opaque grid of the C ABI: the resident pair plus its captured graphs.
struct GraphEntry {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    int64_t launches = 0;
    int64_t residuals = 0;
};

struct pskel_grid : pskel::Grid {
    std::vector<GraphEntry> graphs;
    // pinned staging for uploads and downloads: two buffers, filled and sent
    // in turn (allocated by the first transfer, at most 16 MiB each)
    void* stage[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

namespace pskel {
// Issue sweeps without waiting for them (the host-staged pipeline; pskel_api.cpp).
pskel_status_t run_async(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t iterations,
                         const pskel_launch_cfg_t& cfg, hipStream_t stream, int64_t* launches);
}  // namespace pskel
