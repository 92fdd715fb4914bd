// C ABI of the stencil library: device residency, launch sequencing, graphs.
//
// Host-side replacement of the reference's StencilBase GPU runners
// (include/PSkelStencil.hpp:1047-1097 runIterativeGPU, :1765-1792 runCUDA) and
// of ArrayBase's device residency (include/PSkelArray.hpp:72-95, 469-608).
// Differences that are deliberate:
//   * buffers stay resident across calls; sweeps only swap pointers;
//   * no host synchronisation inside the sweep loop (the reference syncs after
//     every launch, include/PSkelStencil.hpp:1791);
//   * both buffers are initialised (the reference never uploads `output`,
//     include/PSkelStencil.hpp:1076);
//   * failures return a status instead of exit() (include/PSkelDefs.h:50-56).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pskel_internal.hpp"

namespace pskel {

namespace {
thread_local std::string g_error;

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

int elem_size(int dtype)
{
    switch (dtype) {
    case PSKEL_DTYPE_U8: return 1;
    case PSKEL_DTYPE_F32: return 4;
    case PSKEL_DTYPE_F64: return 8;
    case PSKEL_DTYPE_BF16:
    case PSKEL_DTYPE_F16: return 2;
    default: return 0;
    }
}

// Host-side storage conversions (round to nearest even), matching the device's.
uint16_t f32_to_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

uint16_t f32_to_f16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) {
        if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu));
        return (uint16_t)(sign | 0x7c00u);
    }
    if (a < 0x38800000u) {
        float f;
        std::memcpy(&f, &a, 4);
        return (uint16_t)(sign | (uint32_t)std::nearbyint(f * 16777216.0f));
    }
    const uint32_t e = (a >> 23) - 127u + 15u;
    const uint32_t m = a & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t h)
{
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    const uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    const uint32_t m = (uint32_t)h & 0x3ffu;
    uint32_t u;
    if (e == 0) {
        const float v = (float)m * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    if (e == 31) u = sign | 0x7f800000u | (m << 13);
    else u = sign | ((e - 15u + 127u) << 23) | (m << 13);
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

}  // namespace

void set_error(const std::string& msg) { g_error = msg; }

pskel_status_t fail(pskel_status_t st, const std::string& msg)
{
    g_error = msg;
    return st;
}

pskel_status_t hip_fail(hipError_t e, const char* what)
{
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return PSKEL_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PSKEL_ERR_NO_DEVICE;
    if (e == hipErrorLaunchFailure || e == hipErrorInvalidConfiguration) return PSKEL_ERR_LAUNCH;
    return PSKEL_ERR_RUNTIME;
}

int desc_radius(const pskel_stencil_desc_t& d, int ndim)
{
    int r = 0;
    for (int i = 0; i < d.npoints; ++i)
        for (int k = 0; k < ndim; ++k) r = std::max(r, std::abs(d.offsets[ndim * i + k]));
    return d.radius > 0 ? d.radius : r;
}

static int desc_reach(const pskel_stencil_desc_t& d, int ndim)
{
    int r = 0;
    for (int i = 0; i < d.npoints; ++i)
        for (int k = 0; k < ndim; ++k) r = std::max(r, std::abs(d.offsets[ndim * i + k]));
    return r;
}

int classify(const pskel_stencil_desc_t& d, int ndim)
{
    const int* o = d.offsets;
    if (ndim == 2) {
        static const int j5[8] = {-1, 0, 1, 0, 0, -1, 0, 1};
        if (d.kind == PSKEL_KIND_SCALED_SUM && d.npoints == 4 && std::equal(j5, j5 + 8, o)) return PAT_J5;
        if (d.kind == PSKEL_KIND_WEIGHTED_SUM && d.npoints == 9) {
            bool ok = true;
            for (int i = 0; i < 9 && ok; ++i) ok = o[2 * i] == i / 3 - 1 && o[2 * i + 1] == i % 3 - 1;
            if (ok) return PAT_BOX9;
        }
        if (d.kind == PSKEL_KIND_CLOUDSIM) return PAT_CLOUD;
        if (d.kind == PSKEL_KIND_LIFE && d.npoints == 8) {
            int seen = 0;
            for (int i = 0; i < 8; ++i) {
                const int dh = o[2 * i], dw = o[2 * i + 1];
                if (dh < -1 || dh > 1 || dw < -1 || dw > 1 || (dh == 0 && dw == 0)) return PAT_GENERIC;
                seen |= 1 << ((dh + 1) * 3 + (dw + 1));
            }
            if (seen == (0x1ff & ~(1 << 4))) return PAT_LIFE8;
        }
    } else {
        static const int j7[18] = {-1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1};
        if (d.kind == PSKEL_KIND_SCALED_SUM && d.npoints == 6 && std::equal(j7, j7 + 18, o) &&
            d.boundary == PSKEL_BOUNDARY_COPY)
            return PAT_J7;
    }
    return PAT_GENERIC;
}

void fill_points(Points& p, const pskel_stencil_desc_t& d, int ndim)
{
    std::memset(&p, 0, sizeof(p));
    p.np = d.npoints;
    p.radius = desc_radius(d, ndim);
    p.kind = d.kind;
    p.boundary = d.boundary;
    p.coeff = d.coeff;
    p.halo = d.halo_value;
    for (int i = 0; i < d.npoints; ++i) {
        p.dh[i] = d.offsets[ndim * i];
        p.dw[i] = d.offsets[ndim * i + 1];
        p.dd[i] = ndim == 3 ? d.offsets[3 * i + 2] : 0;
        p.wt[i] = d.weights[i];
    }
}

}  // namespace pskel

using namespace pskel;

// ---------------------------------------------------------------------------
// the opaque grid type of the C ABI
// ---------------------------------------------------------------------------

namespace {

pskel_status_t check_grid(const pskel_grid_t* g)
{
    if (!g) return fail(PSKEL_ERR_INVALID_ARGUMENT, "grid is NULL");
    return PSKEL_OK;
}

Sweep2D make_sweep2d(const pskel_grid_t* g, const void* src, void* dst, const void* ring, int radius)
{
    Sweep2D s;
    s.src = src;
    s.dst = dst;
    s.ring_other = ring;
    s.W = (int)g->W;
    s.H = (int)g->H;
    s.pitch = g->pitch;
    s.radius = radius;
    return s;
}

Sweep3D make_sweep3d(const pskel_grid_t* g, const void* src, void* dst, int radius, double* partials)
{
    Sweep3D s;
    s.rounding = g->rounding;
    s.seed = g->seed;
    s.sweep = g->sweeps_done;
    s.plane0 = g->plane_origin;
    s.src = src;
    s.dst = dst;
    s.W = (int)g->W;
    s.H = (int)g->H;
    s.D = (int)g->D;
    s.pitch = g->pitch;
    s.radius = radius;
    s.store_offset = g->store_offset;
    s.partials = partials;
    s.c0 = 0.0f;
    return s;
}

// ---------------------------------------------------------------------------
// Host rows <-> device pitched rows.  Transfers go through two pinned staging
// buffers used in turn: the host side of one chunk of rows (a copy or a storage
// conversion, on several threads) overlaps the DMA of the previous chunk.
// A pageable source would be copied by the runtime one pitched row at a time.
// ---------------------------------------------------------------------------
constexpr size_t kStageBytes = size_t(16) << 20;

size_t cells(const pskel_grid_t* g) { return g->W * g->H * g->D; }
size_t row_cells(const pskel_grid_t* g) { return g->ndim == 2 ? g->W : g->D; }

pskel_status_t ensure_stage(pskel_grid_t* g)
{
    const size_t rb = row_cells(g) * g->elem;
    const size_t want = std::max(rb, std::min(kStageBytes, rb * g->nrows));
    // no copy of an earlier transfer (one that failed part-way) still reads or
    // writes the staging buffers
    PSKEL_HIP(hipStreamSynchronize(g->stream));
    if (g->stage_bytes >= want) return PSKEL_OK;
    for (int b = 0; b < 2; ++b) {
        if (g->stage[b]) (void)hipHostFree(g->stage[b]);
        g->stage[b] = nullptr;
    }
    g->stage_bytes = 0;
    for (int b = 0; b < 2; ++b) {
        PSKEL_HIP(hipHostMalloc(&g->stage[b], want, hipHostMallocDefault));
        if (!g->stage_ev[b]) PSKEL_HIP(hipEventCreateWithFlags(&g->stage_ev[b], hipEventDisableTiming));
    }
    g->stage_bytes = want;
    return PSKEL_OK;
}

// rows [r0, r0 + n) between a staging buffer (packed rows) and device memory
hipError_t stage_copy(pskel_grid_t* g, void* dbuf, void* stage, size_t r0, size_t n, hipMemcpyKind kind)
{
    const size_t rb = row_cells(g) * g->elem, pb = g->pitch * g->elem;
    char* dev = (char*)dbuf + r0 * pb;
    if (kind == hipMemcpyHostToDevice)
        return pb == rb ? hipMemcpyAsync(dev, stage, n * rb, kind, g->stream)
                        : hipMemcpy2DAsync(dev, pb, stage, rb, rb, n, kind, g->stream);
    return pb == rb ? hipMemcpyAsync(stage, dev, n * rb, kind, g->stream)
                    : hipMemcpy2DAsync(stage, rb, dev, pb, rb, n, kind, g->stream);
}

// fill(stage, first_cell, ncells) packs host cells into the staging buffer
template <class Fill>
pskel_status_t upload_rows(pskel_grid_t* g, void* dbuf, Fill fill)
{
    pskel_status_t st = ensure_stage(g);
    if (st != PSKEL_OK) return st;
    const size_t rc = row_cells(g), per = g->stage_bytes / (rc * g->elem);
    size_t i = 0;
    for (size_t r0 = 0; r0 < g->nrows; r0 += per, ++i) {
        const size_t n = std::min(per, g->nrows - r0);
        const int b = (int)(i & 1);
        if (i >= 2) PSKEL_HIP(hipEventSynchronize(g->stage_ev[b]));  // its previous chunk has left
        fill((unsigned char*)g->stage[b], r0 * rc, n * rc);
        PSKEL_HIP(stage_copy(g, dbuf, g->stage[b], r0, n, hipMemcpyHostToDevice));
        PSKEL_HIP(hipEventRecord(g->stage_ev[b], g->stream));
    }
    PSKEL_HIP(hipStreamSynchronize(g->stream));
    return PSKEL_OK;
}

// drain(stage, first_cell, ncells) unpacks the staging buffer into host cells
template <class Drain>
pskel_status_t download_rows(pskel_grid_t* g, const void* dbuf, Drain drain)
{
    pskel_status_t st = ensure_stage(g);
    if (st != PSKEL_OK) return st;
    const size_t rc = row_cells(g), per = g->stage_bytes / (rc * g->elem);
    const size_t chunks = (g->nrows + per - 1) / per;
    auto issue = [&](size_t i) -> hipError_t {
        const size_t r0 = i * per;
        hipError_t e = stage_copy(g, (void*)dbuf, g->stage[i & 1], r0, std::min(per, g->nrows - r0),
                                  hipMemcpyDeviceToHost);
        return e == hipSuccess ? hipEventRecord(g->stage_ev[i & 1], g->stream) : e;
    };
    if (chunks) PSKEL_HIP(issue(0));
    for (size_t i = 0; i < chunks; ++i) {
        if (i + 1 < chunks) PSKEL_HIP(issue(i + 1));  // into the buffer drained last time round
        PSKEL_HIP(hipEventSynchronize(g->stage_ev[i & 1]));
        const size_t r0 = i * per;
        drain((const unsigned char*)g->stage[i & 1], r0 * rc, std::min(per, g->nrows - r0) * rc);
    }
    return PSKEL_OK;
}

constexpr size_t kConvertGrain = size_t(256) << 10;  // cells per thread, at least

pskel_status_t download_buf(pskel_grid_t* g, const void* dbuf, void* host, int32_t host_dtype)
{
    if (host_dtype == g->storage) {
        const size_t e = g->elem;
        return download_rows(g, dbuf, [&](const unsigned char* sb, size_t c0, size_t n) {
            par_copy((unsigned char*)host + c0 * e, sb, n * e);
        });
    }
    if (g->storage == PSKEL_DTYPE_U8) return fail(PSKEL_ERR_INVALID_ARGUMENT, "u8 grid: host dtype must be U8");
    if (host_dtype != PSKEL_DTYPE_F32 && host_dtype != PSKEL_DTYPE_F64)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "host dtype must be the storage type, F32 or F64");
    const float off = g->store_offset;
    const int storage = g->storage;
    return download_rows(g, dbuf, [&](const unsigned char* sb, size_t c0, size_t n) {
        par_range(n, kConvertGrain, [&](size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) {
                float v;
                if (storage == PSKEL_DTYPE_F32) std::memcpy(&v, sb + i * 4, 4);
                else {
                    uint16_t h;
                    std::memcpy(&h, sb + i * 2, 2);
                    v = (storage == PSKEL_DTYPE_BF16 ? bf16_to_f32(h) : f16_to_f32(h)) + off;
                }
                if (host_dtype == PSKEL_DTYPE_F32) ((float*)host)[c0 + i] = v;
                else ((double*)host)[c0 + i] = (double)v;
            }
        });
    });
}

pskel_status_t ensure_partials(pskel_grid_t* g, size_t n)
{
    if (g->partials_cap >= n) return PSKEL_OK;
    if (g->d_partials) (void)hipFree(g->d_partials);
    g->d_partials = nullptr;
    g->partials_cap = 0;
    PSKEL_HIP(hipMalloc(&g->d_partials, n * sizeof(double)));
    g->partials_cap = n;
    return PSKEL_OK;
}

pskel_status_t ensure_res(pskel_grid_t* g, size_t n)
{
    if (g->res_cap >= n) return PSKEL_OK;
    if (g->d_res) (void)hipFree(g->d_res);
    g->d_res = nullptr;
    g->res_cap = 0;
    PSKEL_HIP(hipMalloc(&g->d_res, std::max<size_t>(n, 1) * sizeof(double)));
    g->res_cap = std::max<size_t>(n, 1);
    return PSKEL_OK;
}

// Generic diff residual ||a - b||_2 (FP64) for any grid: partials per block.
}  // namespace

namespace pskel {
hipError_t launch_diff_partials(const void* a, const void* b, int storage, size_t nrows, size_t rowlen,
                                size_t pitch, double* partials, size_t nblocks, hipStream_t st);
size_t diff_blocks(size_t nrows);
}  // namespace pskel

namespace {

struct RunPlan {
    int path = 0;     // 0 plain, 1 tuned, 2 temporal
    int pattern = PAT_GENERIC;
    int radius = 0;
};

// Sweeps fused in the next launch: up to the temporal depth, cut so that a launch
// ends on every residual step.  The residual of sweep n is ||u^n - u^(n-1)||.
// A 3D fused pair reduces it inside the launch (its second sweep against its
// first); every other launch leaves only u^n and the state before the launch,
// so in 2D a launch of k > 1 sweeps that would end on a residual step stops one
// sweep short and the residual step runs as a single sweep.  A 2D launch of an
// even number of sweeps writes a spare buffer; without one (odd2d) a launch
// fuses one sweep less.
size_t next_k(size_t done, size_t iterations, int kmax, size_t every, int ndim, bool odd2d)
{
    size_t k = std::min<size_t>((size_t)kmax, iterations - done);
    if (every) {
        k = std::min(k, (done / every + 1) * every - done);
        if (ndim == 2 && k > 1 && (done + k) % every == 0) k -= 1;
    }
    if (odd2d && ndim == 2 && k % 2 == 0) k -= 1;
    return k;
}

}  // namespace

// Ping-pong state after a launch of k fused sweeps.  Odd k, and any 3D
// launch, wrote the other buffer: swap.  A 2D launch of even k wrote the spare
// (the other buffer carries the other ring): the spare becomes current and
// the old current becomes the spare.
void pskel::advance_state(pskel::Grid& g, size_t k)
{
    if (k % 2 || g.ndim == 3) {
        std::swap(g.buf[0], g.buf[1]);
        g.parity ^= 1;
    } else {
        std::swap(g.buf[0], g.buf[2]);
        g.spare_parity = g.parity;
    }
}

namespace {

// Issue all launches for `iterations` sweeps; advances the grid state.
pskel_status_t issue(pskel_grid_t* g, const pskel_stencil_desc_t& d, const Points& pts, size_t iterations,
                     const pskel_launch_cfg_t& cfg, const RunPlan& plan, hipStream_t st,
                     int64_t* launches, int64_t* nres)
{
    const int ndim = g->ndim;
    const int kmax = (plan.path == 2) ? std::max(1, cfg.time_tile) : 1;
    const size_t every = cfg.residual_every > 0 ? (size_t)cfg.residual_every : 0;
    size_t done = 0;
    int64_t nl = 0, nr = 0;
    while (done < iterations) {
        const size_t k = next_k(done, iterations, kmax, every, ndim, g->odd_2d_launches);
        const bool want_res = every && ((done + k) % every == 0);
        const void* src = g->buf[0];
        void* dst;
        const bool inplace = (k % 2) == 0 && ndim == 2;  // 3D writes every cell: no ring to keep
        if (!inplace) {
            dst = g->buf[1];
        } else {
            if (!g->buf[2]) return fail(PSKEL_ERR_RUNTIME, "spare buffer missing");
            if (ndim == 2 && g->spare_parity != g->parity) {
                Sweep2D rs = make_sweep2d(g, g->buf[0], g->buf[2], nullptr, plan.radius);
                PSKEL_HIP(launch_copy_ring2d(rs, g->elem, st));
                g->spare_parity = g->parity;
            }
            dst = g->buf[2];
        }
        hipError_t e = hipSuccess;
        bool fused_res = false;
        // a temporal run's leftover single sweeps use the tuned kernel's own
        // default tile (tile_y/tile_z describe the temporal tile there)
        pskel_launch_cfg_t one = cfg;
        if (plan.path == 2 && k == 1) {
            one.tile_y = 0;
            one.tile_z = 0;
            if (ndim == 3) {
                one.tile_x = 0;
                one.block_x = 0;
                one.block_y = 0;
            }
        }
        if (ndim == 2) {
            Sweep2D s = make_sweep2d(g, src, dst, g->buf[1], plan.radius);
            const int bx = cfg.block_x > 0 ? cfg.block_x : 64;
            const int by = cfg.block_y > 0 ? cfg.block_y : 4;
            if (plan.pattern == PAT_CLOUD) {
                CloudArgs ca;
                int fd;
                size_t fw, fh, fz;
                field_info(g->fields[0], &fd, &fw, &fh, &fz, &ca.wx, &ca.fpitch);
                field_info(g->fields[1], &fd, &fw, &fh, &fz, &ca.wy, &ca.fpitch);
                ca.dt = d.coeff;
                if (plan.path >= 1) e = launch_cloud_tuned(s, ca, one, st);
                else e = launch_cloud_plain(s, ca, bx, by, st);
            } else if (g->storage == PSKEL_DTYPE_U8) {
                if (k > 1) e = launch_life_bits(s, (int)k, cfg, st);
                else if (plan.path >= 1) e = launch_life_tuned(s, one, st);
                else e = launch_life_plain(s, pts, bx, by, st);
            } else {
                if (k > 1)
                    e = cfg.tile_x > 0 ? launch_2d_march_f32(s, pts, plan.pattern, (int)k, cfg, st)
                                       : launch_2d_ttile_f32(s, pts, plan.pattern, (int)k, cfg, st);
                else if (plan.path >= 1) e = launch_2d_tuned_f32(s, pts, plan.pattern, one, st);
                else e = launch_2d_plain_f32(s, pts, bx, by, st);
            }
        } else {
            Sweep3D s = make_sweep3d(g, src, dst, plan.radius, want_res ? g->d_partials : nullptr);
            if (g->storage == PSKEL_DTYPE_BF16 || g->storage == PSKEL_DTYPE_F16)
                s.c0 = scaled_sum_offset_term(pts.coeff, pts.np, g->store_offset);
            size_t nb;
            if (k >= 2) {
                nb = ttile3d_blocks(s, g->storage, cfg, (int)k);
                e = launch_3d_ttile(s, pts, g->storage, cfg, (int)k, st);
            } else if (plan.path >= 1) {
                nb = tuned3d_blocks(s, g->storage, one);
                e = launch_3d_tuned(s, pts, g->storage, one, st);
            } else {
                const int bx = cfg.block_x > 0 ? cfg.block_x : 64;
                const int by = cfg.block_y > 0 ? cfg.block_y : 4;
                nb = plain3d_blocks(s, bx, by);
                e = launch_3d_plain(s, pts, g->storage, bx, by, st);
            }
            if (e == hipSuccess && want_res) {
                e = launch_reduce_partials(g->d_partials, nb, g->d_res + nr, st);
                fused_res = true;
            }
        }
        if (e != hipSuccess) return hip_fail(e, "stencil launch");
        nl++;
        advance_state(*g, k);
        g->sweeps_done += k;
        if (want_res) {
            if (!fused_res) {
                const size_t rowlen = g->ndim == 2 ? g->W : g->D;
                const size_t nb = diff_blocks(g->nrows);
                e = launch_diff_partials(g->buf[0], inplace ? g->buf[2] : g->buf[1], g->storage, g->nrows,
                                         rowlen, g->pitch, g->d_partials, nb, st);
                if (e == hipSuccess) e = launch_reduce_partials(g->d_partials, nb, g->d_res + nr, st);
                if (e != hipSuccess) return hip_fail(e, "residual launch");
            }
            nr++;
        }
        done += k;
    }
    *launches = nl;
    *nres = nr;
    return PSKEL_OK;
}

}  // namespace

namespace {

// Validate a run and choose its kernels (shared by the resident runner and the
// host-staged pipeline).
pskel_status_t plan_run(const pskel_grid_t* g, const pskel_stencil_desc_t* d, const pskel_launch_cfg_t& cfg,
                        RunPlan& plan)
{
    const int ndim = g->ndim;
    if (d->npoints < 1 || d->npoints > PSKEL_MAX_POINTS)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "npoints must be in [1, 64]");
    if (d->kind != PSKEL_KIND_SCALED_SUM && d->kind != PSKEL_KIND_WEIGHTED_SUM && d->kind != PSKEL_KIND_LIFE &&
        d->kind != PSKEL_KIND_CLOUDSIM)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown stencil kind");
    if (d->kind == PSKEL_KIND_CLOUDSIM) {
        if (ndim != 2 || g->storage != PSKEL_DTYPE_F32)
            return fail(PSKEL_ERR_UNSUPPORTED, "CLOUDSIM runs on 2D f32 grids");
        for (int k = 0; k < 2; ++k) {
            int fd;
            size_t fw, fh, fz, fp;
            const float* data;
            if (!field_info(g->fields[k], &fd, &fw, &fh, &fz, &data, &fp))
                return fail(PSKEL_ERR_INVALID_ARGUMENT, "CLOUDSIM needs wind fields in slots 0 and 1");
            // the kernel reads field (x = h, y = w) for every interior cell
            if (fd != 2 || fw < g->H || fh < g->W)
                return fail(PSKEL_ERR_INVALID_ARGUMENT, "wind fields must be 2D with width >= grid height "
                                                        "and height >= grid width");
        }
        if (desc_radius(*d, 2) != 1) return fail(PSKEL_ERR_INVALID_ARGUMENT, "CLOUDSIM has mask range 1");
    }
    if ((d->kind == PSKEL_KIND_LIFE) != (g->storage == PSKEL_DTYPE_U8))
        return fail(PSKEL_ERR_UNSUPPORTED, "life runs on u8 grids and u8 grids run life");
    if (ndim == 3 && d->boundary != PSKEL_BOUNDARY_COPY && d->boundary != PSKEL_BOUNDARY_HALO)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown boundary rule");
    const int reach = desc_reach(*d, ndim);
    const int radius = desc_radius(*d, ndim);
    if (radius < reach)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "mask range smaller than the farthest offset");
    if (cfg.variant != PSKEL_VARIANT_PLAIN && cfg.variant != PSKEL_VARIANT_TUNED)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown variant");
    if (cfg.residual_every < 0 || cfg.time_tile < 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "negative cfg field");
    const int bx = cfg.block_x > 0 ? cfg.block_x : 64;
    const int by = cfg.block_y > 0 ? cfg.block_y : 4;
    if (cfg.variant == PSKEL_VARIANT_PLAIN && (bx * by > 1024 || bx > 1024 || by > 1024))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "block has more than 1024 threads");
    if (ndim == 3 && cfg.variant == PSKEL_VARIANT_PLAIN && cfg.residual_every > 0 && (bx * by) % 64 != 0)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "residual needs a block of whole 64-lane waves");

    plan = RunPlan();
    plan.radius = radius;
    plan.pattern = classify(*d, ndim);
    plan.path = 0;
    if (cfg.variant == PSKEL_VARIANT_TUNED) {
        const bool tunable = (ndim == 2 && (plan.pattern == PAT_J5 || plan.pattern == PAT_BOX9 ||
                                            plan.pattern == PAT_LIFE8 || plan.pattern == PAT_CLOUD)) ||
                             (ndim == 3 && plan.pattern == PAT_J7);
        if (tunable) plan.path = 1;
        // field stencils run one sweep per launch (no temporal tiles)
        if (tunable && ndim == 2 && cfg.time_tile > 1 && plan.pattern != PAT_CLOUD) {
            // life runs up to 8 generations per launch on bit-packed cells
            const int kmax = plan.pattern == PAT_LIFE8 ? 8 : 4;
            if (cfg.time_tile > kmax) return fail(PSKEL_ERR_UNSUPPORTED, "temporal depth above 4 (8 for life)");
            // a residual interval shorter than the depth only shortens the
            // launches: none crosses a residual step (next_k).  The bit-packed
            // life tiles keep a ring of one cell: a larger mask range sweeps
            // one generation per launch.
            plan.path = (plan.pattern == PAT_LIFE8 && radius != 1) ? 1 : 2;
            // tile_x > 0 selects the register march (4 cells per lane) over the LDS tile
            if (plan.pattern != PAT_LIFE8 && cfg.tile_x != 0) {
                if (cfg.tile_x != 4 || cfg.tile_y != 0 || cfg.tile_z < 0)
                    return fail(PSKEL_ERR_UNSUPPORTED, "2D register march: tile (4, 0, rows per segment >= 0)");
                if (!march2d_ok((int)g->W, (int)g->H, g->pitch))
                    return fail(PSKEL_ERR_UNSUPPORTED, "2D register march: grids wider than 8 cells, under 2 GiB");
                // its ring handling is that of mask range 1 (rows 0 and H-1, columns 0 and W-1)
                if (radius != 1) return fail(PSKEL_ERR_UNSUPPORTED, "2D register march: mask range 1");
            }
        }
        if (cfg.time_tile > 3 && ndim == 3) return fail(PSKEL_ERR_UNSUPPORTED, "3D temporal depth above 3");
        // two or three fused sweeps per launch run the 7-point Jacobi; for any
        // other 3D stencil the depth is a hint and sweeps run singly
        if (ndim == 3 && cfg.time_tile >= 2 && tunable) {
            // the fused kernel addresses a plane through a buffer of < 2^31 bytes
            if ((uint64_t)g->W * g->pitch * g->elem > 0x7fffffffull)
                return fail(PSKEL_ERR_UNSUPPORTED, "3D temporal tiles need planes under 2 GiB");
            // a run of depth 3 ends with a pair when its length or a residual
            // interval leaves two sweeps: the shape must exist at both depths
            if (!ttile3d_shape_ok(cfg, g->storage, cfg.time_tile) || !ttile3d_shape_ok(cfg, g->storage, 2))
                return fail(PSKEL_ERR_UNSUPPORTED,
                            "3D temporal tile shape (cells per lane, rows per wave, waves along d and w, depth)");
            plan.path = 2;
        }
        if (!tunable) plan.path = 0;  // no specialised kernel for this point set: the plain kernel runs
    }
    if (plan.path == 0) {
        // plain launches keep the caller's block; it must have whole waves when the
        // residual reduction uses wave shuffles
        if (bx <= 0 || by <= 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "block dims must be positive");
    }
    return PSKEL_OK;
}

}  // namespace

namespace {

// FP64 partials a run with residuals needs: one per block of the largest
// launch its sweeps use (3D sweeps reduce inside the stencil launch), or of
// the separate difference kernel (2D).
size_t residual_partials(const pskel_grid_t* g, const RunPlan& plan, const pskel_launch_cfg_t& cfg)
{
    size_t nb = diff_blocks(g->nrows);
    if (g->ndim == 3) {
        const int bx = cfg.block_x > 0 ? cfg.block_x : 64;
        const int by = cfg.block_y > 0 ? cfg.block_y : 4;
        Sweep3D s3 = make_sweep3d(g, nullptr, nullptr, plan.radius, nullptr);
        pskel_launch_cfg_t one = cfg;
        if (plan.path == 2) one.tile_x = one.tile_y = one.tile_z = one.block_x = one.block_y = 0;
        nb = std::max(nb, plan.path >= 1 ? tuned3d_blocks(s3, g->storage, one) : plain3d_blocks(s3, bx, by));
        if (plan.path == 2)
            for (int k = 2; k <= std::max(2, cfg.time_tile); ++k)
                nb = std::max(nb, ttile3d_blocks(s3, g->storage, cfg, k));
    }
    return nb;
}

}  // namespace

// Issue `iterations` sweeps on `stream` and return without waiting for them:
// the host-staged pipeline overlaps them with copies on other streams.  No
// residuals and no graph; the grid's buffers must already exist (the spare
// too, for 2D temporal tiles of even depth).
pskel_status_t pskel::run_async(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t iterations,
                                const pskel_launch_cfg_t& cfg, hipStream_t stream, int64_t* launches)
{
    *launches = 0;
    if (cfg.residual_every > 0 || cfg.use_graph) return fail(PSKEL_ERR_UNSUPPORTED, "residuals or graphs here");
    RunPlan plan;
    pskel_status_t st = plan_run(g, d, cfg, plan);
    if (st != PSKEL_OK) return st;
    if (iterations == 0) return PSKEL_OK;
    if (plan.path == 2 && g->ndim == 2 && !g->buf[2]) return fail(PSKEL_ERR_RUNTIME, "spare buffer missing");
    Points pts;
    fill_points(pts, *d, g->ndim);
    int64_t nr = 0;
    return issue(g, *d, pts, iterations, cfg, plan, stream, launches, &nr);
}

extern "C" {

int32_t pskel_abi_version(void) { return PSKEL_AMD_ABI_VERSION; }

size_t pskel_abi_struct_size(int32_t which)
{
    switch (which) {
    case 0: return sizeof(pskel_stencil_desc_t);
    case 1: return sizeof(pskel_launch_cfg_t);
    case 2: return sizeof(pskel_run_stats_t);
    case 3: return sizeof(pskel_grid_desc_t);
    case 4: return sizeof(pskel_grid_info_t);
    case 5: return sizeof(pskel_field_desc_t);
    default: return 0;
    }
}

const char* pskel_status_string(pskel_status_t st)
{
    switch (st) {
    case PSKEL_OK: return "ok";
    case PSKEL_ERR_INVALID_ARGUMENT: return "invalid argument";
    case PSKEL_ERR_OUT_OF_MEMORY: return "out of device memory";
    case PSKEL_ERR_NO_DEVICE: return "no HIP device";
    case PSKEL_ERR_LAUNCH: return "kernel launch failed";
    case PSKEL_ERR_UNSUPPORTED: return "unsupported configuration";
    case PSKEL_ERR_RUNTIME: return "HIP runtime error";
    }
    return "unknown status";
}

const char* pskel_last_error(void) { return g_error.c_str(); }

pskel_status_t pskel_device_count(int32_t* count)
{
    if (!count) return fail(PSKEL_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return PSKEL_OK;
}

pskel_status_t pskel_device_synchronize(void)
{
    PSKEL_HIP(hipDeviceSynchronize());
    return PSKEL_OK;
}

pskel_status_t pskel_device_info(int32_t device, char* name, size_t name_cap, size_t* mem_free,
                                 size_t* mem_total, int32_t* num_cus)
{
    PSKEL_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PSKEL_HIP(hipGetDeviceProperties(&prop, device));
    if (name && name_cap) {
        std::snprintf(name, name_cap, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    size_t f = 0, t = 0;
    PSKEL_HIP(hipMemGetInfo(&f, &t));
    if (mem_free) *mem_free = f;
    if (mem_total) *mem_total = t;
    return PSKEL_OK;
}

pskel_status_t pskel_grid_create(const pskel_grid_desc_t* desc, pskel_grid_t** out)
{
    if (!desc || !out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (desc->ndim != 2 && desc->ndim != 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    const int st = desc->storage;
    if (st != PSKEL_DTYPE_U8 && st != PSKEL_DTYPE_F32 && st != PSKEL_DTYPE_BF16 && st != PSKEL_DTYPE_F16)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "storage must be U8, F32, BF16 or F16");
    if (desc->ndim == 2 && (st == PSKEL_DTYPE_BF16 || st == PSKEL_DTYPE_F16))
        return fail(PSKEL_ERR_UNSUPPORTED, "16-bit storage is implemented for 3D grids");
    if (desc->ndim == 3 && st == PSKEL_DTYPE_U8)
        return fail(PSKEL_ERR_UNSUPPORTED, "u8 storage is implemented for 2D grids");
    const size_t W = desc->width, H = desc->height, D = desc->ndim == 2 ? 1 : desc->depth;
    if (W == 0 || H == 0 || D == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "zero-sized grid");
    if (W > (1u << 30) || H > (1u << 30) || D > (1u << 30))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "dimension too large");
    hipError_t e = hipSetDevice(desc->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");

    pskel_grid_t* g = new (std::nothrow) pskel_grid_t();
    if (!g) return fail(PSKEL_ERR_OUT_OF_MEMORY, "host allocation");
    g->ndim = desc->ndim;
    g->storage = st;
    g->elem = elem_size(st);
    g->W = W;
    g->H = H;
    g->D = D;
    g->device = desc->device;
    // NaN: the offset of 16-bit storage follows the data (the mean of each INPUT upload)
    g->auto_offset = std::isnan(desc->store_offset) && (st == PSKEL_DTYPE_BF16 || st == PSKEL_DTYPE_F16);
    g->store_offset = std::isnan(desc->store_offset) ? 0.0f : desc->store_offset;
    if (desc->rounding != PSKEL_ROUND_NEAREST && desc->rounding != PSKEL_ROUND_STOCHASTIC) {
        delete g;
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown rounding mode");
    }
    // the random word of a cell hashes its 32-bit in-plane index w*D + d
    if (desc->rounding == PSKEL_ROUND_STOCHASTIC && (uint64_t)desc->width * desc->depth >= (1ull << 32)) {
        delete g;
        return fail(PSKEL_ERR_UNSUPPORTED, "stochastic rounding needs width * depth < 2^32");
    }
    g->rounding = desc->rounding;
    g->seed = desc->seed;
    // rows are padded to whole 1 KiB segments: the widest tuned tile, so vector
    // loads of a tile never leave the allocation and every row is aligned
    const size_t row = desc->ndim == 2 ? W : D;
    g->pitch = round_up(row * g->elem, 1024) / g->elem;
    g->nrows = desc->ndim == 2 ? H : H * W;
    g->bytes = g->pitch * g->elem * g->nrows;

    auto cleanup = [&](pskel_status_t s) {
        pskel_grid_destroy(g);
        return s;
    };
    if ((e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup(hip_fail(e, "hipStreamCreate"));
    g->own_stream = true;
    if ((e = hipEventCreate(&g->ev0)) != hipSuccess) return cleanup(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreate(&g->ev1)) != hipSuccess) return cleanup(hip_fail(e, "hipEventCreate"));
    for (int i = 0; i < 2; ++i) {
        if ((e = hipMalloc(&g->buf[i], g->bytes)) != hipSuccess) return cleanup(hip_fail(e, "hipMalloc grid buffer"));
        if ((e = launch_fill_zero(g->buf[i], g->bytes, g->stream)) != hipSuccess)
            return cleanup(hip_fail(e, "zero fill"));
    }
    if ((e = hipStreamSynchronize(g->stream)) != hipSuccess) return cleanup(hip_fail(e, "hipStreamSynchronize"));
    *out = g;
    return PSKEL_OK;
}

pskel_status_t pskel_grid_destroy(pskel_grid_t* g)
{
    if (!g) return PSKEL_OK;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (auto& ge : g->graphs)
        if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
    for (int i = 0; i < 3; ++i)
        if (g->buf[i]) (void)hipFree(g->buf[i]);
    for (int b = 0; b < 2; ++b) {
        if (g->stage[b]) (void)hipHostFree(g->stage[b]);
        if (g->stage_ev[b]) (void)hipEventDestroy(g->stage_ev[b]);
    }
    if (g->d_partials) (void)hipFree(g->d_partials);
    if (g->d_res) (void)hipFree(g->d_res);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return PSKEL_OK;
}

pskel_status_t pskel_grid_upload(pskel_grid_t* g, int32_t which, const void* host, int32_t host_dtype)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (!host) return fail(PSKEL_ERR_INVALID_ARGUMENT, "host pointer is NULL");
    if (which != PSKEL_BUFFER_INPUT && which != PSKEL_BUFFER_OUTPUT)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "which must be INPUT or OUTPUT");
    PSKEL_HIP(hipSetDevice(g->device));
    // uploading resets the ping-pong: INPUT is the current state, OUTPUT the other
    g->parity = 0;
    if (which == PSKEL_BUFFER_INPUT) g->sweeps_done = 0;
    g->spare_parity = -1;
    void* dbuf = which == PSKEL_BUFFER_INPUT ? g->buf[0] : g->buf[1];
    const size_t n = cells(g);
    if (g->storage == PSKEL_DTYPE_U8) {
        // u8 grids run life, whose byte-parallel kernels count neighbours of 0/1 cells
        if (host_dtype != PSKEL_DTYPE_U8) return fail(PSKEL_ERR_INVALID_ARGUMENT, "u8 grid: host dtype must be U8");
        const uint8_t* c = (const uint8_t*)host;
        for (size_t i = 0; i < n; ++i)
            if (c[i] > 1) return fail(PSKEL_ERR_INVALID_ARGUMENT, "life cells must be 0 or 1");
    }
    if (host_dtype == g->storage) {
        if (g->auto_offset && which == PSKEL_BUFFER_INPUT)
            return fail(PSKEL_ERR_INVALID_ARGUMENT, "an automatic storage offset is the mean of F32 or F64 data");
        const size_t e = g->elem;
        return upload_rows(g, dbuf, [&](unsigned char* sb, size_t c0, size_t m) {
            par_copy(sb, (const unsigned char*)host + c0 * e, m * e);
        });
    }
    if (host_dtype != PSKEL_DTYPE_F32 && host_dtype != PSKEL_DTYPE_F64)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "host dtype must be the storage type, F32 or F64");
    if (g->auto_offset && which == PSKEL_BUFFER_INPUT) {
        // the mean of the finite values, in double, rounded once
        double sum = 0.0;
        size_t cnt = 0;
        for (size_t i = 0; i < n; ++i) {
            const double v = host_dtype == PSKEL_DTYPE_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i];
            if (std::isfinite(v)) {
                sum += v;
                ++cnt;
            }
        }
        g->store_offset = cnt ? (float)(sum / (double)cnt) : 0.0f;
    }
    const float off = g->store_offset;
    const int storage = g->storage;
    return upload_rows(g, dbuf, [&](unsigned char* sb, size_t c0, size_t m) {
        par_range(m, kConvertGrain, [&](size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) {
                const size_t c = c0 + i;
                const float v = host_dtype == PSKEL_DTYPE_F32 ? ((const float*)host)[c] : (float)((const double*)host)[c];
                if (storage == PSKEL_DTYPE_F32) std::memcpy(sb + i * 4, &v, 4);
                else {
                    const uint16_t h = storage == PSKEL_DTYPE_BF16 ? f32_to_bf16(v - off) : f32_to_f16(v - off);
                    std::memcpy(sb + i * 2, &h, 2);
                }
            }
        });
    });
}

pskel_status_t pskel_grid_download(pskel_grid_t* g, void* host, int32_t host_dtype)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (!host) return fail(PSKEL_ERR_INVALID_ARGUMENT, "host pointer is NULL");
    PSKEL_HIP(hipSetDevice(g->device));
    return download_buf(g, g->buf[0], host, host_dtype);
}

pskel_status_t pskel_grid_download_previous(pskel_grid_t* g, void* host, int32_t host_dtype)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (!host) return fail(PSKEL_ERR_INVALID_ARGUMENT, "host pointer is NULL");
    PSKEL_HIP(hipSetDevice(g->device));
    return download_buf(g, g->buf[1], host, host_dtype);
}

pskel_status_t pskel_grid_info(const pskel_grid_t* g, pskel_grid_info_t* info)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (!info) return fail(PSKEL_ERR_INVALID_ARGUMENT, "info is NULL");
    info->ndim = g->ndim;
    info->storage = g->storage;
    info->elem_bytes = g->elem;
    info->parity = g->parity;
    info->width = g->W;
    info->height = g->H;
    info->depth = g->D;
    info->pitch = g->pitch;
    info->bytes_per_buffer = g->bytes;
    info->store_offset = g->store_offset;
    info->current = g->buf[0];
    info->previous = g->buf[1];
    info->stream = g->stream;
    return PSKEL_OK;
}

pskel_status_t pskel_run_iterative(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t iterations,
                                   const pskel_launch_cfg_t* cfg_in, pskel_run_stats_t* stats)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (!d) return fail(PSKEL_ERR_INVALID_ARGUMENT, "descriptor is NULL");
    pskel_launch_cfg_t cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    if (cfg_in) cfg = *cfg_in;
    if (stats) {
        stats->elapsed_ms = 0.0;
        stats->launches = 0;
        stats->sweeps = 0;
        stats->residual_count = 0;
    }
    // ---- validation -------------------------------------------------------
    RunPlan plan;
    if ((st = plan_run(g, d, cfg, plan)) != PSKEL_OK) return st;
    const int ndim = g->ndim;
    const int radius = plan.radius;
    if (stats) stats->sweeps = (int64_t)iterations;
    if (iterations == 0) return PSKEL_OK;

    PSKEL_HIP(hipSetDevice(g->device));
    hipStream_t stream = cfg.stream ? (hipStream_t)cfg.stream : g->stream;
    Points pts;
    fill_points(pts, *d, ndim);

    // ---- resources ----------------------------------------------------------
    const size_t every = cfg.residual_every > 0 ? (size_t)cfg.residual_every : 0;
    const size_t nres_max = every ? iterations / every : 0;
    if (every) {
        if ((st = ensure_partials(g, residual_partials(g, plan, cfg))) != PSKEL_OK) return st;
        if ((st = ensure_res(g, nres_max)) != PSKEL_OK) return st;
    }
    if (plan.path == 2 && ndim == 3) {
        // the fused kernel sizes its grid from device queries: make them here,
        // outside any stream capture below
        Sweep3D s3 = make_sweep3d(g, nullptr, nullptr, radius, nullptr);
        for (int k = 2; k <= std::max(2, cfg.time_tile); ++k) (void)ttile3d_blocks(s3, g->storage, cfg, k);
    }
    if (plan.path == 2 && ndim == 2 && !g->buf[2]) {
        PSKEL_HIP(hipMalloc(&g->buf[2], g->bytes));
        PSKEL_HIP(launch_copy(g->buf[2], g->buf[0], g->bytes, stream));
        g->spare_parity = g->parity;
    }

    int64_t nl = 0, nr = 0;
    if (!cfg.use_graph) {
        PSKEL_HIP(hipEventRecord(g->ev0, stream));
        st = issue(g, *d, pts, iterations, cfg, plan, stream, &nl, &nr);
        if (st != PSKEL_OK) return st;
        PSKEL_HIP(hipEventRecord(g->ev1, stream));
    } else {
        // graph cache key: everything that shapes the launch sequence
        pskel_launch_cfg_t kc = cfg;
        kc.stream = nullptr;
        const int sp = g->spare_parity * 2 + g->parity;
        const void* ptrs[5] = {g->buf[0], g->buf[1], g->buf[2], g->d_res, g->d_partials};
        // stochastic rounding bakes the sweep indices into the launches
        const uint64_t sweep0 = g->rounding == PSKEL_ROUND_STOCHASTIC ? g->sweeps_done : 0;
        std::vector<unsigned char> key;
        auto put = [&key](const void* p, size_t n) {
            key.insert(key.end(), (const unsigned char*)p, (const unsigned char*)p + n);
        };
        put(d, sizeof(*d));
        put(&kc, sizeof(kc));
        put(&iterations, sizeof(iterations));
        put(ptrs, sizeof(ptrs));
        put(&sp, sizeof(sp));
        put(&sweep0, sizeof(sweep0));
        put(g->fields, sizeof(g->fields));
        GraphEntry* hit = nullptr;
        for (auto& ge : g->graphs)
            if (ge.key == key) hit = &ge;
        if (!hit) {
            // capture on a private stream so the caller's stream is never in capture mode
            pskel::Grid saved = *g;
            hipStream_t cap;
            PSKEL_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                (void)hipStreamDestroy(cap);
                return hip_fail(e, "hipStreamBeginCapture");
            }
            st = issue(g, *d, pts, iterations, cfg, plan, cap, &nl, &nr);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(cap, &graph);
            (void)hipStreamDestroy(cap);
            // the capture advanced the state; restore it, the replay advances it for real
            static_cast<pskel::Grid&>(*g) = saved;
            if (st != PSKEL_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            GraphEntry ge;
            ge.key = key;
            ge.launches = nl;
            ge.residuals = nr;
            e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            if (g->graphs.size() >= 8) {
                (void)hipGraphExecDestroy(g->graphs.front().exec);
                g->graphs.erase(g->graphs.begin());
            }
            g->graphs.push_back(std::move(ge));
            hit = &g->graphs.back();
        }
        PSKEL_HIP(hipEventRecord(g->ev0, stream));
        PSKEL_HIP(hipGraphLaunch(hit->exec, stream));
        PSKEL_HIP(hipEventRecord(g->ev1, stream));
        nl = hit->launches;
        nr = hit->residuals;
        // advance the host-side state exactly as the captured sequence did
        const int kmax = (plan.path == 2) ? std::max(1, cfg.time_tile) : 1;
        for (size_t done = 0; done < iterations;) {
            const size_t k = next_k(done, iterations, kmax, every, g->ndim, g->odd_2d_launches);
            if (k % 2 == 0 && g->ndim == 2) g->spare_parity = g->parity;  // ring copy in the graph
            advance_state(*g, k);
            g->sweeps_done += k;
            done += k;
        }
    }
    PSKEL_HIP(hipEventSynchronize(g->ev1));
    float ms = 0.f;
    PSKEL_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    if (stats) {
        stats->elapsed_ms = ms;
        stats->launches = nl;
        stats->residual_count = nr;
        if (nr > 0 && stats->residuals && stats->residual_capacity > 0) {
            const size_t ncopy = (size_t)std::min<int64_t>(nr, stats->residual_capacity);
            PSKEL_HIP(hipMemcpy(stats->residuals, g->d_res, ncopy * sizeof(double), hipMemcpyDeviceToHost));
        }
    }
    return PSKEL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free runs (pskel_run_iterative_noalloc).  The FP64 residual
// reductions need scratch: one partial per block of a launch, then the
// reduced residuals of the call.  It lives in a static device array, whose
// copy on each GPU stays allocated until the process ends.  Host code reaches
// a GPU's copy only through the address hipGetSymbolAddress returns for it
// (resolved once per device, with that device current); the array's name is
// not a device address on the host.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchPartials = 65536;   // partials of one launch (blocks)
constexpr size_t kScratchResiduals = 4096;   // residuals of one call
constexpr int kScratchDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kScratchPartials + kScratchResiduals];

namespace {
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kScratchDevices] = {};

pskel_status_t scratch_of(int device, double** out)
{
    if (device < 0 || device >= kScratchDevices) return fail(PSKEL_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        PSKEL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PSKEL_ERR_RUNTIME, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    return PSKEL_OK;
}
}  // namespace

extern "C" pskel_status_t pskel_run_iterative_noalloc(const pskel_grid_desc_t* desc, void* in, void* out,
                                                      size_t pitch, const pskel_stencil_desc_t* d,
                                                      size_t iterations, const pskel_launch_cfg_t* cfg_in,
                                                      pskel_run_stats_t* stats)
{
    if (!desc || !d || !in || !out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (in == out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "in and out must be distinct buffers");
    if (stats) {
        stats->elapsed_ms = 0.0;
        stats->launches = 0;
        stats->sweeps = 0;
        stats->residual_count = 0;
    }
    const int stg = desc->storage;
    if (desc->ndim != 2 && desc->ndim != 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    if (stg != PSKEL_DTYPE_U8 && stg != PSKEL_DTYPE_F32 && stg != PSKEL_DTYPE_BF16 && stg != PSKEL_DTYPE_F16)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "storage must be U8, F32, BF16 or F16");
    if (desc->ndim == 2 && (stg == PSKEL_DTYPE_BF16 || stg == PSKEL_DTYPE_F16))
        return fail(PSKEL_ERR_UNSUPPORTED, "16-bit storage is implemented for 3D grids");
    if (desc->ndim == 3 && stg == PSKEL_DTYPE_U8) return fail(PSKEL_ERR_UNSUPPORTED, "u8 storage is implemented for 2D grids");
    const size_t W = desc->width, H = desc->height, D = desc->ndim == 2 ? 1 : desc->depth;
    if (W == 0 || H == 0 || D == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "zero-sized grid");
    if (W > (1u << 30) || H > (1u << 30) || D > (1u << 30)) return fail(PSKEL_ERR_INVALID_ARGUMENT, "dimension too large");
    if (desc->rounding != PSKEL_ROUND_NEAREST && desc->rounding != PSKEL_ROUND_STOCHASTIC)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown rounding mode");
    if (std::isnan(desc->store_offset))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "the buffers hold storage words: give their storage offset");
    if (desc->rounding == PSKEL_ROUND_STOCHASTIC && (uint64_t)W * D >= (1ull << 32))
        return fail(PSKEL_ERR_UNSUPPORTED, "stochastic rounding needs width * depth < 2^32");
    const int elem = elem_size(stg);
    const size_t row = desc->ndim == 2 ? W : D;
    if (pitch < row || (pitch * elem) % 1024 != 0)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "pitch must cover a row and be a whole number of 1 KiB");
    if ((uintptr_t)in % 256 != 0 || (uintptr_t)out % 256 != 0)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "buffers must be 256-byte aligned");
    pskel_launch_cfg_t cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    if (cfg_in) cfg = *cfg_in;
    if (cfg.use_graph) return fail(PSKEL_ERR_UNSUPPORTED, "no graphs without allocation (instantiating one allocates)");
    if (d->kind == PSKEL_KIND_CLOUDSIM) return fail(PSKEL_ERR_UNSUPPORTED, "field arguments are attached to a grid");

    // a grid over the caller's buffers: no device memory of its own, no spare
    pskel_grid_t g;
    g.ndim = desc->ndim;
    g.storage = stg;
    g.elem = elem;
    g.W = W;
    g.H = H;
    g.D = D;
    g.pitch = pitch;
    g.nrows = desc->ndim == 2 ? H : H * W;
    g.bytes = pitch * elem * g.nrows;
    g.store_offset = desc->store_offset;
    g.device = desc->device;
    g.rounding = desc->rounding;
    g.seed = desc->seed;
    g.buf[0] = in;
    g.buf[1] = out;
    g.odd_2d_launches = true;
    RunPlan plan;
    pskel_status_t st = plan_run(&g, d, cfg, plan);
    if (st != PSKEL_OK) return st;
    const size_t every = cfg.residual_every > 0 ? (size_t)cfg.residual_every : 0;
    const size_t nres = every ? iterations / every : 0;
    if (nres > kScratchResiduals)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "at most 4096 residuals per call (iterations / residual_every)");
    PSKEL_HIP(hipSetDevice(desc->device));
    double* scr = nullptr;
    if ((st = scratch_of(desc->device, &scr)) != PSKEL_OK) return st;
    if (every) {
        const size_t nb = residual_partials(&g, plan, cfg);
        if (nb > kScratchPartials)
            return fail(PSKEL_ERR_INVALID_ARGUMENT, "residuals of launches of more than 65536 blocks need a grid");
        g.d_partials = scr;
        g.partials_cap = kScratchPartials;
        g.d_res = scr + kScratchPartials;
        g.res_cap = kScratchResiduals;
    }
    hipStream_t stream = (hipStream_t)cfg.stream;  // NULL: the null stream
    Points pts;
    fill_points(pts, *d, g.ndim);
    int64_t nl = 0, nr = 0;
    if (iterations > 0 && (st = issue(&g, *d, pts, iterations, cfg, plan, stream, &nl, &nr)) != PSKEL_OK) return st;
    // the result goes to `out`, as runIterativeGPU leaves it in output
    if (g.buf[0] != out) PSKEL_HIP(launch_copy(out, g.buf[0], g.bytes, stream));
    if (stats) {
        stats->launches = nl;
        stats->sweeps = (int64_t)iterations;
        stats->residual_count = nr;
        if (nr > 0 && stats->residuals && stats->residual_capacity > 0) {
            const size_t ncopy = (size_t)std::min<int64_t>(nr, stats->residual_capacity);
            PSKEL_HIP(hipMemcpyAsync(stats->residuals, g.d_res, ncopy * sizeof(double), hipMemcpyDeviceToHost, stream));
            PSKEL_HIP(hipStreamSynchronize(stream));
        }
    }
    return PSKEL_OK;
}

extern "C" {

pskel_status_t pskel_grid_swap(pskel_grid_t* g)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    std::swap(g->buf[0], g->buf[1]);
    g->parity ^= 1;
    g->sweeps_done += 1;
    return PSKEL_OK;
}

pskel_status_t pskel_grid_set_field(pskel_grid_t* g, int32_t slot, pskel_field_t* f)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (slot < 0 || slot >= 4) return fail(PSKEL_ERR_INVALID_ARGUMENT, "field slot must be 0..3");
    g->fields[slot] = f;
    return PSKEL_OK;
}

pskel_status_t pskel_grid_residual(pskel_grid_t* g, double* out)
{
    pskel_status_t st = check_grid(g);
    if (st != PSKEL_OK) return st;
    if (!out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "out is NULL");
    PSKEL_HIP(hipSetDevice(g->device));
    const size_t nb = diff_blocks(g->nrows);
    if ((st = ensure_partials(g, nb)) != PSKEL_OK) return st;
    if ((st = ensure_res(g, 1)) != PSKEL_OK) return st;
    const size_t rowlen = g->ndim == 2 ? g->W : g->D;
    hipError_t e = launch_diff_partials(g->buf[0], g->buf[1], g->storage, g->nrows, rowlen, g->pitch,
                                        g->d_partials, nb, g->stream);
    if (e == hipSuccess) e = launch_reduce_partials(g->d_partials, nb, g->d_res, g->stream);
    if (e != hipSuccess) return hip_fail(e, "residual launch");
    PSKEL_HIP(hipMemcpyAsync(out, g->d_res, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    PSKEL_HIP(hipStreamSynchronize(g->stream));
    return PSKEL_OK;
}

pskel_status_t pskel_device_copy(void* dst, const void* src, size_t bytes, void* stream, double* ms)
{
    if ((!dst || !src) && bytes) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (((uintptr_t)dst | (uintptr_t)src) % 16 != 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "pointers must be 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t a, b;
    PSKEL_HIP(hipEventCreate(&a));
    PSKEL_HIP(hipEventCreate(&b));
    PSKEL_HIP(hipEventRecord(a, st));
    hipError_t e = launch_copy(dst, src, bytes, st);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return hip_fail(e, "copy launch");
    }
    PSKEL_HIP(hipEventRecord(b, st));
    PSKEL_HIP(hipEventSynchronize(b));
    float t = 0.f;
    PSKEL_HIP(hipEventElapsedTime(&t, a, b));
    if (ms) *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return PSKEL_OK;
}

pskel_status_t pskel_device_alloc(size_t bytes, void** out)
{
    if (!out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    PSKEL_HIP(hipMalloc(out, bytes));
    return PSKEL_OK;
}

pskel_status_t pskel_device_free(void* ptr)
{
    if (ptr) PSKEL_HIP(hipFree(ptr));
    return PSKEL_OK;
}

pskel_status_t pskel_memcpy(void* dst, const void* src, size_t bytes, int32_t to_device)
{
    PSKEL_HIP(hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    return PSKEL_OK;
}

}  // extern "C"
