// Host-staged (out-of-core) sweeps: the grid stays in host memory and slabs of
// it visit the device.
//
// Reference: StencilBase::runIterativeTilingGPU (include/PSkelStencil.hpp:
// 1663-1760) and the tiled branch of runIterativeAutoGPU (:2014-2074): the
// iterations are cut into passes of `inner` sweeps; every pass stages each
// tile plus a halo of range*inner cells (StencilTiling::tile,
// include/PSkelStencilTiling.h:85-148) on the device, runs the sweeps there and
// copies the tile's core back.  Here a tile is a slab of whole rows (2D) or
// whole h-planes (3D), which keeps every host copy contiguous.
//
// Each slab becomes a small resident grid whose two buffers are staged from
// the two host arrays of the global ping-pong (so each carries its own ring,
// as the resident runner's buffers do).  A slab's edge rows/planes are not
// updated correctly on the device (2D sweeps never write them; 3D sweeps
// treat them as faces), and that error travels range cells per sweep: after
// `inner` sweeps it has not reached the core, which is exact.  Cores of a pass
// are collected in a third array, which then becomes the host array of the
// state's parity, so later slabs of the same pass still read the pass's
// starting state.  The result equals `iterations` resident sweeps bit for bit,
// stochastic rounding included (the slab carries its global plane origin and
// the global sweep index).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>

#include "pskel_internal.hpp"

using namespace pskel;

namespace {

int storage_bytes(int dtype)
{
    switch (dtype) {
    case PSKEL_DTYPE_U8: return 1;
    case PSKEL_DTYPE_F32: return 4;
    case PSKEL_DTYPE_BF16:
    case PSKEL_DTYPE_F16: return 2;
    default: return 0;
    }
}

struct HostBuf {
    unsigned char* p = nullptr;
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

// rows [row0, row0 + n) of a grid's current buffer -> host (contiguous rows)
pskel_status_t download_rows(pskel_grid_t* g, void* host, size_t row0, size_t n)
{
    const size_t row_bytes = (g->ndim == 2 ? g->W : g->D) * g->elem;
    const unsigned char* src = (const unsigned char*)g->buf[0] + row0 * g->pitch * g->elem;
    PSKEL_HIP(hipMemcpy2DAsync(host, row_bytes, src, g->pitch * g->elem, row_bytes, n, hipMemcpyDeviceToHost,
                               g->stream));
    PSKEL_HIP(hipStreamSynchronize(g->stream));
    return PSKEL_OK;
}

}  // namespace

extern "C" pskel_status_t pskel_run_iterative_host(const pskel_grid_desc_t* gd, const void* host_in, void* host_out,
                                                   int32_t host_dtype, const pskel_stencil_desc_t* d,
                                                   size_t iterations, size_t tile, size_t inner,
                                                   const pskel_launch_cfg_t* cfg_in, pskel_run_stats_t* stats)
{
    if (!gd || !host_in || !host_out || !d) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (gd->ndim != 2 && gd->ndim != 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    const int elem = storage_bytes(gd->storage);
    if (elem == 0) return fail(PSKEL_ERR_UNSUPPORTED, "storage type");
    if (host_dtype != gd->storage)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "host-staged sweeps take host arrays in the storage type");
    if (tile == 0 || inner == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "tile and inner must be positive");
    pskel_launch_cfg_t cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    if (cfg_in) cfg = *cfg_in;
    if (cfg.residual_every > 0) return fail(PSKEL_ERR_UNSUPPORTED, "residuals of host-staged sweeps");
    if (d->kind == PSKEL_KIND_CLOUDSIM) return fail(PSKEL_ERR_UNSUPPORTED, "field stencils stay resident");
    if (stats) {
        stats->elapsed_ms = 0.0;
        stats->launches = 0;
        stats->sweeps = (int64_t)iterations;
        stats->residual_count = 0;
    }
    const size_t H = gd->height;
    const size_t plane_elems = gd->width * (gd->ndim == 3 ? gd->depth : 1);
    const size_t plane_bytes = plane_elems * (size_t)elem;
    const size_t total = H * plane_bytes;
    if (H == 0 || plane_elems == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "empty grid");
    // A slab's wrong edge zone is its ring (range r, never updated on the
    // device) and then grows by the h-reach of the mask per sweep, so a halo of
    // r * inner (r >= the h-reach) keeps the core exact.
    if (d->npoints < 1 || d->npoints > PSKEL_MAX_POINTS) return fail(PSKEL_ERR_INVALID_ARGUMENT, "npoints");
    const size_t radius = (size_t)desc_radius(*d, gd->ndim);

    const auto t0 = std::chrono::steady_clock::now();
    PSKEL_HIP(hipSetDevice(gd->device));
    // the two arrays of the global ping-pong plus the cores of the current pass
    HostBuf a, b, sc;
    PSKEL_HIP(hipHostMalloc((void**)&a.p, total, hipHostMallocDefault));
    PSKEL_HIP(hipHostMalloc((void**)&b.p, total, hipHostMallocDefault));
    PSKEL_HIP(hipHostMalloc((void**)&sc.p, total, hipHostMallocDefault));
    std::memcpy(a.p, host_in, total);
    std::memcpy(b.p, host_out, total);
    unsigned char* buf[2] = {a.p, b.p};
    unsigned char* scratch = sc.p;

    std::map<size_t, pskel_grid_t*> grids;  // one device grid per slab height
    auto cleanup = [&]() {
        for (auto& kv : grids) pskel_grid_destroy(kv.second);
        grids.clear();
    };
    int64_t launches = 0;
    size_t done = 0;
    while (done < iterations) {
        const size_t sub = std::min(inner, iterations - done);
        const int srcp = (int)(done % 2);  // host array holding the state
        const size_t halo = radius * sub;
        for (size_t h0 = 0; h0 < H; h0 += tile) {
            const size_t h1 = std::min(H, h0 + tile);
            const size_t s0 = h0 > halo ? h0 - halo : 0;
            const size_t s1 = std::min(H, h1 + halo);
            const size_t rows = s1 - s0;
            pskel_grid_t* g;
            auto it = grids.find(rows);
            if (it != grids.end()) {
                g = it->second;
            } else {
                pskel_grid_desc_t sd = *gd;
                sd.height = rows;
                pskel_status_t st = pskel_grid_create(&sd, &g);
                if (st != PSKEL_OK) {
                    cleanup();
                    return st;
                }
                grids[rows] = g;
            }
            pskel_status_t st = pskel_grid_upload(g, PSKEL_BUFFER_INPUT, buf[srcp] + s0 * plane_bytes, host_dtype);
            // 3D sweeps write every cell (faces by copy), so the other buffer's
            // contents are never read; 2D sweeps keep its ring
            if (st == PSKEL_OK && gd->ndim == 2)
                st = pskel_grid_upload(g, PSKEL_BUFFER_OUTPUT, buf[srcp ^ 1] + s0 * plane_bytes, host_dtype);
            if (st != PSKEL_OK) {
                cleanup();
                return st;
            }
            g->sweeps_done = done;           // global sweep index (stochastic rounding)
            g->plane_origin = (uint32_t)s0;  // global plane index (stochastic rounding)
            pskel_run_stats_t rs{};
            st = pskel_run_iterative(g, d, sub, &cfg, &rs);
            if (st == PSKEL_OK) {
                const size_t rows_per_plane = g->ndim == 2 ? 1 : g->W;
                st = download_rows(g, scratch + h0 * plane_bytes, (h0 - s0) * rows_per_plane,
                                   (h1 - h0) * rows_per_plane);
            }
            if (st != PSKEL_OK) {
                cleanup();
                return st;
            }
            launches += rs.launches;
        }
        // the cores are the state after the pass, in the array of its parity
        const int dstp = (int)((done + sub) % 2);
        std::swap(buf[dstp], scratch);
        done += sub;
    }
    cleanup();
    // runIterativeCPU's result placement: N even leaves the state in the input
    // array and copies it to the output (include/PSkelStencil.hpp:932)
    std::memcpy(host_out, buf[iterations % 2], total);
    if (stats) {
        stats->elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->launches = launches;
    }
    return PSKEL_OK;
}
