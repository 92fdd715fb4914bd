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
// Exactness.  A slab's edge rows/planes are not updated correctly on the
// device (2D sweeps never write them; 3D sweeps treat them as faces), and that
// error travels range cells per sweep: after `inner` sweeps it has not reached
// the core, which is exact.  The result equals `iterations` resident sweeps bit
// for bit, stochastic rounding included (the slab carries its global plane
// origin and the global sweep index).
//
// Memory.  The state is one pinned host array, updated in place: the core of
// slab i is written back only once every slab of the same pass that reads
// those rows as halo has been uploaded (a lag of ceil(halo / tile) slabs).
// The device holds lag + 2 slab slots.  2D sweeps never write the grid's ring
// (the outer `range` rows and columns), and each buffer of the ping-pong keeps
// its own ring: the rings of both host arrays are kept on the device and
// patched into the slab's second buffer, whose other cells the sweeps
// overwrite or the halo absorbs.
//
// Overlap.  Three streams: uploads, sweeps, downloads.  Slab i+1 uploads while
// slab i sweeps and slab i-1 downloads; events order each slot's reuse, the
// in-place write-back, and each pass's reads after the previous pass's writes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

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

struct Slab {
    size_t h0, h1;  // core rows/planes
    size_t s0, s1;  // staged rows/planes (core plus halo, clamped)
};

std::vector<Slab> make_slabs(size_t H, size_t tile, size_t halo)
{
    std::vector<Slab> v;
    for (size_t h0 = 0; h0 < H; h0 += tile) {
        Slab s;
        s.h0 = h0;
        s.h1 = std::min(H, h0 + tile);
        s.s0 = h0 > halo ? h0 - halo : 0;
        s.s1 = std::min(H, s.h1 + halo);
        v.push_back(s);
    }
    return v;
}

// last slab whose staged rows start before row h (slabs are sorted by s0)
size_t last_reader(const std::vector<Slab>& v, size_t h)
{
    size_t i = 0;
    while (i + 1 < v.size() && v[i + 1].s0 < h) ++i;
    return i;
}

// Device bytes the pipeline holds: lag + 2 slots of the widest staged slab,
// each with 2 buffers (3 in 2D, whose temporal tiles of even depth use a spare).
size_t pipeline_device_bytes(int ndim, size_t W, size_t H, size_t D, int elem, size_t tile, size_t halo)
{
    const std::vector<Slab> widest = make_slabs(H, tile, halo);
    size_t max_rows = 0, lag = 0;
    for (size_t j = 0; j < widest.size(); ++j) {
        max_rows = std::max(max_rows, widest[j].s1 - widest[j].s0);
        lag = std::max(lag, last_reader(widest, widest[j].h1) - j);
    }
    const size_t nslots = std::min(widest.size(), lag + 2);
    const size_t row_bytes = (ndim == 2 ? W : D) * (size_t)elem;
    const size_t pitch_bytes = (row_bytes + 1023) / 1024 * 1024;
    const size_t plane_rows = ndim == 2 ? 1 : W;
    return nslots * pitch_bytes * plane_rows * max_rows * (ndim == 2 ? 3 : 2);
}

// Everything one call owns; the destructor drains the streams before it frees.
struct Pipeline {
    int device = 0;
    unsigned char* state = nullptr;  // pinned host array holding the grid
    hipStream_t up = nullptr, run = nullptr, down = nullptr;
    std::vector<hipEvent_t> events;
    std::vector<void*> dmem;
    pskel_grid_t* view = nullptr;  // the slab being swept (buffers point into a slot)

    ~Pipeline()
    {
        (void)hipSetDevice(device);
        for (hipStream_t s : {up, run, down})
            if (s) (void)hipStreamSynchronize(s);
        delete view;  // its buffers belong to dmem
        for (void* p : dmem) (void)hipFree(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipStream_t s : {up, run, down})
            if (s) (void)hipStreamDestroy(s);
        if (state) (void)hipHostFree(state);
    }
    pskel_status_t event(hipEvent_t* e)
    {
        PSKEL_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
        events.push_back(*e);
        return PSKEL_OK;
    }
    pskel_status_t alloc(void** p, size_t bytes)
    {
        PSKEL_HIP(hipMalloc(p, bytes));
        dmem.push_back(*p);
        return PSKEL_OK;
    }
};

// host rows -> device pitched rows, or back (one copy when rows are unpadded)
hipError_t copy_rows(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t n,
                     hipMemcpyKind kind, hipStream_t s)
{
    if (n == 0 || width == 0) return hipSuccess;
    if (dpitch == width && spitch == width) return hipMemcpyAsync(dst, src, width * n, kind, s);
    return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, n, kind, s);
}

// The ring of one host array of a 2D grid: `range` columns at each side of
// every row and `range` rows at the top and bottom, kept on the device.
struct Ring2D {
    size_t lw = 0, rs = 0, rw = 0;  // left width, right start, right width (columns)
    size_t tn = 0, bs = 0, bn = 0;  // top rows, bottom start, bottom rows
    unsigned char* left = nullptr;  // H x lw
    unsigned char* right = nullptr; // H x rw
    unsigned char* top = nullptr;   // tn x W
    unsigned char* bottom = nullptr;// bn x W
};

pskel_status_t make_ring(Pipeline& P, Ring2D& R, const unsigned char* host, size_t W, size_t H, size_t r,
                         size_t elem)
{
    R.lw = std::min(r, W);
    R.rs = W > r ? W - r : 0;
    R.rw = W - R.rs;
    R.tn = std::min(r, H);
    R.bs = H > r ? H - r : 0;
    R.bn = H - R.bs;
    const size_t nl = H * R.lw, nr = H * R.rw, nt = R.tn * W, nb = R.bn * W;
    if (nl + nr + nt + nb == 0) return PSKEL_OK;  // range 0: no ring
    std::vector<unsigned char> tmp((nl + nr + nt + nb) * elem);
    unsigned char* o = tmp.data();
    for (size_t h = 0; h < H; ++h) {
        std::memcpy(o + h * R.lw * elem, host + h * W * elem, R.lw * elem);
        std::memcpy(o + (nl + h * R.rw) * elem, host + (h * W + R.rs) * elem, R.rw * elem);
    }
    std::memcpy(o + (nl + nr) * elem, host, nt * elem);
    std::memcpy(o + (nl + nr + nt) * elem, host + R.bs * W * elem, nb * elem);
    void* d = nullptr;
    pskel_status_t st = P.alloc(&d, tmp.size());
    if (st != PSKEL_OK) return st;
    PSKEL_HIP(hipMemcpy(d, tmp.data(), tmp.size(), hipMemcpyHostToDevice));
    R.left = (unsigned char*)d;
    R.right = R.left + nl * elem;
    R.top = R.right + nr * elem;
    R.bottom = R.top + nt * elem;
    return PSKEL_OK;
}

// ring cells of rows [s0, s1) of the grid -> a slab buffer (row s0 at row 0)
hipError_t patch_ring(const Ring2D& R, void* dbuf, size_t pitch_bytes, size_t W, size_t s0, size_t s1, size_t elem,
                      hipStream_t s)
{
    unsigned char* d = (unsigned char*)dbuf;
    const size_t n = s1 - s0;
    hipError_t e = copy_rows(d, pitch_bytes, R.left + s0 * R.lw * elem, R.lw * elem, R.lw * elem, n,
                             hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = copy_rows(d + R.rs * elem, pitch_bytes, R.right + s0 * R.rw * elem, R.rw * elem, R.rw * elem, n,
                      hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && s0 < R.tn) {
        const size_t a = s0, b = std::min(s1, R.tn);
        e = copy_rows(d + (a - s0) * pitch_bytes, pitch_bytes, R.top + a * W * elem, W * elem, W * elem, b - a,
                      hipMemcpyDeviceToDevice, s);
    }
    if (e == hipSuccess && s1 > R.bs) {
        const size_t a = std::max(s0, R.bs), b = s1;
        e = copy_rows(d + (a - s0) * pitch_bytes, pitch_bytes, R.bottom + (a - R.bs) * W * elem, W * elem, W * elem,
                      b - a, hipMemcpyDeviceToDevice, s);
    }
    return e;
}

bool life_cells_ok(const void* p, size_t n)
{
    const uint8_t* c = (const uint8_t*)p;
    for (size_t i = 0; i < n; ++i)
        if (c[i] > 1) return false;
    return true;
}

}  // namespace

namespace {
// The pipeline; `pass_ms` (optional) receives the time of the passes alone,
// from the first upload to the last write-back, without the set-up (pinned
// state, slots, streams) and the host copies in and out.
pskel_status_t run_host(const pskel_grid_desc_t* gd, const void* host_in, void* host_out, int32_t host_dtype,
                        const pskel_stencil_desc_t* d, size_t iterations, size_t tile, size_t inner,
                        const pskel_launch_cfg_t* cfg_in, pskel_run_stats_t* stats, double* pass_ms)
{
    if (!gd || !host_in || !host_out || !d) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    const int ndim = gd->ndim;
    if (ndim != 2 && ndim != 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    const int elem = storage_bytes(gd->storage);
    if (elem == 0) return fail(PSKEL_ERR_UNSUPPORTED, "storage type");
    if (ndim == 2 && elem == 2) return fail(PSKEL_ERR_UNSUPPORTED, "16-bit storage is implemented for 3D grids");
    if (ndim == 3 && elem == 1) return fail(PSKEL_ERR_UNSUPPORTED, "u8 storage is implemented for 2D grids");
    if (std::isnan(gd->store_offset))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "host arrays hold storage words: give their storage offset");
    if (host_dtype != gd->storage)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "host-staged sweeps take host arrays in the storage type");
    if (tile == 0 || inner == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "tile and inner must be positive");
    pskel_launch_cfg_t cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    if (cfg_in) cfg = *cfg_in;
    if (cfg.residual_every > 0) return fail(PSKEL_ERR_UNSUPPORTED, "residuals of host-staged sweeps");
    if (d->kind == PSKEL_KIND_CLOUDSIM) return fail(PSKEL_ERR_UNSUPPORTED, "field stencils stay resident");
    cfg.use_graph = 0;
    cfg.stream = nullptr;
    const size_t W = gd->width, H = gd->height, D = ndim == 3 ? gd->depth : 1;
    if (W == 0 || H == 0 || D == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "empty grid");
    if (W > (1u << 30) || H > (1u << 30) || D > (1u << 30))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "dimension too large");
    if (gd->rounding != PSKEL_ROUND_NEAREST && gd->rounding != PSKEL_ROUND_STOCHASTIC)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown rounding mode");
    if (gd->rounding == PSKEL_ROUND_STOCHASTIC && (uint64_t)W * D >= (1ull << 32))
        return fail(PSKEL_ERR_UNSUPPORTED, "stochastic rounding needs width * depth < 2^32");
    if (d->npoints < 1 || d->npoints > PSKEL_MAX_POINTS) return fail(PSKEL_ERR_INVALID_ARGUMENT, "npoints");
    if (stats) {
        stats->elapsed_ms = 0.0;
        stats->launches = 0;
        stats->sweeps = (int64_t)iterations;
        stats->residual_count = 0;
    }
    const size_t plane_rows = ndim == 2 ? 1 : W;  // contiguous rows per h-plane
    const size_t row_bytes = (ndim == 2 ? W : D) * (size_t)elem;
    const size_t plane_bytes = plane_rows * row_bytes;
    const size_t total = H * plane_bytes;
    if (gd->storage == PSKEL_DTYPE_U8 &&
        (!life_cells_ok(host_in, W * H) || !life_cells_ok(host_out, W * H)))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "life cells must be 0 or 1");
    // A slab's wrong edge zone is its ring (range r, never updated on the
    // device) and then grows by the h-reach of the mask per sweep, so a halo of
    // r * inner (r >= the h-reach) keeps the core exact.
    const size_t radius = (size_t)desc_radius(*d, ndim);

    const auto t0 = std::chrono::steady_clock::now();
    PSKEL_HIP(hipSetDevice(gd->device));
    Pipeline P;
    P.device = gd->device;
    PSKEL_HIP(hipHostMalloc((void**)&P.state, total, hipHostMallocDefault));
    par_copy(P.state, host_in, total);
    if (iterations == 0) {  // N = 0: output := input
        par_copy(host_out, P.state, total);
        return PSKEL_OK;
    }
    PSKEL_HIP(hipStreamCreateWithFlags(&P.up, hipStreamNonBlocking));
    PSKEL_HIP(hipStreamCreateWithFlags(&P.run, hipStreamNonBlocking));
    PSKEL_HIP(hipStreamCreateWithFlags(&P.down, hipStreamNonBlocking));

    // slots: the largest slab, lag + 2 of them
    const size_t inner_eff = std::min(inner, iterations);
    const size_t halo_max = radius * inner_eff;
    const std::vector<Slab> widest = make_slabs(H, tile, halo_max);
    size_t max_rows = 0, lag = 0;
    for (size_t j = 0; j < widest.size(); ++j) {
        max_rows = std::max(max_rows, widest[j].s1 - widest[j].s0);
        lag = std::max(lag, last_reader(widest, widest[j].h1) - j);
    }
    const size_t nslots = std::min(widest.size(), lag + 2);
    const size_t pitch = (row_bytes + 1023) / 1024 * 1024 / (size_t)elem;  // as pskel_grid_create
    const size_t pitch_bytes = pitch * (size_t)elem;
    const size_t buf_bytes = pitch_bytes * plane_rows * max_rows;
    const int nbuf = ndim == 2 ? 3 : 2;  // 2D temporal tiles of even depth use a spare
    std::vector<unsigned char*> slot(nslots);
    pskel_status_t st;
    for (size_t k = 0; k < nslots; ++k) {
        void* p = nullptr;
        if ((st = P.alloc(&p, buf_bytes * nbuf)) != PSKEL_OK) return st;
        // stale cells only ever reach a slab's edge zone; start them at 0
        PSKEL_HIP(launch_fill_zero(p, buf_bytes * nbuf, P.run));
        slot[k] = (unsigned char*)p;
    }
    Ring2D ring[2];
    if (ndim == 2) {
        if ((st = make_ring(P, ring[0], (const unsigned char*)host_in, W, H, radius, elem)) != PSKEL_OK) return st;
        if ((st = make_ring(P, ring[1], (const unsigned char*)host_out, W, H, radius, elem)) != PSKEL_OK) return st;
    }
    PSKEL_HIP(hipStreamSynchronize(P.run));

    hipEvent_t up_done;
    std::vector<hipEvent_t> run_done(nslots), slot_free(nslots), rows_done(widest.size());
    std::vector<bool> slot_used(nslots, false);
    if ((st = P.event(&up_done)) != PSKEL_OK) return st;
    for (size_t k = 0; k < nslots; ++k) {
        if ((st = P.event(&run_done[k])) != PSKEL_OK) return st;
        if ((st = P.event(&slot_free[k])) != PSKEL_OK) return st;
    }
    for (auto& e : rows_done)
        if ((st = P.event(&e)) != PSKEL_OK) return st;

    P.view = new (std::nothrow) pskel_grid_t();
    if (!P.view) return fail(PSKEL_ERR_OUT_OF_MEMORY, "host allocation");
    pskel_grid_t* v = P.view;
    v->ndim = ndim;
    v->storage = gd->storage;
    v->elem = elem;
    v->W = W;
    v->D = D;
    v->pitch = pitch;
    v->store_offset = gd->store_offset;
    v->device = gd->device;
    v->rounding = gd->rounding;
    v->seed = gd->seed;
    v->stream = P.run;

    struct Pending {
        size_t slab, slot;
        const unsigned char* src;  // the slab's current buffer after its sweeps
    };
    int64_t launches = 0;
    size_t done = 0, seq = 0;
    const auto t_pass = std::chrono::steady_clock::now();
    while (done < iterations) {
        const size_t sub = std::min(inner, iterations - done);
        const int srcp = (int)(done % 2);  // parity of the state's ring (2D)
        const std::vector<Slab> slabs = make_slabs(H, tile, radius * sub);
        std::vector<Pending> pending;
        size_t next_dl = 0;
        for (size_t i = 0; i < slabs.size(); ++i, ++seq) {
            const Slab& sl = slabs[i];
            const size_t k = seq % nslots;
            const size_t rows = sl.s1 - sl.s0;
            // upload: after this slot's previous slab left, and (after the first
            // pass) after the previous pass wrote back every row this slab reads
            if (slot_used[k]) PSKEL_HIP(hipStreamWaitEvent(P.up, slot_free[k], 0));
            if (done > 0) PSKEL_HIP(hipStreamWaitEvent(P.up, rows_done[(sl.s1 - 1) / tile], 0));
            unsigned char* b0 = slot[k];
            PSKEL_HIP(copy_rows(b0, pitch_bytes, P.state + sl.s0 * plane_bytes, row_bytes, row_bytes,
                                rows * plane_rows, hipMemcpyHostToDevice, P.up));
            PSKEL_HIP(hipEventRecord(up_done, P.up));
            slot_used[k] = true;
            // sweeps
            PSKEL_HIP(hipStreamWaitEvent(P.run, up_done, 0));
            v->H = rows;
            v->nrows = rows * plane_rows;
            v->bytes = pitch_bytes * v->nrows;
            v->buf[0] = b0;
            v->buf[1] = b0 + buf_bytes;
            v->buf[2] = nbuf > 2 ? b0 + 2 * buf_bytes : nullptr;
            v->parity = 0;
            v->spare_parity = -1;
            v->sweeps_done = done;                // global sweep index (stochastic rounding)
            v->plane_origin = (uint32_t)sl.s0;    // global plane index (stochastic rounding)
            if (ndim == 2) PSKEL_HIP(patch_ring(ring[srcp ^ 1], v->buf[1], pitch_bytes, W, sl.s0, sl.s1, elem, P.run));
            int64_t nl = 0;
            if ((st = run_async(v, d, sub, cfg, P.run, &nl)) != PSKEL_OK) return st;
            launches += nl;
            PSKEL_HIP(hipEventRecord(run_done[k], P.run));
            pending.push_back({i, k, (const unsigned char*)v->buf[0]});
            // write back every core that no later slab of this pass reads
            while (next_dl < pending.size() &&
                   (i + 1 == slabs.size() || last_reader(slabs, slabs[pending[next_dl].slab].h1) <= i)) {
                const Pending& pd = pending[next_dl++];
                const Slab& c = slabs[pd.slab];
                PSKEL_HIP(hipStreamWaitEvent(P.down, run_done[pd.slot], 0));
                PSKEL_HIP(hipStreamWaitEvent(P.down, up_done, 0));
                PSKEL_HIP(copy_rows(P.state + c.h0 * plane_bytes, row_bytes,
                                    pd.src + (c.h0 - c.s0) * plane_rows * pitch_bytes, pitch_bytes, row_bytes,
                                    (c.h1 - c.h0) * plane_rows, hipMemcpyDeviceToHost, P.down));
                PSKEL_HIP(hipEventRecord(slot_free[pd.slot], P.down));
                PSKEL_HIP(hipEventRecord(rows_done[pd.slab], P.down));
            }
        }
        done += sub;
    }
    PSKEL_HIP(hipStreamSynchronize(P.down));
    if (pass_ms) *pass_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_pass).count();
    // the state holds the result with runIterativeCPU's placement: N even
    // leaves the input's ring, N odd the output's (include/PSkelStencil.hpp:932)
    par_copy(host_out, P.state, total);
    if (stats) {
        stats->elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->launches = launches;
    }
    return PSKEL_OK;
}
}  // namespace

extern "C" pskel_status_t pskel_run_iterative_host(const pskel_grid_desc_t* gd, const void* host_in, void* host_out,
                                                   int32_t host_dtype, const pskel_stencil_desc_t* d,
                                                   size_t iterations, size_t tile, size_t inner,
                                                   const pskel_launch_cfg_t* cfg_in, pskel_run_stats_t* stats)
{
    return run_host(gd, host_in, host_out, host_dtype, d, iterations, tile, inner, cfg_in, stats, nullptr);
}

// ---------------------------------------------------------------------------
// Out-of-core launch search: runIterativeAutoGPU's tiled branch
// (include/PSkelStencil.hpp:2014-2074) picks the tile and its sub-iterations
// with a GA over an analytic objective under a memory-fit constraint
// (:1883-1895, :1943-1957).  Here each candidate (slab height, sweeps per
// pass) that fits the device budget is timed through the pipeline itself on a
// sample of the grid: its first min(H, 4 * slab) rows or planes, one pass
// (the faster of two), scaled to the planes a pass moves on the whole grid.
// ---------------------------------------------------------------------------
namespace {
struct HostSearchKey {
    std::vector<unsigned char> bytes;
    bool operator<(const HostSearchKey& o) const { return bytes < o.bytes; }
};
std::mutex g_host_search_mutex;
std::map<HostSearchKey, std::pair<size_t, size_t>> g_host_search_cache;
}  // namespace

extern "C" pskel_status_t pskel_tuner_host_search(const pskel_grid_desc_t* gd, const void* host_in, int32_t host_dtype,
                                                  const pskel_stencil_desc_t* d, size_t iterations,
                                                  size_t device_bytes, size_t* tile_out, size_t* inner_out,
                                                  double* ms_per_sweep)
{
    if (!gd || !host_in || !d || !tile_out || !inner_out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    const int ndim = gd->ndim;
    if (ndim != 2 && ndim != 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    const int elem = storage_bytes(gd->storage);
    if (elem == 0 || host_dtype != gd->storage)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "host-staged sweeps take host arrays in the storage type");
    if (d->npoints < 1 || d->npoints > PSKEL_MAX_POINTS) return fail(PSKEL_ERR_INVALID_ARGUMENT, "npoints");
    const size_t W = gd->width, H = gd->height, D = ndim == 3 ? gd->depth : 1;
    if (W == 0 || H == 0 || D == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "empty grid");
    if (iterations == 0) iterations = 1;
    if (device_bytes == 0) {
        size_t fr = 0, tot = 0;
        PSKEL_HIP(hipSetDevice(gd->device));
        PSKEL_HIP(hipMemGetInfo(&fr, &tot));
        device_bytes = fr / 2;
    }
    const size_t radius = (size_t)desc_radius(*d, ndim);
    const size_t plane_cells = ndim == 2 ? W : W * D;
    const size_t plane_bytes = plane_cells * (size_t)elem;
    HostSearchKey key;
    auto put = [&key](const void* p, size_t n) {
        key.bytes.insert(key.bytes.end(), (const unsigned char*)p, (const unsigned char*)p + n);
    };
    put(gd, sizeof(*gd));
    put(d, sizeof(*d));
    put(&iterations, sizeof(iterations));
    put(&device_bytes, sizeof(device_bytes));
    {
        std::lock_guard<std::mutex> lock(g_host_search_mutex);
        auto it = g_host_search_cache.find(key);
        if (it != g_host_search_cache.end()) {
            *tile_out = it->second.first;
            *inner_out = it->second.second;
            if (ms_per_sweep) *ms_per_sweep = 0.0;
            return PSKEL_OK;
        }
    }
    double best = -1.0;
    size_t bt = 0, bi = 0;
    pskel_launch_cfg_t cfg{};
    cfg.variant = PSKEL_VARIANT_TUNED;
    for (size_t tile : {16, 32, 64, 128, 256, 512, 1024}) {
        if (tile > std::max<size_t>(H / 2, 1) && tile != 16) continue;
        for (size_t inner : {2, 4, 8, 16}) {
            if (inner > iterations && inner != 2) continue;
            const size_t in_eff = std::min(inner, iterations);
            if (pipeline_device_bytes(ndim, W, H, D, elem, tile, radius * in_eff) > device_bytes) continue;
            const size_t hs = std::min(H, 4 * tile);
            pskel_grid_desc_t gs = *gd;
            gs.height = hs;
            std::vector<unsigned char> a((const unsigned char*)host_in, (const unsigned char*)host_in + hs * plane_bytes);
            std::vector<unsigned char> b(a);
            // the faster of two passes over the sample, timed without the
            // call's set-up and host copies (the same for every candidate,
            // and noisy: pinning the state)
            double ms = -1.0;
            for (int rep = 0; rep < 2; ++rep) {
                double pm = 0.0;
                if (run_host(&gs, a.data(), b.data(), host_dtype, d, in_eff, tile, in_eff, &cfg, nullptr, &pm) !=
                    PSKEL_OK)
                    break;  // the pipeline cannot run this shape here
                ms = ms < 0.0 ? pm : std::min(ms, pm);
            }
            if (ms < 0.0) continue;
            // The sample's end slabs have their halos clamped, so it moves fewer
            // planes per core plane than the whole grid does: scale by the
            // planes a pass moves (staged up plus core down) on each.
            const size_t halo = radius * in_eff;
            auto moved = [&](size_t h) {
                size_t up = 0;
                for (const Slab& sl : make_slabs(h, tile, halo)) up += sl.s1 - sl.s0;
                return (double)(up + h) / (double)h;
            };
            const double per_cu = ms * (moved(H) / moved(hs)) / ((double)hs * (double)plane_cells * (double)in_eff);
            if (best < 0.0 || per_cu < best) {
                best = per_cu;
                bt = tile;
                bi = in_eff;
            }
        }
    }
    if (best < 0.0) return fail(PSKEL_ERR_OUT_OF_MEMORY, "no slab fits the device budget");
    set_error("");
    *tile_out = bt;
    *inner_out = bi;
    if (ms_per_sweep) *ms_per_sweep = best * (double)H * (double)plane_cells;
    std::lock_guard<std::mutex> lock(g_host_search_mutex);
    g_host_search_cache[key] = {bt, bi};
    return PSKEL_OK;
}
