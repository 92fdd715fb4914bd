// Batched persistent sweeps: many uneven 2D grids advanced together.
//
// The reference runs one Stencil at a time, one launch + sync per sweep
// (include/PSkelStencil.hpp:1078-1082, 1790-1791); a workload of hundreds of
// small grids is then launch-bound.  Here one persistent kernel per sweep
// covers every tile of every grid: a fixed set of blocks walks a work list of
// (grid, tile) items round-robin (static order: deterministic, every block exits
// after its last item), each item running the same register-tiled body as the
// standalone tuned kernels, with the grid's kernel family (Jacobi f32, 3x3
// weighted f32, life u8) and rows per lane chosen per grid.  The sequence of
// sweeps is captured once in a hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <vector>

#include "k_tiles2d.hpp"

using namespace pskel;

namespace {

enum BatchKind : int { BK_J5 = 0, BK_BOX9 = 1, BK_LIFE = 2 };

// One grid of the batch (device table entry).
struct BatchEntry {
    void* buf[2];     // the launch with parity p reads buf[p] and writes buf[p ^ 1]
    int W, H;
    int pitch;
    int radius;
    int kind;
    int ry;           // rows per lane
    int tiles_x;      // tiles along w
    int item0;        // first work item of this grid
    float coeff;
    float wt[9];
};

constexpr int kBatchWaves = 4;  // block = (64, 4)

template <int RY>
__device__ __forceinline__ void run_item(const BatchEntry& E, int t, int parity)
{
    Sweep2D s;
    s.src = E.buf[parity];
    s.dst = E.buf[parity ^ 1];
    s.ring_other = E.buf[parity ^ 1];
    s.W = E.W;
    s.H = E.H;
    s.pitch = (size_t)E.pitch;
    s.radius = E.radius;
    const int tx = t % E.tiles_x, ty = t / E.tiles_x;
    if (E.kind == BK_LIFE) {
        tile_life<RY>(s, tx, ty);
    } else if (E.kind == BK_J5) {
        tile_f32<PAT_J5, RY>(s, E, tx, ty);
    } else {
        tile_f32<PAT_BOX9, RY>(s, E, tx, ty);
    }
}

__global__ __launch_bounds__(256) void k_batch(const BatchEntry* __restrict__ tab, const int* __restrict__ item_grid,
                                               int nitems, int parity)
{
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int g = item_grid[item];
        const BatchEntry& E = tab[g];
        const int t = item - E.item0;
        switch (E.ry) {
        case 1: run_item<1>(E, t, parity); break;
        case 2: run_item<2>(E, t, parity); break;
        case 4: run_item<4>(E, t, parity); break;
        default: run_item<8>(E, t, parity); break;
        }
    }
}

int tile_cols(int kind) { return kind == BK_LIFE ? 1024 : 256; }

}  // namespace

struct pskel_batch {
    std::vector<pskel_grid_t*> grids;
    std::vector<BatchEntry> host_tab;
    BatchEntry* d_tab = nullptr;
    int* d_items = nullptr;
    int nitems = 0;
    int nblocks = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t graph_sweeps = 0;
};

namespace {

// Refresh the per-grid buffer pointers from the grids' current states.
pskel_status_t refresh_table(pskel_batch_t* b)
{
    for (size_t i = 0; i < b->grids.size(); ++i) {
        b->host_tab[i].buf[0] = b->grids[i]->buf[0];
        b->host_tab[i].buf[1] = b->grids[i]->buf[1];
    }
    PSKEL_HIP(hipMemcpyAsync(b->d_tab, b->host_tab.data(), b->host_tab.size() * sizeof(BatchEntry),
                             hipMemcpyHostToDevice, b->stream));
    return PSKEL_OK;
}

hipError_t issue_batch(pskel_batch_t* b, size_t sweeps, hipStream_t st)
{
    for (size_t it = 0; it < sweeps; ++it) {
        hipLaunchKernelGGL(k_batch, dim3(b->nblocks), dim3(kWave, kBatchWaves), 0, st, b->d_tab, b->d_items,
                           b->nitems, (int)(it & 1));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

pskel_status_t pskel_batch_create(pskel_grid_t* const* grids, const pskel_stencil_desc_t* descs,
                                  const pskel_launch_cfg_t* cfgs, size_t n, pskel_batch_t** out)
{
    if (!grids || !descs || !out || n == 0 || n > (1u << 20)) return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad batch");
    *out = nullptr;
    for (size_t i = 0; i < n; ++i)
        if (!grids[i]) return fail(PSKEL_ERR_INVALID_ARGUMENT, "batch grid is NULL");
    auto* b = new pskel_batch_t();
    b->device = grids[0]->device;
    std::vector<int> items;
    for (size_t i = 0; i < n; ++i) {
        pskel_grid_t* g = grids[i];
        if (!g || g->ndim != 2 || g->device != b->device) {
            delete b;
            return fail(PSKEL_ERR_UNSUPPORTED, "batches hold 2D grids of one device");
        }
        const pskel_stencil_desc_t& d = descs[i];
        const int pat = classify(d, 2);
        BatchEntry e{};
        if (pat == PAT_J5 && g->storage == PSKEL_DTYPE_F32) e.kind = BK_J5;
        else if (pat == PAT_BOX9 && g->storage == PSKEL_DTYPE_F32) e.kind = BK_BOX9;
        else if (pat == PAT_LIFE8 && g->storage == PSKEL_DTYPE_U8) e.kind = BK_LIFE;
        else {
            delete b;
            return fail(PSKEL_ERR_UNSUPPORTED, "batched grids run Jacobi 5-point, 3x3 weighted (f32) or life (u8)");
        }
        const int radius = desc_radius(d, 2);
        if (radius < 1) {
            delete b;
            return fail(PSKEL_ERR_INVALID_ARGUMENT, "mask range");
        }
        int ry = cfgs && cfgs[i].tile_y > 0 ? cfgs[i].tile_y : 4;
        if (ry != 1 && ry != 2 && ry != 4 && ry != 8) {
            delete b;
            return fail(PSKEL_ERR_UNSUPPORTED, "batched rows per lane must be 1, 2, 4 or 8");
        }
        e.W = (int)g->W;
        e.H = (int)g->H;
        e.pitch = (int)g->pitch;
        e.radius = radius;
        e.ry = ry;
        e.coeff = d.coeff;
        for (int k = 0; k < 9; ++k) e.wt[k] = d.weights[k];
        e.tiles_x = (e.W + tile_cols(e.kind) - 1) / tile_cols(e.kind);
        const int tiles_y = (e.H + kBatchWaves * ry - 1) / (kBatchWaves * ry);
        e.item0 = (int)items.size();
        for (int t = 0; t < e.tiles_x * tiles_y; ++t) items.push_back((int)i);
        b->grids.push_back(g);
        b->host_tab.push_back(e);
    }
    b->nitems = (int)items.size();
    hipError_t e = hipSetDevice(b->device);
    int cus = 256;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    // persistent: a few resident blocks per CU, never more blocks than items
    b->nblocks = std::max(1, std::min(b->nitems, cus * 4));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipMalloc(&b->d_tab, b->host_tab.size() * sizeof(BatchEntry));
    if (e == hipSuccess) e = hipMalloc(&b->d_items, items.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        pskel_status_t st = hip_fail(e, "batch setup");
        pskel_batch_destroy(b);
        return st;
    }
    *out = b;
    return PSKEL_OK;
}

pskel_status_t pskel_batch_destroy(pskel_batch_t* b)
{
    if (!b) return PSKEL_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PSKEL_OK;
}

pskel_status_t pskel_batch_run(pskel_batch_t* b, size_t iterations, int32_t use_graph, pskel_run_stats_t* stats)
{
    if (!b) return fail(PSKEL_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (stats) {
        stats->elapsed_ms = 0.0;
        stats->launches = 0;
        stats->sweeps = (int64_t)iterations;
        stats->residual_count = 0;
    }
    if (iterations == 0) return PSKEL_OK;
    PSKEL_HIP(hipSetDevice(b->device));
    // the grids' own streams may still be working on their buffers
    for (pskel_grid_t* g : b->grids) PSKEL_HIP(hipStreamSynchronize(g->stream));
    pskel_status_t st = refresh_table(b);
    if (st != PSKEL_OK) return st;
    if (use_graph && (!b->exec || b->graph_sweeps != iterations)) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        hipStream_t cap;
        PSKEL_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) e = issue_batch(b, iterations, cap);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(cap, &graph);
        (void)hipStreamDestroy(cap);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(e, "batch graph");
        b->graph_sweeps = iterations;
    }
    PSKEL_HIP(hipEventRecord(b->ev0, b->stream));
    if (use_graph) PSKEL_HIP(hipGraphLaunch(b->exec, b->stream));
    else {
        const hipError_t e = issue_batch(b, iterations, b->stream);
        if (e != hipSuccess) return hip_fail(e, "batch launch");
    }
    PSKEL_HIP(hipEventRecord(b->ev1, b->stream));
    PSKEL_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    PSKEL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    for (pskel_grid_t* g : b->grids) {
        for (size_t it = 0; it < iterations % 2; ++it) advance_state(*g, 1);
        g->sweeps_done += iterations;
    }
    if (stats) {
        stats->elapsed_ms = ms;
        stats->launches = (int64_t)iterations;
    }
    return PSKEL_OK;
}

pskel_status_t pskel_batch_info(const pskel_batch_t* b, int32_t* nitems, int32_t* nblocks)
{
    if (!b) return fail(PSKEL_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (nitems) *nitems = b->nitems;
    if (nblocks) *nblocks = b->nblocks;
    return PSKEL_OK;
}

}  // extern "C"
