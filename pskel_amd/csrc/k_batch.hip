// Batched persistent sweeps: many uneven 2D grids advanced together.
//
// The reference runs one Stencil at a time, one launch + sync per sweep
// (include/PSkelStencil.hpp:1078-1082, 1790-1791); a workload of hundreds of
// small grids is then launch-bound.  Here one persistent kernel per launch
// covers every tile of every grid: a fixed set of blocks walks a work list of
// (grid, tile) items round-robin (static order: deterministic, every block exits
// after its last item), each item running the body of the standalone kernels
// with the grid's kernel family (Jacobi f32, 3x3 weighted f32, life u8) and
// tile shape.  The sequence of launches is captured once in a hipGraph and
// replayed; each family's kernels run on a stream of their own, side by side
// with the other families' (no grid belongs to two).
//
// Temporal depth.  Each grid fuses K sweeps per launch (K odd: 1 or 3 for the
// f32 stencils, through the LDS temporal tile; 1, 3, 5 or 7 for life, through
// the bit-packed tile; K = 1 runs the register tiles).  An odd number of sweeps
// per launch moves the state to the other buffer, so grid g's launch l reads
// buffer l & 1 and writes the other, with no third buffer.  N sweeps of grid g
// run as N / K launches of K, then the rest r = N mod K as one launch of r
// (odd) or two of r - 1 and 1 (even r): n_g launches.  Work items are ordered by
// n_g, longest first, so launch l's items are a prefix of the list.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "k_tiles2d.hpp"
#define PSKEL_MARCH2D_INSTANTIATE
#include "k_march2d.hpp"

using namespace pskel;

namespace {

enum BatchKind : int { BK_J5 = 0, BK_BOX9 = 1, BK_LIFE = 2 };

// One grid of the batch (device table entry).
struct BatchEntry {
    void* buf[2];     // launch l reads buf[l & 1] and writes buf[(l & 1) ^ 1]
    int W, H;
    int pitch;
    int radius;
    int kind;
    int ry;           // K = 1: rows per lane; K > 1: tile rows (f32: CY; life: bit-tile rows R;
                      // register march: rows per segment)
    int march;        // f32 at K = 3 through the register march (one wave per tile and segment)
    int depth;        // K, sweeps fused per launch (odd)
    int full;         // launches of K sweeps
    int tail0, tail1; // sweeps of the launches after them (odd; 0: none)
    int tiles_x;      // tiles along w
    int ntiles;       // tiles of the grid
    int item0;        // first work item of this grid
    float coeff;
    float wt[9];
};

constexpr int kBatchWaves = 4;  // block = (64, 4)
constexpr int kTTileCX = 120;   // columns of an f32 temporal tile
constexpr int kTTileK = 3;      // depth of the f32 temporal tiles
constexpr int kLifeRows = 32;   // rows of a bit-packed life tile
constexpr int kBatchLds = ttile_lds<kTTileK, kTTileCX, 32>();  // floats per LDS buffer (the largest tile)

__device__ __forceinline__ int sweeps_of(const BatchEntry& E, int launch)
{
    if (launch < E.full) return E.depth;
    if (launch == E.full) return E.tail0;
    if (launch == E.full + 1) return E.tail1;
    return 0;
}

template <int RY>
__device__ __forceinline__ void run_item(const BatchEntry& E, const Sweep2D& s, int t)
{
    const int tx = t % E.tiles_x, ty = t / E.tiles_x;
    if (E.kind == BK_LIFE) {
        tile_life<RY>(s, tx, ty);
    } else if (E.kind == BK_J5) {
        tile_f32<PAT_J5, RY>(s, E, tx, ty);
    } else {
        tile_f32<PAT_BOX9, RY>(s, E, tx, ty);
    }
}

template <int PAT>
__device__ __forceinline__ void run_ttile(const BatchEntry& E, const Sweep2D& s, int t, int k, float* A, float* B)
{
    const int tid = (int)threadIdx.y * kWave + (int)threadIdx.x;
    const int tx = t % E.tiles_x, ty = t / E.tiles_x;
    if (E.ry == 16) ttile_body<PAT, kTTileK, kTTileCX, 16>(s, E, tx, ty, k, tid, A, B);
    else ttile_body<PAT, kTTileK, kTTileCX, 32>(s, E, tx, ty, k, tid, A, B);
}

__device__ __forceinline__ void run_life_bits(const BatchEntry& E, const Sweep2D& s, int item_t, int k)
{
    // an item is four bit tiles, one per wave
    const int t = item_t * kBatchWaves + (int)threadIdx.y;
    if (t >= E.ntiles) return;  // a whole wave; the tile body does not synchronise
    const int lane = (int)threadIdx.x;
    switch (E.depth) {
    case 3: life_bits_tile<3, kLifeRows>(s, E.tiles_x, t, lane, k); break;
    case 5: life_bits_tile<5, kLifeRows>(s, E.tiles_x, t, lane, k); break;
    default: life_bits_tile<7, kLifeRows>(s, E.tiles_x, t, lane, k); break;
    }
}

// A register-march grid: launches of 3 sweeps run one wave per (tile,
// segment) unit, four units per item; its single-sweep tail launches cover the
// grid with register tiles of 4 rows per lane (256 columns x 16 rows), each
// of the grid's items taking every items-th tile.
__device__ __forceinline__ void run_march(const BatchEntry& E, const Sweep2D& s, int item_t, int k)
{
    const int lane = (int)threadIdx.x, wave = (int)threadIdx.y;
    if (k == 3) {
        const int g = item_t * kBatchWaves + wave;
        if (g >= E.ntiles) return;  // the whole wave; the march does not synchronise
        if (E.kind == BK_J5) march2d_wave<PAT_J5, 3, 3>(s, E, E.tiles_x, E.ry, g, lane);
        else march2d_wave<PAT_BOX9, 3, 3>(s, E, E.tiles_x, E.ry, g, lane);
        return;
    }
    const int ni = (E.ntiles + kBatchWaves - 1) / kBatchWaves;  // the grid's items
    const int rtx = (E.W + 255) / 256, nrt = rtx * ((E.H + 4 * kBatchWaves - 1) / (4 * kBatchWaves));
    for (int rt = item_t; rt < nrt; rt += ni) {
        if (E.kind == BK_J5) tile_f32<PAT_J5, 4>(s, E, rt % rtx, rt / rtx);
        else tile_f32<PAT_BOX9, 4>(s, E, rt % rtx, rt / rtx);
    }
}

// Work families, one kernel each so each keeps the registers its body needs
// (register and SWAR tiles, f32 LDS temporal tiles, bit-packed life tiles,
// register-march units).
enum BatchFamily : int { BF_REG = 0, BF_TTILE = 1, BF_BITS = 2, BF_MARCH = 3, BF_COUNT = 4 };

template <int FAM>
__global__ __launch_bounds__(256) void k_batch(const BatchEntry* __restrict__ tab, const int* __restrict__ item_grid,
                                               int nitems, int launch)
{
    __shared__ float lds[FAM == BF_TTILE ? 2 * kBatchLds : 1];
    const int parity = launch & 1;
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int g = item_grid[item];
        const BatchEntry& E = tab[g];
        const int k = sweeps_of(E, launch);  // uniform across the block
        if (k == 0) continue;
        const int t = item - E.item0;
        Sweep2D s;
        s.src = E.buf[parity];
        s.dst = E.buf[parity ^ 1];
        s.ring_other = E.buf[parity ^ 1];
        s.W = E.W;
        s.H = E.H;
        s.pitch = (size_t)E.pitch;
        s.radius = E.radius;
        if constexpr (FAM == BF_REG) {
            switch (E.ry) {
            case 1: run_item<1>(E, s, t); break;
            case 2: run_item<2>(E, s, t); break;
            case 4: run_item<4>(E, s, t); break;
            default: run_item<8>(E, s, t); break;
            }
        } else if constexpr (FAM == BF_BITS) {
            run_life_bits(E, s, t, k);
        } else if constexpr (FAM == BF_MARCH) {
            run_march(E, s, t, k);
        } else {
            if (E.kind == BK_J5) run_ttile<PAT_J5>(E, s, t, k, lds, lds + kBatchLds);
            else run_ttile<PAT_BOX9>(E, s, t, k, lds, lds + kBatchLds);
        }
    }
}

int family_of(const BatchEntry& e)
{
    if (e.depth == 1) return BF_REG;
    if (e.kind == BK_LIFE) return BF_BITS;
    return e.march ? BF_MARCH : BF_TTILE;
}

// the schedule of N sweeps at depth K: full launches of K, then the tail
void schedule(BatchEntry& e, size_t n)
{
    const int K = e.depth;
    e.full = (int)(n / K);
    const int r = (int)(n % K);
    e.tail0 = e.tail1 = 0;
    if (r % 2 == 1) e.tail0 = r;
    else if (r > 0) {
        e.tail0 = r - 1;
        e.tail1 = 1;
    }
}

int launches_of(const BatchEntry& e) { return e.full + (e.tail0 > 0) + (e.tail1 > 0); }

// items of one grid: register and SWAR tiles and f32 temporal tiles are one
// block each, bit-packed life tiles and register-march units one wave each
// (four per item)
int items_of(const BatchEntry& e)
{
    return (e.kind == BK_LIFE && e.depth > 1) || e.march ? (e.ntiles + kBatchWaves - 1) / kBatchWaves : e.ntiles;
}

}  // namespace

struct pskel_batch {
    std::vector<pskel_grid_t*> grids;
    std::vector<BatchEntry> host_tab;
    BatchEntry* d_tab = nullptr;
    int* d_items = nullptr;          // the work items of each family, one list after the other
    int fam_items[BF_COUNT] = {};    // items per family
    int fam_first[BF_COUNT] = {};    // offset of a family's list in d_items
    int fam_blocks[BF_COUNT] = {};   // persistent blocks of a family's kernel
    std::vector<int> launch_items[BF_COUNT];  // items of launch l: a prefix of the family's list
    int nlaunch = 0;
    int nitems = 0;
    int nblocks = 0;
    int device = 0;
    size_t sched_sweeps = 0;        // iteration count the schedule and item order are for
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // one stream per kernel family (default priority): no grid belongs to two
    // families, so the families' launch sequences run side by side, forked
    // from and joined back into the issuing stream by events
    hipStream_t fam_stream[BF_COUNT] = {};
    hipEvent_t fork = nullptr, join[BF_COUNT] = {};
    hipGraphExec_t exec = nullptr;
    size_t graph_sweeps = 0;
};

namespace {

// Schedules and item order for `n` sweeps: in each family's list, grids with
// more launches first.
pskel_status_t plan_batch(pskel_batch_t* b, size_t n)
{
    const size_t ng = b->host_tab.size();
    for (auto& e : b->host_tab) schedule(e, n);
    std::vector<int> order(ng);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return launches_of(b->host_tab[x]) > launches_of(b->host_tab[y]); });
    b->nlaunch = 0;
    for (int g : order) b->nlaunch = std::max(b->nlaunch, launches_of(b->host_tab[g]));
    std::vector<int> items((size_t)b->nitems);
    int fill[BF_COUNT];
    for (int f = 0; f < BF_COUNT; ++f) {
        fill[f] = 0;
        b->launch_items[f].assign((size_t)b->nlaunch, 0);
    }
    for (int g : order) {
        BatchEntry& e = b->host_tab[g];
        const int f = family_of(e);
        e.item0 = fill[f];
        const int ni = items_of(e);
        for (int t = 0; t < ni; ++t) items[(size_t)(b->fam_first[f] + fill[f]++)] = g;
        for (int l = 0; l < launches_of(e); ++l) b->launch_items[f][l] += ni;
    }
    for (int f = 0; f < BF_COUNT; ++f)
        if (fill[f] != b->fam_items[f]) return fail(PSKEL_ERR_RUNTIME, "batch item count changed");
    PSKEL_HIP(hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    PSKEL_HIP(hipStreamSynchronize(b->stream));
    b->sched_sweeps = n;
    return PSKEL_OK;
}

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

// Each family's launches in order on the family's stream, forked from `st`
// and joined back into it (under capture: parallel branches of the graph).
hipError_t issue_batch(pskel_batch_t* b, hipStream_t st)
{
    const dim3 block(kWave, kBatchWaves);
    hipError_t e = hipEventRecord(b->fork, st);
    for (int f = 0; f < BF_COUNT && e == hipSuccess; ++f) {
        if (b->fam_items[f] == 0) continue;
        hipStream_t fs = b->fam_stream[f];
        e = hipStreamWaitEvent(fs, b->fork, 0);
        const int* items = b->d_items + b->fam_first[f];
        for (int l = 0; l < b->nlaunch && e == hipSuccess; ++l) {
            const int n = b->launch_items[f][l];
            if (n == 0) continue;
            const dim3 grid((unsigned)std::min(n, b->fam_blocks[f]));
            if (f == BF_REG) hipLaunchKernelGGL(k_batch<BF_REG>, grid, block, 0, fs, b->d_tab, items, n, l);
            else if (f == BF_TTILE) hipLaunchKernelGGL(k_batch<BF_TTILE>, grid, block, 0, fs, b->d_tab, items, n, l);
            else if (f == BF_BITS) hipLaunchKernelGGL(k_batch<BF_BITS>, grid, block, 0, fs, b->d_tab, items, n, l);
            else hipLaunchKernelGGL(k_batch<BF_MARCH>, grid, block, 0, fs, b->d_tab, items, n, l);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(b->join[f], fs);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, b->join[f], 0);
    }
    return e;
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
    int nitems = 0;
    for (size_t i = 0; i < n; ++i) {
        pskel_grid_t* g = grids[i];
        auto reject = [&](pskel_status_t st, const char* msg) {
            delete b;
            return fail(st, msg);
        };
        if (g->ndim != 2 || g->device != b->device) return reject(PSKEL_ERR_UNSUPPORTED, "batches hold 2D grids of one device");
        const pskel_stencil_desc_t& d = descs[i];
        const int pat = classify(d, 2);
        BatchEntry e{};
        if (pat == PAT_J5 && g->storage == PSKEL_DTYPE_F32) e.kind = BK_J5;
        else if (pat == PAT_BOX9 && g->storage == PSKEL_DTYPE_F32) e.kind = BK_BOX9;
        else if (pat == PAT_LIFE8 && g->storage == PSKEL_DTYPE_U8) e.kind = BK_LIFE;
        else return reject(PSKEL_ERR_UNSUPPORTED, "batched grids run Jacobi 5-point, 3x3 weighted (f32) or life (u8)");
        const int radius = desc_radius(d, 2);
        if (radius < 1) return reject(PSKEL_ERR_INVALID_ARGUMENT, "mask range");
        const int depth = cfgs && cfgs[i].time_tile > 1 ? cfgs[i].time_tile : 1;
        e.depth = depth;
        e.W = (int)g->W;
        e.H = (int)g->H;
        e.pitch = (int)g->pitch;
        e.radius = radius;
        e.coeff = d.coeff;
        for (int k = 0; k < 9; ++k) e.wt[k] = d.weights[k];
        if (depth == 1) {
            const int ry = cfgs && cfgs[i].tile_y > 0 ? cfgs[i].tile_y : 4;
            if (ry != 1 && ry != 2 && ry != 4 && ry != 8)
                return reject(PSKEL_ERR_UNSUPPORTED, "batched rows per lane must be 1, 2, 4 or 8");
            e.ry = ry;
            const int cols = e.kind == BK_LIFE ? 1024 : 256;
            e.tiles_x = (e.W + cols - 1) / cols;
            e.ntiles = e.tiles_x * ((e.H + kBatchWaves * ry - 1) / (kBatchWaves * ry));
        } else if (e.kind == BK_LIFE) {
            // bit-packed tiles of 32 rows, 3, 5 or 7 generations per launch
            if (depth != 3 && depth != 5 && depth != 7)
                return reject(PSKEL_ERR_UNSUPPORTED, "batched life depths are 1, 3, 5 or 7");
            if (radius != 1) return reject(PSKEL_ERR_UNSUPPORTED, "batched life temporal tiles need mask range 1");
            e.ry = kLifeRows;
            constexpr int CW = 62 * 32;
            const int cr = kLifeRows - 2 * depth;
            e.tiles_x = std::max(1, (e.W - 1 + CW - 1) / CW);
            e.ntiles = e.tiles_x * std::max(1, (e.H - 1 + cr - 1) / cr);
        } else if (cfgs && cfgs[i].tile_x == kMarch2DVd) {
            // register march, 3 sweeps per launch: tiles of 62 lane vectors along w,
            // segments of tile_z rows (0: 32)
            if (depth != 3) return reject(PSKEL_ERR_UNSUPPORTED, "the batched register march runs 3 sweeps per launch");
            if (radius != 1 || !march2d_fits(e.W, e.H, g->pitch))
                return reject(PSKEL_ERR_UNSUPPORTED, "the batched register march needs mask range 1 and a grid over 8 cells wide");
            const March2DShape m = march2d_shape(e.W, e.H, cfgs[i].tile_z > 0 ? cfgs[i].tile_z : 32);
            if (m.units < 1) return reject(PSKEL_ERR_UNSUPPORTED, "the batched register march needs interior rows");
            e.march = 1;
            e.ry = m.seg;
            e.tiles_x = m.ntw;
            e.ntiles = m.units;
        } else {
            // LDS temporal tiles of 120 columns and 16 or 32 rows, 3 sweeps per launch
            if (depth != kTTileK) return reject(PSKEL_ERR_UNSUPPORTED, "batched f32 depths are 1 or 3");
            const int cy = cfgs && cfgs[i].tile_y > 0 ? cfgs[i].tile_y : 16;
            if (cy != 16 && cy != 32) return reject(PSKEL_ERR_UNSUPPORTED, "batched f32 temporal tiles have 16 or 32 rows");
            e.ry = cy;
            e.tiles_x = (e.W + kTTileCX - 1) / kTTileCX;
            e.ntiles = e.tiles_x * ((e.H + cy - 1) / cy);
        }
        nitems += items_of(e);
        b->fam_items[family_of(e)] += items_of(e);
        b->grids.push_back(g);
        b->host_tab.push_back(e);
    }
    b->nitems = nitems;
    for (int f = 1; f < BF_COUNT; ++f) b->fam_first[f] = b->fam_first[f - 1] + b->fam_items[f - 1];
    hipError_t e = hipSetDevice(b->device);
    int cus = 256;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    // persistent: at most as many blocks as the device holds at once (the
    // family kernel's occupancy: registers and LDS), never more than items, so
    // no block waits for a slot with a full share of items
    b->nblocks = std::max(1, std::min(b->nitems, cus * 4));
    const void* fam_kernel[BF_COUNT] = {reinterpret_cast<const void*>(&k_batch<BF_REG>),
                                        reinterpret_cast<const void*>(&k_batch<BF_TTILE>),
                                        reinterpret_cast<const void*>(&k_batch<BF_BITS>),
                                        reinterpret_cast<const void*>(&k_batch<BF_MARCH>)};
    for (int f = 0; f < BF_COUNT; ++f) {
        int per_cu = 4;
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fam_kernel[f], kWave * kBatchWaves, 0);
        b->fam_blocks[f] = std::max(1, std::min(b->fam_items[f], cus * std::max(1, per_cu)));
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int f = 0; f < BF_COUNT && e == hipSuccess; ++f) {
        e = hipStreamCreateWithFlags(&b->fam_stream[f], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[f], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipMalloc(&b->d_tab, b->host_tab.size() * sizeof(BatchEntry));
    if (e == hipSuccess) e = hipMalloc(&b->d_items, (size_t)b->nitems * sizeof(int));
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
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (int f = 0; f < BF_COUNT; ++f) {
        if (b->fam_stream[f]) (void)hipStreamSynchronize(b->fam_stream[f]);
        if (b->fam_stream[f]) (void)hipStreamDestroy(b->fam_stream[f]);
        if (b->join[f]) (void)hipEventDestroy(b->join[f]);
    }
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
    if (iterations > (size_t)0x7fffffff) return fail(PSKEL_ERR_INVALID_ARGUMENT, "too many iterations");
    PSKEL_HIP(hipSetDevice(b->device));
    // the grids' own streams may still be working on their buffers
    for (pskel_grid_t* g : b->grids) PSKEL_HIP(hipStreamSynchronize(g->stream));
    pskel_status_t st;
    if (b->sched_sweeps != iterations) {
        if ((st = plan_batch(b, iterations)) != PSKEL_OK) return st;
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if ((st = refresh_table(b)) != PSKEL_OK) return st;
    if (use_graph && (!b->exec || b->graph_sweeps != iterations)) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        hipStream_t cap;
        PSKEL_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) e = issue_batch(b, cap);
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
        const hipError_t e = issue_batch(b, b->stream);
        if (e != hipSuccess) return hip_fail(e, "batch launch");
    }
    PSKEL_HIP(hipEventRecord(b->ev1, b->stream));
    PSKEL_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    PSKEL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    // every launch of a grid fused an odd number of sweeps: the state moved to
    // the other buffer once per launch
    for (size_t i = 0; i < b->grids.size(); ++i) {
        pskel_grid_t* g = b->grids[i];
        if (launches_of(b->host_tab[i]) % 2) advance_state(*g, 1);
        g->sweeps_done += iterations;
    }
    if (stats) {
        stats->elapsed_ms = ms;
        stats->launches = (int64_t)b->nlaunch;
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
