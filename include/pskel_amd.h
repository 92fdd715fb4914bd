/*
 * pskel_amd.h -- C ABI of the MI355X (gfx950) iterative-stencil library.
 *
 * This is the boundary the PSkel drop-in headers (PSkel.h and friends in this
 * directory) call for every GPU method.  Plain C types only: pointers, sizes,
 * fixed-layout structs, status codes.  No exceptions cross it and nothing in it
 * calls exit() -- the reference's gpuErrchk prints and exits
 * (include/PSkelDefs.h:50-56); here every failure is a status plus a message
 * (pskel_last_error).
 *
 * Reference paths below are relative to the PSkel source tree.
 */
#ifndef PSKEL_AMD_H
#define PSKEL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSKEL_AMD_ABI_VERSION 1
#define PSKEL_MAX_POINTS 64

typedef enum pskel_status {
    PSKEL_OK = 0,
    PSKEL_ERR_INVALID_ARGUMENT = 1,
    PSKEL_ERR_OUT_OF_MEMORY = 2,
    PSKEL_ERR_NO_DEVICE = 3,
    PSKEL_ERR_LAUNCH = 4,
    PSKEL_ERR_UNSUPPORTED = 5,
    PSKEL_ERR_RUNTIME = 6
} pskel_status_t;

/* Element types.  U8/F32 are what Array2D<unsigned char>/Array2D<float> hold;
 * BF16/F16 are device storage formats (host side stays F32, or raw 16-bit words);
 * F64 is accepted only as a host type for upload/download conversions. */
typedef enum pskel_dtype {
    PSKEL_DTYPE_U8 = 1,
    PSKEL_DTYPE_F32 = 2,
    PSKEL_DTYPE_F64 = 3,
    PSKEL_DTYPE_BF16 = 4,
    PSKEL_DTYPE_F16 = 5
} pskel_dtype_t;

/* Stencil families.  A user stencilKernel written against Mask::get in mask
 * order (include/PSkelStencil.h:121-156, include/PSkelMask.hpp:186-199, 292-305)
 * computes one of these:
 *   SCALED_SUM   out = coeff * (((p0 + p1) + p2) + ...)
 *   WEIGHTED_SUM out = ((w0*p0 + w1*p1) + w2*p2) + ...     (w_i = Mask weight i)
 *   LIFE         n = p0 + p1 + ...; out = (n == 3 || (n == 2 && centre == 1))
 *   CLOUDSIM     2D wind advection + conduction with two field arguments, the
 *                CLOUDSIM_KERNEL formula (include/PSkelStencil.hpp:349-387) on
 *                the global neighbours of every interior cell (h, w):
 *                  c = in(h,w); n = in(h-1,w); e = in(h,w+1); w_ = in(h,w-1); s = in(h+1,w)
 *                  sum = 4*c - (((n + e) + w_) + s)
 *                  xw = wind_x(h,w); yw = wind_y(h,w)   (field slots 0 and 1)
 *                  xf = xw > 0 ? 1 : -1; yf = yw > 0 ? 1 : -1
 *                  wind = (-(xf*xw) * ((c - in(h,w+xf)) * 10)) - ((yf*yw) * ((c - in(h+yf,w)) * 10))
 *                  out = (c + (-0.0243 * (sum * 0.25)) * dt) + wind * dt     (dt = coeff)
 * Each product and sum rounds on its own (no fused multiply-add), so results
 * are bit-identical to runIterativeCPU built the same way. */
typedef enum pskel_kind {
    PSKEL_KIND_SCALED_SUM = 1,
    PSKEL_KIND_WEIGHTED_SUM = 2,
    PSKEL_KIND_LIFE = 3,
    PSKEL_KIND_CLOUDSIM = 4
} pskel_kind_t;

/* 3D boundary rule (Stencil3D visits every cell, include/PSkelStencil.hpp:2095-2113):
 *   COPY  cells within `radius` of a face keep their value (out = in)
 *   HALO  every cell is updated; neighbours outside read halo_value
 *         (Mask3D::get, include/PSkelMask.hpp:186-199)
 * 2D grids always update the interior [r, N-r) only (Stencil2D::runOpenMP,
 * include/PSkelStencil.hpp:2235-2255); the ring keeps each buffer's values. */
typedef enum pskel_boundary {
    PSKEL_BOUNDARY_COPY = 0,
    PSKEL_BOUNDARY_HALO = 1
} pskel_boundary_t;

/* Mask2D/Mask3D + the kernel family (include/PSkelMask.h:47-262).
 * offsets: (dh, dw) per point for 2D, (dh, dw, dd) for 3D, as Mask::set takes them. */
typedef struct pskel_stencil_desc {
    int32_t kind;
    int32_t npoints;
    int32_t offsets[3 * PSKEL_MAX_POINTS];
    float weights[PSKEL_MAX_POINTS];
    float coeff;
    int32_t radius;     /* 0: Mask::getRange() of the offsets (include/PSkelMask.hpp:308-334) */
    int32_t boundary;   /* pskel_boundary_t, 3D only */
    float halo_value;   /* 3D HALO */
    int32_t reserved[4];
} pskel_stencil_desc_t;

typedef enum pskel_variant {
    PSKEL_VARIANT_PLAIN = 0, /* one lane per cell, caller's block shape: the runCUDA contract */
    PSKEL_VARIANT_TUNED = 1  /* register/LDS-tiled kernels, optional temporal tiles */
} pskel_variant_t;

/* Launch configuration (replaces the GPUBlockSizeX/Y, tilingWidth/Height and
 * innerIterations arguments of runIterativeGPU / runIterativeTilingGPU,
 * include/PSkelStencil.h:258-297). Zero fields take defaults. */
typedef struct pskel_launch_cfg {
    int32_t variant;        /* pskel_variant_t */
    int32_t block_x;        /* plain: threads along the contiguous axis (w in 2D, d in 3D) */
    int32_t block_y;        /* plain: threads along the next axis (h in 2D, w in 3D) */
    int32_t tile_x;         /* tuned: cells per lane along the contiguous axis (vector width) */
    int32_t tile_y;         /* tuned: rows per lane (2D) / w-rows per wave (3D) */
    int32_t tile_z;         /* tuned 3D: planes per block segment along h; 2D: waves per block */
    int32_t time_tile;      /* tuned: sweeps fused per launch (temporal depth k); 0/1 = none.
                               2D: k <= 4 (8 for life): tile (0, CY, 0) LDS tiles of CY in
                               {8,16,32,64} rows (0: 32), or tile (4, 0, S) the register march,
                               S rows per segment (0: automatic; grids wider than 8 cells);
                               3D 7-point Jacobi: 2 (any storage) or 3 (f32: tile (4,4,0) block
                               (512,2), the default, or tile (4,2,0) block (768,2)); a launch
                               never passes a residual step */
    int32_t residual_every; /* FP64 residual ||u^k - u^(k-1)||_2 every n sweeps; 0 = never */
    int32_t use_graph;      /* capture the launch sequence in a hipGraph and replay it once */
    int32_t reserved[3];
    void* stream;           /* hipStream_t; NULL: the grid's own stream */
} pskel_launch_cfg_t;

typedef struct pskel_run_stats {
    double elapsed_ms;        /* device time between events around the launch sequence */
    int64_t launches;         /* kernel launches issued for the sweeps */
    int64_t sweeps;           /* sweeps performed (== iterations) */
    double* residuals;        /* caller buffer for FP64 residuals (may be NULL) */
    int64_t residual_capacity;
    int64_t residual_count;   /* residuals produced (written up to capacity) */
} pskel_run_stats_t;

typedef struct pskel_grid pskel_grid_t;

/* Rounding of BF16/F16 stores.  STOCHASTIC rounds the magnitude up with
 * probability equal to the dropped fraction of an ulp, from random bits that
 * are the XOR of a hash of (seed, in-plane cell index w*D + d) and a hash of
 * (seed, index of the sweep since the last INPUT upload, plane h): unbiased, so
 * slow relaxation does not stall below half an ulp, pairwise independent between
 * distinct (cell, sweep) pairs, and reproducible run to run. */
typedef enum pskel_rounding {
    PSKEL_ROUND_NEAREST = 0,
    PSKEL_ROUND_STOCHASTIC = 1
} pskel_rounding_t;

/* store_offset of a grid whose 16-bit storage offset follows its data */
#define PSKEL_STORE_OFFSET_AUTO (__builtin_nanf(""))

typedef struct pskel_grid_desc {
    int32_t ndim;        /* 2 or 3 */
    int32_t storage;     /* pskel_dtype_t on the device: U8, F32, BF16 or F16 */
    size_t width;        /* Array2D/Array3D width  (include/PSkelArray.h:295-384) */
    size_t height;       /* height */
    size_t depth;        /* depth (1 for 2D); the contiguous axis in 3D (include/PSkelArray.hpp:642) */
    float store_offset;  /* BF16/F16 storage holds (u - store_offset); NaN (PSKEL_STORE_OFFSET_AUTO):
                            each INPUT upload of F32/F64 data sets it to the data's mean */
    int32_t device;      /* HIP device ordinal */
    int32_t rounding;    /* pskel_rounding_t of BF16/F16 stores */
    uint32_t seed;       /* stochastic rounding seed */
} pskel_grid_desc_t;

typedef enum pskel_buffer {
    PSKEL_BUFFER_INPUT = 0,  /* the Stencil's `input` array */
    PSKEL_BUFFER_OUTPUT = 1  /* the Stencil's `output` array */
} pskel_buffer_t;

typedef struct pskel_grid_info {
    int32_t ndim, storage, elem_bytes, parity;
    size_t width, height, depth, pitch, bytes_per_buffer;
    float store_offset;  /* the offset 16-bit storage holds u - store_offset against */
    void* current;   /* device pointer of the buffer holding the latest state */
    void* previous;  /* device pointer of the other ping-pong buffer */
    void* stream;    /* the grid's hipStream_t */
} pskel_grid_info_t;

/* ---- library ------------------------------------------------------------ */
int32_t pskel_abi_version(void);
/* sizeof of the ABI structs, for bindings to verify their layout:
 * 0 stencil_desc, 1 launch_cfg, 2 run_stats, 3 grid_desc, 4 grid_info,
 * 5 field_desc; else 0. */
size_t pskel_abi_struct_size(int32_t which);
const char* pskel_status_string(pskel_status_t st);
/* Message of the last failure on the calling thread ("" if none). */
const char* pskel_last_error(void);
pskel_status_t pskel_device_count(int32_t* count);
pskel_status_t pskel_device_synchronize(void);
/* Device name and HBM size of `device` (cudaGetDeviceProperties in runGPU,
 * include/PSkelStencil.hpp:733-737; cudaMemGetInfo in runAutoGPU, :834). */
pskel_status_t pskel_device_info(int32_t device, char* name, size_t name_cap,
                                 size_t* mem_free, size_t* mem_total, int32_t* num_cus);

/* ---- device residency (replaces ArrayBase::deviceAlloc/deviceFree,
 *      include/PSkelArray.hpp:72-95, and copyToDevice/copyToHost/copyFromDevice,
 *      include/PSkelArray.hpp:469-608).  Buffers persist across runs: sweeps
 *      ping-pong between them with no host copies.  Both buffers are zeroed at
 *      creation; the reference leaves the device output uninitialised
 *      (include/PSkelStencil.hpp:1076). */
pskel_status_t pskel_grid_create(const pskel_grid_desc_t* desc, pskel_grid_t** out);
pskel_status_t pskel_grid_destroy(pskel_grid_t* grid);
/* Copy a host array (tightly packed, Array2D/Array3D layout) into the INPUT or
 * OUTPUT buffer.  Uploading INPUT makes it the current state (parity 0).
 * host_dtype: the storage type (raw copy) or F32/F64 (converted, rounding to
 * nearest even for BF16/F16 storage).  With an automatic storage offset an
 * INPUT upload takes F32/F64 data and first sets the offset to the mean of its
 * finite values; OUTPUT uploads encode against the offset in force (upload
 * INPUT first: 3D sweeps overwrite every cell of OUTPUT anyway). */
pskel_status_t pskel_grid_upload(pskel_grid_t* grid, int32_t which, const void* host,
                                 int32_t host_dtype);
/* Copy the current state (the result of the last run) into a host array. */
pskel_status_t pskel_grid_download(pskel_grid_t* grid, void* host, int32_t host_dtype);
/* Copy the previous ping-pong state (the state one sweep before the result
 * when the last launch was a single sweep). */
pskel_status_t pskel_grid_download_previous(pskel_grid_t* grid, void* host, int32_t host_dtype);
pskel_status_t pskel_grid_info(const pskel_grid_t* grid, pskel_grid_info_t* info);

/* ---- the hot path: StencilBase::runIterativeGPU (include/PSkelStencil.hpp:1047-1097)
 *      -> runCUDA (:1765-1792) -> stencilTilingCU (:572-652).  Runs `iterations`
 *      sweeps from the current state.  Successive calls resume where the last one
 *      stopped, so run(k) then run(n-k) is bit-identical to run(n).
 *      iterations == 0 leaves the state unchanged (the reference's output := input,
 *      include/PSkelStencil.hpp:932). */
pskel_status_t pskel_run_iterative(pskel_grid_t* grid, const pskel_stencil_desc_t* desc,
                                   size_t iterations, const pskel_launch_cfg_t* cfg,
                                   pskel_run_stats_t* stats);

/* ---- host-staged sweeps for grids that stay in host memory:
 *      runIterativeTilingGPU (include/PSkelStencil.hpp:1663-1760) and the tiled
 *      branch of runIterativeAutoGPU (:2014-2074).  Passes of `inner` sweeps;
 *      each stages slabs of `tile` rows (2D) or h-planes (3D) plus a halo of
 *      range*inner on the device and copies the slab's core back.  `host_in` is
 *      read only; `host_out` holds the initial output (its ring) and receives
 *      the result, equal to `iterations` resident sweeps bit for bit.  Host
 *      arrays are in the storage type (raw words for BF16/F16); `desc` gives
 *      the full grid (`desc->height` rows or planes).  No residuals, no field
 *      stencils. */
pskel_status_t pskel_run_iterative_host(const pskel_grid_desc_t* desc, const void* host_in, void* host_out,
                                        int32_t host_dtype, const pskel_stencil_desc_t* stencil,
                                        size_t iterations, size_t tile, size_t inner,
                                        const pskel_launch_cfg_t* cfg, pskel_run_stats_t* stats);

/* ---- allocation-free run.  No counterpart in the reference: it is for
 *      callers that cannot allocate device memory (serving loops, harnesses
 *      that forbid hipMalloc).  Runs `iterations` sweeps of `stencil`, as
 *      pskel_run_iterative does, on two caller-owned device buffers instead
 *      of a grid; it calls no hipMalloc, hipFree or other allocating routine.
 *      `desc` gives the dims, storage, offset, rounding, seed and device
 *      (16-bit buffers hold the storage words of u - store_offset).  Each
 *      buffer holds nrows rows of `pitch` elements (nrows = height in 2D,
 *      height*width in 3D; row h (2D) or (h, w) (3D) at row*pitch); pitch must
 *      cover a row (width in 2D, depth in 3D) and be a whole number of 1 KiB;
 *      both buffers 256-byte aligned.  `in` is the initial state and is
 *      overwritten with intermediate states, as runIterativeCPU overwrites
 *      `input` (include/PSkelStencil.hpp:903-932); the result of the
 *      iterations is left in `out` (iterations == 0 copies `in`).  2D
 *      temporal tiles fuse an odd number of sweeps per launch (even depths
 *      need a third buffer), which gives the same result.  Stochastic rounding
 *      counts sweeps from 0 at every call.  Launches go to cfg->stream (NULL:
 *      the null stream); the call returns once they are queued, or, when
 *      stats->residuals is given, once the residuals are copied.  Residuals
 *      use a static scratch array of each GPU (65536 FP64 block partials +
 *      4096 results, 544 KiB, held until the process ends): at most 4096
 *      residuals per call and launches of at most 65536 blocks with
 *      residuals, else PSKEL_ERR_INVALID_ARGUMENT.  Calls on one GPU share
 *      that array: calls that overlap in time on one GPU (other streams or
 *      threads) race on it.  Not supported here: graphs (use_graph),
 *      field-argument stencils; stats->elapsed_ms stays 0 (no events). */
pskel_status_t pskel_run_iterative_noalloc(const pskel_grid_desc_t* desc, void* in, void* out, size_t pitch,
                                           const pskel_stencil_desc_t* stencil, size_t iterations,
                                           const pskel_launch_cfg_t* cfg, pskel_run_stats_t* stats);

/* Out-of-core launch search: the slab height (`tile`) and sweeps per pass
 * (`inner`) for pskel_run_iterative_host.  Replaces the tiled branch of
 * runIterativeAutoGPU (include/PSkelStencil.hpp:2014-2074), whose GA picks the
 * tile and its sub-iterations by an analytic objective under a memory-fit
 * constraint (:1883-1895, :1943-1957).  Candidates: slabs of 16 .. 1024
 * rows/planes (at most half the grid) and 2, 4, 8 or 16 sweeps per pass whose
 * device slots fit in `device_bytes` (0: half the free memory); each is timed
 * through pskel_run_iterative_host on a sample of `host_in` (its first
 * min(height, 4 * tile) rows/planes, one pass, copied: host_in is only read).
 * *tile_out, *inner_out: the fastest per cell-update; ms_per_sweep (may be
 * NULL): its time per sweep scaled to the whole grid (0 when the result comes
 * from the process cache: a shape, stencil, iteration count and budget are
 * searched once). */
pskel_status_t pskel_tuner_host_search(const pskel_grid_desc_t* desc, const void* host_in, int32_t host_dtype,
                                       const pskel_stencil_desc_t* stencil, size_t iterations, size_t device_bytes,
                                       size_t* tile_out, size_t* inner_out, double* ms_per_sweep);

/* Record one sweep launched outside the library on the grid's buffers (the
 * drop-in headers' device path for user stencil kernels with no descriptor):
 * the previous buffer becomes the current state, as runIterativeGPU's
 * in/out swap does (include/PSkelStencil.hpp:1078-1082), and counts it in the
 * sweep index that stochastic rounding draws from. */
pskel_status_t pskel_grid_swap(pskel_grid_t* grid);

/* FP64 ||current - previous||_2 over every cell (the convergence test the
 * reference leaves to the caller). */
pskel_status_t pskel_grid_residual(pskel_grid_t* grid, double* out);

/* ---- field arguments: Args / Args2D / Args3D (include/PSkelArgs.h:31-132,
 *      include/PSkelArgs.hpp:49-161), device-resident instead of the
 *      reference's zero-copy mapped host memory, so sweeps read them from HBM.
 *      A field has the Args extents (width, height, depth) = (x, y, z) ranges;
 *      its host data is in the Args layout, x fastest: element (x, y, z) at
 *      (z*height + y)*width + x (include/PSkelArgs.hpp:116-123, 155-160).  On
 *      the device the axes are reversed so that a kernel reading args.f(h, w)
 *      or args.f(h, w, d), as the reference's kernels call them (x = h,
 *      y = w, z = d; include/PSkelStencil.hpp:366-369), reads along the grid's
 *      contiguous axis: element (x, y) at x*pitch + y, element (x, y, z) at
 *      (x*height + y)*pitch + z, element x of a 1D field at x. */
typedef struct pskel_field pskel_field_t;

typedef struct pskel_field_desc {
    int32_t ndim;        /* 1, 2 or 3: Args, Args2D, Args3D */
    int32_t dtype;       /* PSKEL_DTYPE_F32 */
    size_t width;        /* x extent */
    size_t height;       /* y extent (1 for 1D) */
    size_t depth;        /* z extent (1 for 1D and 2D) */
    int32_t device;      /* HIP device ordinal */
    int32_t reserved[3];
} pskel_field_desc_t;

pskel_status_t pskel_field_create(const pskel_field_desc_t* desc, pskel_field_t** out);
pskel_status_t pskel_field_destroy(pskel_field_t* field);
/* host -> device, host in the Args layout (x fastest) */
pskel_status_t pskel_field_upload(pskel_field_t* field, const void* host);
/* device -> host, into the Args layout */
pskel_status_t pskel_field_download(const pskel_field_t* field, void* host);
/* device base pointer and pitch (elements) of the reversed layout above */
pskel_status_t pskel_field_info(const pskel_field_t* field, void** device_ptr, size_t* pitch);
/* Attach a field to slot 0..3 of a grid (not owned; NULL detaches).  CLOUDSIM
 * reads wind_x from slot 0 and wind_y from slot 1; both must be 2D fields with
 * width >= grid height and height >= grid width. */
pskel_status_t pskel_grid_set_field(pskel_grid_t* grid, int32_t slot, pskel_field_t* field);

/* ---- batched persistent sweeps: many uneven 2D grids advanced together by
 *      persistent kernels (a fixed set of blocks walking a list of (grid,
 *      tile) work items), the launch sequence captured in one hipGraph.  The
 *      reference runs one Stencil at a time with a launch and a host sync per
 *      sweep (include/PSkelStencil.hpp:1078-1082, 1790-1791).
 *      Grids: 2D, one device, Jacobi 5-point or 3x3 weighted (F32) or life
 *      (U8).  cfgs[i].time_tile is grid i's temporal depth, the sweeps it fuses
 *      per launch: 1 (or 0) or 3 for F32 (LDS temporal tiles), 1, 3, 5 or 7
 *      for life (bit-packed tiles, mask range 1); odd, so every launch moves a
 *      grid's state to its other buffer.  N sweeps run as N / k launches of k
 *      and the rest as one odd launch or two (r - 1, then 1).  cfgs[i].tile_y:
 *      at depth 1 the rows per lane (1, 2, 4, 8; default 4), at depth 3 in F32
 *      the tile rows (16 or 32; default 16).  F32 at depth 3 with tile_x = 4
 *      runs the register march instead (mask range 1, grids over 8 cells wide):
 *      one wave per tile of 62 x 4 columns and segment of tile_z rows (0: 32),
 *      its single-sweep tail launches through register tiles.  cfgs may be
 *      NULL (depth 1, 4 rows). */
typedef struct pskel_batch pskel_batch_t;
pskel_status_t pskel_batch_create(pskel_grid_t* const* grids, const pskel_stencil_desc_t* descs,
                                  const pskel_launch_cfg_t* cfgs, size_t n, pskel_batch_t** out);
/* `iterations` sweeps of every grid; use_graph: replay a captured hipGraph
 * (captured on first use for this iteration count).  stats->launches: the
 * launch steps (the longest grid's schedule). */
pskel_status_t pskel_batch_run(pskel_batch_t* b, size_t iterations, int32_t use_graph,
                               pskel_run_stats_t* stats);
pskel_status_t pskel_batch_info(const pskel_batch_t* b, int32_t* nitems, int32_t* nblocks);
pskel_status_t pskel_batch_destroy(pskel_batch_t* b);

/* ---- launch tuner (replaces runIterativeAutoGPU's GAlib search over an
 *      analytic objective, include/PSkelStencil.hpp:1865-2076).  Candidates
 *      are the launch configurations of the kernel family a descriptor maps
 *      to; a learned cost model (pskel_mlp_t, below) predicts log time per
 *      sweep from pskel_tuner_features and the tuner takes the argmin.
 *      Exhaustive search over measured times (pskel_tuner_measure) is the
 *      fallback and the yardstick of the model's regret. */
typedef struct pskel_tune_query {
    int32_t ndim;
    int32_t storage;     /* pskel_dtype_t of the grid */
    size_t width, height, depth;
    size_t iterations;
} pskel_tune_query_t;

#define PSKEL_TUNER_MAX_CANDIDATES 256
#define PSKEL_TUNER_NFEAT 26

typedef struct pskel_mlp pskel_mlp_t;

pskel_status_t pskel_tuner_candidates(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d,
                                      pskel_launch_cfg_t* out, size_t cap, size_t* n);
/* Model input of one (query, descriptor, candidate): PSKEL_TUNER_NFEAT floats. */
pskel_status_t pskel_tuner_features(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d,
                                    const pskel_launch_cfg_t* c, float* feat);
/* The candidate with the lowest predicted time.  model NULL: the model installed
 * with pskel_tuner_set_default_model, else the model shipped next to the
 * library (<libdir>/data/costmodel.bin, loaded on first use; the environment
 * variable PSKEL_TUNER_MODEL names another file, or "none"), else a fixed
 * heuristic per kernel family (the drop-in headers search by measurement,
 * pskel_tuner_search, instead). */
pskel_status_t pskel_tuner_choose(const pskel_mlp_t* model, const pskel_tune_query_t* q,
                                  const pskel_stencil_desc_t* d, pskel_launch_cfg_t* out);
pskel_status_t pskel_tuner_set_default_model(pskel_mlp_t* model);
/* The fixed per-family heuristic alone: what pskel_tuner_choose falls back to
 * without a model (for measuring the model against it). */
pskel_status_t pskel_tuner_heuristic(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d,
                                     pskel_launch_cfg_t* out);
/* Device time per sweep of `sweeps` sweeps of candidate c on grid g (median of
 * reps runs after one warm-up; the grid state advances). */
pskel_status_t pskel_tuner_measure(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t sweeps,
                                   const pskel_launch_cfg_t* c, int32_t reps, double* ms_per_sweep);

/* 1 in *out when pskel_tuner_choose(NULL, ...) has a cost model (installed, or
 * the file shipped next to the library, or the file named by the environment
 * variable PSKEL_TUNER_MODEL; PSKEL_TUNER_MODEL=none: no model), else 0. */
pskel_status_t pskel_tuner_has_model(int32_t* out);

typedef enum pskel_search {
    PSKEL_SEARCH_EXHAUSTIVE = 0, /* time every candidate */
    PSKEL_SEARCH_GENETIC = 1     /* simple GA over candidate indices, each distinct candidate timed once */
} pskel_search_t;

/* Measured search over the candidates of grid g's shape, storage and stencil
 * for `iterations` sweeps: the tuner's fallback when there is no cost model,
 * and the yardstick of the model's regret.  Replaces the GAlib search of
 * runIterativeAutoGPU (include/PSkelStencil.hpp:1865-2005), which scored an
 * analytic objective and timed nothing.  Each candidate is timed with
 * pskel_tuner_measure (`sweeps` sweeps, 0: min(iterations, 24); median of
 * `reps` after a warm-up); GENETIC runs population 16 for 8 generations
 * (binary genome, sigma-truncation fitness, one-point crossover 0.6, bit
 * mutation 0.01) from a fixed seed.  The grid's state advances: upload again
 * before the real run.  Results are kept for the process: the same shape,
 * stencil, method and sweep count are searched once.  out: the fastest
 * candidate; ms_per_sweep and measured (may be NULL): its time per sweep and
 * the number of candidates timed by this call (0 when cached). */
pskel_status_t pskel_tuner_search(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t iterations,
                                  int32_t method, size_t sweeps, int32_t reps, pskel_launch_cfg_t* out,
                                  double* ms_per_sweep, int32_t* measured);

/* ---- cost model: a small MLP trained and evaluated on the device by the
 *      library's own kernels (forward, backward, Adam).  Inputs are
 *      standardised per feature and the target by (y_mean, y_std), both part
 *      of the model.  Parameters: one f32 vector in layer order
 *      [W_1 (out x in, row-major), b_1, W_2, b_2, ...]; tanh hidden layers,
 *      one linear output.  No counterpart in the reference, which searches
 *      with GAlib (include/PSkelStencil.hpp:1865-2005). */
pskel_status_t pskel_mlp_create(const int32_t* sizes, int32_t nlayers, uint64_t seed, int32_t device,
                                pskel_mlp_t** out);
pskel_status_t pskel_mlp_destroy(pskel_mlp_t* m);
pskel_status_t pskel_mlp_num_params(const pskel_mlp_t* m, size_t* n);
pskel_status_t pskel_mlp_get_params(const pskel_mlp_t* m, float* params, size_t n);
pskel_status_t pskel_mlp_set_params(pskel_mlp_t* m, const float* params, size_t n);
pskel_status_t pskel_mlp_set_normalization(pskel_mlp_t* m, const float* mean, const float* std,
                                           float y_mean, float y_std);
/* One forward + backward pass: mean squared error of the standardised target
 * and its gradient with respect to the parameters. */
pskel_status_t pskel_mlp_loss_grad(pskel_mlp_t* m, const float* x, const float* y, size_t n,
                                   float* loss, float* grads);
/* Minibatch Adam; final_loss is the mean batch loss of the last epoch. */
pskel_status_t pskel_mlp_train(pskel_mlp_t* m, const float* x, const float* y, size_t n,
                               int32_t epochs, int32_t batch, float lr, uint64_t seed,
                               float* final_loss);
pskel_status_t pskel_mlp_predict(pskel_mlp_t* m, const float* x, size_t n, float* y);
pskel_status_t pskel_mlp_save(const pskel_mlp_t* m, const char* path);
pskel_status_t pskel_mlp_load(const char* path, int32_t device, pskel_mlp_t** out);

/* ---- utilities ------------------------------------------------------------ */
/* Plain device-to-device copy kernel (the bandwidth yardstick for roofline
 * calibration); ms (may be NULL) receives its event-timed duration. */
pskel_status_t pskel_device_copy(void* dst, const void* src, size_t bytes, void* stream,
                                 double* ms);
/* Device buffers for tools and tests (hipMalloc / hipFree / hipMemcpy). */
pskel_status_t pskel_device_alloc(size_t bytes, void** out);
pskel_status_t pskel_device_free(void* ptr);
pskel_status_t pskel_memcpy(void* dst, const void* src, size_t bytes, int32_t to_device);

#ifdef __cplusplus
}
#endif

#endif /* PSKEL_AMD_H */
