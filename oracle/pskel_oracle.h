/*
 * pskel_oracle.h -- CPU restatement of PSkel's iterative stencil runners.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for cpu_baseline, as the timed CPU port).  The product library
 * (libpskel_amd.so) never links, loads or calls it.
 *
 * What it restates (reference paths relative to the PSkel tree):
 *   - StencilBase::runIterativeCPU / runIterativeSequential ping-pong and the
 *     even-N `output.hostMemCopy(input)`      include/PSkelStencil.hpp:864-938
 *   - Stencil2D::runSeq / runOpenMP: interior [r, N-r) only, rows in parallel
 *                                             include/PSkelStencil.hpp:2184-2257
 *   - Stencil3D::runSeq / runOpenMP: every cell, the kernel owns the faces
 *                                             include/PSkelStencil.hpp:2095-2113
 *   - Array2D / Array3D index maps h*W+w and (h*W+w)*D+d
 *                                             include/PSkelArray.hpp:636-677
 *   - Mask2D/Mask3D::getRange (Chebyshev radius), Mask3D::get halo value
 *                                             include/PSkelMask.hpp:186-199, 292-334
 *   - the user stencilKernel contract         include/PSkelStencil.h:121-156
 * The per-cell arithmetic is the stencil family fixed by the descriptor
 * (scaled sum, weighted sum, B3/S23 life), evaluated in mask order with no
 * contraction, which is how a user kernel written against Mask::get reads.
 */
#ifndef PSKEL_ORACLE_H
#define PSKEL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_POINTS 64

enum {
    ORACLE_KIND_SCALED_SUM = 1,   /* out = coeff * (((p0 + p1) + p2) + ...)          */
    ORACLE_KIND_WEIGHTED_SUM = 2, /* out = ((w0*p0 + w1*p1) + w2*p2) + ...            */
    ORACLE_KIND_LIFE = 3          /* n = sum of the points; out = n==3 || (n==2 && c==1) */
};

enum {
    ORACLE_BOUNDARY_COPY = 0, /* 3D: cells within r of a face are copied (out = in) */
    ORACLE_BOUNDARY_HALO = 1  /* 3D: every cell updated, outside reads halo_value   */
};

enum { ORACLE_STORE_F32 = 0, ORACLE_STORE_BF16 = 1, ORACLE_STORE_F16 = 2 };

/* Rounding of 16-bit stores: to nearest even, or stochastic.  Stochastic
 * rounding rounds the magnitude up with probability equal to the fraction of
 * an ulp being dropped (counted to 2^-16 for bf16, 2^-13 for f16), from the top
 * bits of a 32-bit random word r = oracle_sr_bits(seed, sweep, h, w*D + d):
 * sweep is the 0-based index of the sweep producing the value (counted from
 * the upload).  No reference counterpart (the reference has no reduced-precision
 * path); the conversions are pinned by exhaustive probability counts, closed
 * forms and recorded outputs of gfx950's v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32,
 * whose rounding they restate. */
enum { ORACLE_ROUND_NEAREST = 0, ORACLE_ROUND_STOCHASTIC = 1 };
uint32_t oracle_sr_key(uint32_t seed, uint64_t sweep);
uint32_t oracle_sr_cell_key(uint32_t seed);
uint32_t oracle_sr_plane_bits(uint32_t key, uint64_t h);
uint32_t oracle_sr_cell_bits(uint32_t cell_key, uint64_t wd);
uint32_t oracle_sr_bits(uint32_t seed, uint64_t sweep, uint64_t h, uint64_t wd);
uint16_t oracle_f32_to_bf16_sr(float x, uint32_t r);
uint16_t oracle_f32_to_f16_sr(float x, uint32_t r);

typedef struct oracle_desc {
    int32_t kind;
    int32_t npoints;
    int32_t offsets[3 * ORACLE_MAX_POINTS]; /* (dh, dw) or (dh, dw, dd) per point */
    float weights[ORACLE_MAX_POINTS];
    float coeff;
    double coeff_f64;    /* scaled-sum coefficient used by the f64 runners;
                            0 means (double)coeff                           */
    int32_t radius;      /* 0: computed like Mask::getRange()             */
    int32_t boundary;    /* 3D only                                         */
    float halo_value;    /* 3D HALO boundary value                          */
    float store_offset;  /* reduced-precision storage holds (u - offset)    */
} oracle_desc;

/* Mask::getRange(): max over points of max(|dh|,|dw|[,|dd|]). */
int32_t oracle_range(const oracle_desc* d, int ndim);

/* runIterativeCPU on 2D float grids.  `input` and `output` are updated in place
 * exactly like the reference: `input` is clobbered with intermediate states and
 * the result is left in `output`.  threads <= 0 means omp_get_max_threads(). */
void oracle_run_iterative_2d_f32(const oracle_desc* d, float* input, float* output,
                                 size_t width, size_t height, size_t iterations, int threads);
void oracle_run_iterative_2d_f64(const oracle_desc* d, double* input, double* output,
                                 size_t width, size_t height, size_t iterations, int threads);
void oracle_run_iterative_2d_u8(const oracle_desc* d, uint8_t* input, uint8_t* output,
                                size_t width, size_t height, size_t iterations, int threads);

/* runIterativeCPU on 3D grids.  residuals (may be NULL) receives the FP64 norm
 * ||u^k - u^{k-1}||_2 after every sweep k that is a multiple of residual_every
 * (0 disables); the number written is returned. */
/* CLOUDSIM (2D, f32): wind fields in the Args2D host layout, read at
 * wind[w*field_width + h]; interior cells only, as every 2D runner. */
void oracle_run_iterative_2d_cloud(float* input, float* output, size_t width, size_t height,
                                   const float* wind_x, const float* wind_y, size_t field_width, float delta_t,
                                   size_t iterations, int threads);

size_t oracle_run_iterative_3d_f32(const oracle_desc* d, float* input, float* output,
                                   size_t width, size_t height, size_t depth, size_t iterations,
                                   int threads, size_t residual_every, double* residuals);
size_t oracle_run_iterative_3d_f64(const oracle_desc* d, double* input, double* output,
                                   size_t width, size_t height, size_t depth, size_t iterations,
                                   int threads, size_t residual_every, double* residuals);
/* Reduced-precision storage (16-bit words: bf16 or f16 bit patterns) of
 * x = u - desc->store_offset, float arithmetic, round-to-nearest-even on store.
 * A scaled sum adds the stored x of its points in mask order (a HALO point out
 * of range reads halo_value - offset) and stores fma(coeff, sum, c0), with
 * c0 = oracle_scaled_sum_offset_term(coeff, npoints, offset): the offset folded
 * into one term, since coeff * sum(x_i + off) - off = coeff * sum(x_i) + c0.
 * A weighted sum decodes u = x + offset, runs the f32 expression in mask order
 * and stores (result - offset).  No reference counterpart: the reference has no
 * reduced-precision path; this is the build's definition of one, pinned by
 * error bounds against the f32 and f64 runners for any offset. */
float oracle_scaled_sum_offset_term(float coeff, int npoints, float off);
size_t oracle_run_iterative_3d_r16(const oracle_desc* d, int store, uint16_t* input,
                                   uint16_t* output, size_t width, size_t height, size_t depth,
                                   size_t iterations, int threads, size_t residual_every,
                                   double* residuals);
/* Same with a rounding mode (ORACLE_ROUND_*), seed and the index of the first sweep. */
size_t oracle_run_iterative_3d_r16_mode(const oracle_desc* d, int store, int rounding, uint32_t seed,
                                        uint64_t first_sweep, uint16_t* input, uint16_t* output,
                                        size_t width, size_t height, size_t depth, size_t iterations,
                                        int threads, size_t residual_every, double* residuals);

/* Storage conversions, round-to-nearest-even, NaN preserved. */
uint16_t oracle_f32_to_bf16(float x);
float oracle_bf16_to_f32(uint16_t h);
uint16_t oracle_f32_to_f16(float x);
float oracle_f16_to_f32(uint16_t h);

/* FP64 reference of the cost-model MLP: y = W_L * act(...act(W_1 x + b_1)...) + b_L
 * with tanh hidden activations, mean-squared-error loss over a batch.
 * params is the flat parameter vector in layer order [W_1 (out x in, row-major),
 * b_1, W_2, b_2, ...]; grads (may be NULL) receives dLoss/dparams. */
double oracle_mlp_loss_grad(const int* sizes, int nlayers, const double* params,
                            const double* x, const double* y, size_t n, double* grads);
void oracle_mlp_predict(const int* sizes, int nlayers, const double* params, const double* x,
                        size_t n, double* y);

/* Timing helper for the CPU baseline: wall seconds of the last run call. */
double oracle_last_seconds(void);
int oracle_num_procs(void);

#ifdef __cplusplus
}
#endif

#endif
