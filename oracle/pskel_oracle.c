/*
 * pskel_oracle.c -- CPU restatement of PSkel's runIterativeCPU (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this file's library; the product never does.  Built with -ffp-contract=off so
 * every float operation rounds on its own, in the order written here.
 *
 * Reference behaviour restated (paths relative to the PSkel tree):
 *   runIterativeCPU ping-pong + even-N copy  include/PSkelStencil.hpp:880-938
 *   runIterativeSequential                   include/PSkelStencil.hpp:864-878
 *   Stencil2D::runOpenMP interior loop       include/PSkelStencil.hpp:2235-2255
 *   Stencil3D::runOpenMP all-cells loop      include/PSkelStencil.hpp:2104-2113
 *   Mask getRange / Mask3D::get halo         include/PSkelMask.hpp:186-199, 202-236, 308-334
 *   Array index maps                         include/PSkelArray.hpp:636-677
 */
#include "pskel_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static double g_last_seconds = 0.0;

double oracle_last_seconds(void) { return g_last_seconds; }
int oracle_num_procs(void) { return omp_get_num_procs(); }

static int nthreads(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

/* Mask2D::getRange / Mask3D::getRange: include/PSkelMask.hpp:202-236, 308-334 */
int32_t oracle_range(const oracle_desc* d, int ndim)
{
    if (d->radius > 0) return d->radius;
    int32_t r = 0;
    for (int i = 0; i < d->npoints; ++i) {
        for (int k = 0; k < ndim; ++k) {
            int32_t v = d->offsets[ndim * i + k];
            if (v < 0) v = -v;
            if (v > r) r = v;
        }
    }
    return r;
}

/* ------------------------------------------------------------------------ */
/* storage conversions (round to nearest even)                              */
/* ------------------------------------------------------------------------ */
static uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

uint16_t oracle_f32_to_bf16(float x)
{
    uint32_t u = f2u(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet NaN */
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float oracle_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

uint16_t oracle_f32_to_f16(float x)
{
    uint32_t u = f2u(x);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) {
        if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu));
        return (uint16_t)(sign | 0x7c00u);
    }
    if (a < 0x38800000u) { /* below the smallest f16 normal 2^-14: scale is exact */
        float s = u2f(a) * 16777216.0f; /* 2^24 */
        return (uint16_t)(sign | (uint32_t)nearbyintf(s));
    }
    uint32_t e = (a >> 23) - 127u + 15u;
    uint32_t m = a & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++; /* carry may reach inf: correct */
    return (uint16_t)(sign | h);
}

float oracle_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    if (e == 0) {
        float v = (float)m * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
    return u2f(sign | ((e - 15u + 127u) << 23) | (m << 13));
}

/* Stochastic rounding.  32-bit integer finaliser (a murmur3-style avalanche;
 * a bijection with mix32(0) = 0). */
static uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

uint32_t oracle_sr_key(uint32_t seed, uint64_t sweep)
{
    return mix32(seed + 0x9e3779b9u * ((uint32_t)sweep + 1u));
}

/* Random word of cell (h, w, d) in sweep t: A(w*D + d) ^ B(t, h), two
 * independent hash streams: A of the in-plane index under a key of the seed
 * alone (oracle_sr_cell_key), B of the plane under the sweep's key
 * (oracle_sr_key).  Two distinct (cell, sweep) pairs get pairwise independent
 * words (distinct in-plane index: independent A terms; the same one: the B
 * terms differ), so the rounding errors of a linear sweep are uncorrelated as
 * with a fresh draw per cell and sweep. */
uint32_t oracle_sr_cell_key(uint32_t seed)
{
    return mix32(mix32(seed) ^ 0xa511e9b3u);
}

uint32_t oracle_sr_plane_bits(uint32_t key, uint64_t h)
{
    return mix32(mix32(key + 0x632be5abu) ^ (uint32_t)h);
}

uint32_t oracle_sr_cell_bits(uint32_t cell_key, uint64_t wd)
{
    return mix32(cell_key ^ (uint32_t)wd);
}

uint32_t oracle_sr_bits(uint32_t seed, uint64_t sweep, uint64_t h, uint64_t wd)
{
    return oracle_sr_cell_bits(oracle_sr_cell_key(seed), wd) ^
           oracle_sr_plane_bits(oracle_sr_key(seed, sweep), h);
}

/* bf16: add the top 16 bits of r below the kept bits and truncate, so the
 * magnitude rounds up with probability = dropped fraction of an ulp (to
 * 2^-16); the carry may reach the exponent, and overflow gives inf.  NaN keeps
 * its top payload bits with the quiet bit set. */
uint16_t oracle_f32_to_bf16_sr(float x, uint32_t r)
{
    const uint32_t u = f2u(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + (r >> 16)) >> 16);
}

/* f16: the top 13 bits of r are added to the 13 dropped mantissa bits of the
 * f32 magnitude, which is then truncated (exact in f16; 65536 and above give
 * inf).  Below 2^-14 the f16 spacing is 2^-24: with s = |x| 2^24 the result is
 * floor(s) + 1 exactly when floor(frac(s) 2^13) + (r >> 19) >= 2^13, so the
 * dropped fraction counts to 13 bits as in the normal range.  NaN keeps its
 * top payload bits with the quiet bit set. */
uint16_t oracle_f32_to_f16_sr(float x, uint32_t r)
{
    const uint32_t u = f2u(x);
    const uint32_t a = u & 0x7fffffffu;
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t r13 = r >> 19;
    if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu));
    if (a == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
    if (a >= 0x38800000u) {
        const uint32_t t = (a + r13) & ~0x1fffu;
        if (t >= 0x47800000u) return (uint16_t)(sign | 0x7c00u);
        const uint32_t e = (t >> 23) - 127u + 15u;
        return (uint16_t)(sign | (e << 10) | ((t >> 13) & 0x3ffu));
    }
    const double sv = (double)u2f(a) * 16777216.0;   /* exact */
    const double fl = floor(sv);
    const uint32_t k = (uint32_t)floor((sv - fl) * 8192.0);
    return (uint16_t)(sign | (uint32_t)(fl + (k + r13 >= 8192u ? 1.0 : 0.0)));
}

/* ------------------------------------------------------------------------ */
/* 2D: Stencil2D::runOpenMP over the interior [r, N-r) (PSkelStencil.hpp:2235) */
/* ------------------------------------------------------------------------ */
#define DEFINE_SWEEP_2D_FLOAT(NAME, T, COEFF)                                                   \
    static void NAME(const oracle_desc* d, const T* in, T* out, size_t W, size_t H, size_t r,   \
                     int th)                                                                    \
    {                                                                                           \
        ptrdiff_t lin[ORACLE_MAX_POINTS];                                                       \
        T wts[ORACLE_MAX_POINTS];                                                               \
        const int np = d->npoints;                                                              \
        for (int i = 0; i < np; ++i) {                                                          \
            lin[i] = (ptrdiff_t)d->offsets[2 * i] * (ptrdiff_t)W + d->offsets[2 * i + 1];       \
            wts[i] = (T)d->weights[i];                                                          \
        }                                                                                       \
        const T c = (COEFF);                                                                    \
        const int kind = d->kind;                                                               \
        if (H < 2 * r || W < 2 * r) return;                                                     \
        _Pragma("omp parallel for num_threads(th) schedule(static)")                            \
        for (size_t h = r; h < H - r; ++h) {                                                    \
            for (size_t w = r; w < W - r; ++w) {                                                \
                const T* p = in + h * W + w;                                                    \
                T acc;                                                                          \
                if (kind == ORACLE_KIND_SCALED_SUM) {                                           \
                    acc = p[lin[0]];                                                            \
                    for (int i = 1; i < np; ++i) acc = acc + p[lin[i]];                         \
                    acc = c * acc;                                                              \
                } else {                                                                        \
                    acc = wts[0] * p[lin[0]];                                                   \
                    for (int i = 1; i < np; ++i) acc = acc + wts[i] * p[lin[i]];                \
                }                                                                               \
                out[h * W + w] = acc;                                                           \
            }                                                                                   \
        }                                                                                       \
    }

DEFINE_SWEEP_2D_FLOAT(sweep2d_f32, float, d->coeff)
DEFINE_SWEEP_2D_FLOAT(sweep2d_f64, double, (d->coeff_f64 != 0.0 ? d->coeff_f64 : (double)d->coeff))

static void sweep2d_u8(const oracle_desc* d, const uint8_t* in, uint8_t* out, size_t W, size_t H,
                       size_t r, int th)
{
    ptrdiff_t lin[ORACLE_MAX_POINTS];
    const int np = d->npoints;
    for (int i = 0; i < np; ++i)
        lin[i] = (ptrdiff_t)d->offsets[2 * i] * (ptrdiff_t)W + d->offsets[2 * i + 1];
    if (H < 2 * r || W < 2 * r) return;
#pragma omp parallel for num_threads(th) schedule(static)
    for (size_t h = r; h < H - r; ++h) {
        for (size_t w = r; w < W - r; ++w) {
            const uint8_t* p = in + h * W + w;
            int n = 0;
            for (int i = 0; i < np; ++i) n += p[lin[i]];
            /* B3/S23, the rule of the reference's GOL_KERNEL (PSkelStencil.hpp:439-440) */
            out[h * W + w] = (uint8_t)((n == 3 || (n == 2 && p[0] == 1)) ? 1 : 0);
        }
    }
}

/* runIterativeCPU: include/PSkelStencil.hpp:880-938 */
#define DEFINE_RUN_2D(NAME, T, SWEEP)                                                           \
    void NAME(const oracle_desc* d, T* input, T* output, size_t width, size_t height,           \
              size_t iterations, int threads)                                                   \
    {                                                                                           \
        const int th = nthreads(threads);                                                       \
        const size_t r = (size_t)oracle_range(d, 2);                                            \
        double t0 = omp_get_wtime();                                                            \
        for (size_t it = 0; it < iterations; ++it) {                                            \
            if (it % 2 == 0) SWEEP(d, input, output, width, height, r, th);                     \
            else SWEEP(d, output, input, width, height, r, th);                                 \
        }                                                                                       \
        if (iterations % 2 == 0) memcpy(output, input, width * height * sizeof(T));             \
        g_last_seconds = omp_get_wtime() - t0;                                                  \
    }

DEFINE_RUN_2D(oracle_run_iterative_2d_f32, float, sweep2d_f32)
DEFINE_RUN_2D(oracle_run_iterative_2d_f64, double, sweep2d_f64)
DEFINE_RUN_2D(oracle_run_iterative_2d_u8, uint8_t, sweep2d_u8)

/* ------------------------------------------------------------------------ */
/* CLOUDSIM: the wind advection + conduction formula of the reference's       */
/* CLOUDSIM_KERNEL (include/PSkelStencil.hpp:349-387), evaluated on the       */
/* global neighbours of every interior cell in the same order.  The wind is   */
/* read as args.wind_x(h, w), i.e. through Args2D's map (x, y) -> y*width + x  */
/* (include/PSkelArgs.hpp:116-123) with x = h and y = w: wind[w*fw + h].      */
/* ------------------------------------------------------------------------ */
static void sweep2d_cloud(const float* in, float* out, size_t W, size_t H, const float* wx, const float* wy,
                          size_t fw, float dt, int th)
{
    if (H < 2 || W < 2) return;
#pragma omp parallel for num_threads(th) schedule(static)
    for (size_t h = 1; h < H - 1; ++h) {
        for (size_t w = 1; w < W - 1; ++w) {
            const float c = in[h * W + w];
            const float n = in[(h - 1) * W + w];
            const float e = in[h * W + w + 1];
            const float ww = in[h * W + w - 1];
            const float s = in[(h + 1) * W + w];
            const float sum = 4 * c - (n + e + ww + s);
            const float xwind = wx[w * fw + h];
            const float ywind = wy[w * fw + h];
            const int xfactor = (xwind > 0) ? 1 : -1;
            const int yfactor = (ywind > 0) ? 1 : -1;
            const float comp_y = yfactor * ywind;
            const float comp_x = xfactor * xwind;
            const float nb_x = in[h * W + (size_t)((ptrdiff_t)w + xfactor)];
            const float nb_y = in[(size_t)((ptrdiff_t)h + yfactor) * W + w];
            const float temp_wind = (-comp_x * ((c - nb_x) * 10.0f)) - (comp_y * ((c - nb_y) * 10.0f));
            const float conduction = -0.0243f * (sum * 0.25f) * dt;
            const float result = c + conduction;
            out[h * W + w] = result + temp_wind * dt;
        }
    }
}

void oracle_run_iterative_2d_cloud(float* input, float* output, size_t width, size_t height,
                                   const float* wind_x, const float* wind_y, size_t field_width, float delta_t,
                                   size_t iterations, int threads)
{
    const int th = nthreads(threads);
    double t0 = omp_get_wtime();
    for (size_t it = 0; it < iterations; ++it) {
        if (it % 2 == 0) sweep2d_cloud(input, output, width, height, wind_x, wind_y, field_width, delta_t, th);
        else sweep2d_cloud(output, input, width, height, wind_x, wind_y, field_width, delta_t, th);
    }
    if (iterations % 2 == 0) memcpy(output, input, width * height * sizeof(float));
    g_last_seconds = omp_get_wtime() - t0;
}

/* ------------------------------------------------------------------------ */
/* 3D: Stencil3D::runOpenMP visits every cell (PSkelStencil.hpp:2104-2113)   */
/* ------------------------------------------------------------------------ */
typedef struct {
    ptrdiff_t lin[ORACLE_MAX_POINTS];
    int off[3 * ORACLE_MAX_POINTS];
    float wf[ORACLE_MAX_POINTS];
    double wd[ORACLE_MAX_POINTS];
    int np, kind, boundary;
    size_t r;
    float cf;
    double cd;
} plan3d;

static void plan3d_init(plan3d* p, const oracle_desc* d, size_t W, size_t D)
{
    p->np = d->npoints;
    p->kind = d->kind;
    p->boundary = d->boundary;
    p->r = (size_t)oracle_range(d, 3);
    p->cf = d->coeff;
    p->cd = d->coeff_f64 != 0.0 ? d->coeff_f64 : (double)d->coeff;
    for (int i = 0; i < p->np; ++i) {
        p->off[3 * i] = d->offsets[3 * i];
        p->off[3 * i + 1] = d->offsets[3 * i + 1];
        p->off[3 * i + 2] = d->offsets[3 * i + 2];
        p->lin[i] = ((ptrdiff_t)d->offsets[3 * i] * (ptrdiff_t)W + d->offsets[3 * i + 1]) *
                        (ptrdiff_t)D +
                    d->offsets[3 * i + 2];
        p->wf[i] = d->weights[i];
        p->wd[i] = (double)d->weights[i];
    }
}

/* One cell of the stencil with out-of-range neighbours reading `halo`
 * (Mask3D::get, include/PSkelMask.hpp:186-199). */
#define CELL3D_HALO(T, WT, C, p, in, h, w, dd, W, H, D, halo, res)                              \
    do {                                                                                        \
        T acc_ = 0;                                                                             \
        for (int i_ = 0; i_ < (p)->np; ++i_) {                                                  \
            size_t hh_ = (size_t)((ptrdiff_t)(h) + (p)->off[3 * i_]);                           \
            size_t ww_ = (size_t)((ptrdiff_t)(w) + (p)->off[3 * i_ + 1]);                       \
            size_t de_ = (size_t)((ptrdiff_t)(dd) + (p)->off[3 * i_ + 2]);                      \
            T v_ = (ww_ < (W) && hh_ < (H) && de_ < (D)) ? (in)[(hh_ * (W) + ww_) * (D) + de_]  \
                                                          : (halo);                             \
            if ((p)->kind == ORACLE_KIND_SCALED_SUM) acc_ = (i_ == 0) ? v_ : acc_ + v_;         \
            else acc_ = (i_ == 0) ? (p)->WT[i_] * v_ : acc_ + (p)->WT[i_] * v_;                 \
        }                                                                                       \
        if ((p)->kind == ORACLE_KIND_SCALED_SUM) acc_ = (C) * acc_;                             \
        (res) = acc_;                                                                           \
    } while (0)

static double sweep3d_f32(const plan3d* p, const float* in, float* out, size_t W, size_t H,
                          size_t D, int th, int want_res, float halo)
{
    const size_t r = p->r;
    double* rows = want_res ? (double*)calloc(H, sizeof(double)) : NULL;
#pragma omp parallel for num_threads(th) schedule(static)
    for (size_t h = 0; h < H; ++h) {
        double rs = 0.0;
        for (size_t w = 0; w < W; ++w) {
            const size_t base = (h * W + w) * D;
            for (size_t dd = 0; dd < D; ++dd) {
                float v;
                int edge = (h < r || h >= H - r || w < r || w >= W - r || dd < r || dd >= D - r);
                if (edge && p->boundary == ORACLE_BOUNDARY_COPY) {
                    v = in[base + dd];
                } else if (edge) {
                    CELL3D_HALO(float, wf, p->cf, p, in, h, w, dd, W, H, D, halo, v);
                } else {
                    const float* q = in + base + dd;
                    float acc;
                    if (p->kind == ORACLE_KIND_SCALED_SUM) {
                        acc = q[p->lin[0]];
                        for (int i = 1; i < p->np; ++i) acc = acc + q[p->lin[i]];
                        acc = p->cf * acc;
                    } else {
                        acc = p->wf[0] * q[p->lin[0]];
                        for (int i = 1; i < p->np; ++i) acc = acc + p->wf[i] * q[p->lin[i]];
                    }
                    v = acc;
                }
                if (want_res) {
                    double dv = (double)v - (double)in[base + dd];
                    rs += dv * dv;
                }
                out[base + dd] = v;
            }
        }
        if (want_res) rows[h] = rs;
    }
    double s = 0.0;
    if (want_res) {
        for (size_t h = 0; h < H; ++h) s += rows[h];
        free(rows);
    }
    return sqrt(s);
}

static double sweep3d_f64(const plan3d* p, const double* in, double* out, size_t W, size_t H,
                          size_t D, int th, int want_res, double halo)
{
    const size_t r = p->r;
    double* rows = want_res ? (double*)calloc(H, sizeof(double)) : NULL;
#pragma omp parallel for num_threads(th) schedule(static)
    for (size_t h = 0; h < H; ++h) {
        double rs = 0.0;
        for (size_t w = 0; w < W; ++w) {
            const size_t base = (h * W + w) * D;
            for (size_t dd = 0; dd < D; ++dd) {
                double v;
                int edge = (h < r || h >= H - r || w < r || w >= W - r || dd < r || dd >= D - r);
                if (edge && p->boundary == ORACLE_BOUNDARY_COPY) {
                    v = in[base + dd];
                } else if (edge) {
                    CELL3D_HALO(double, wd, p->cd, p, in, h, w, dd, W, H, D, halo, v);
                } else {
                    const double* q = in + base + dd;
                    double acc;
                    if (p->kind == ORACLE_KIND_SCALED_SUM) {
                        acc = q[p->lin[0]];
                        for (int i = 1; i < p->np; ++i) acc = acc + q[p->lin[i]];
                        acc = p->cd * acc;
                    } else {
                        acc = p->wd[0] * q[p->lin[0]];
                        for (int i = 1; i < p->np; ++i) acc = acc + p->wd[i] * q[p->lin[i]];
                    }
                    v = acc;
                }
                if (want_res) {
                    double dv = v - in[base + dd];
                    rs += dv * dv;
                }
                out[base + dd] = v;
            }
        }
        if (want_res) rows[h] = rs;
    }
    double s = 0.0;
    if (want_res) {
        for (size_t h = 0; h < H; ++h) s += rows[h];
        free(rows);
    }
    return sqrt(s);
}

typedef float (*dec_fn)(uint16_t);
typedef uint16_t (*enc_fn)(float);

/* Reduced-precision storage of x = u - off.
 *   scaled sum:   the stored values of the points (an out-of-range point of a
 *                 HALO mask reads halo - off) are summed in mask order, then
 *                 v = fma(coeff, sum, c0) with c0 = (coeff * n - 1) * off
 *                 evaluated in double and rounded once to float: since
 *                 coeff * sum(x_i + off) - off = coeff * sum(x_i) + c0, v is
 *                 the new stored value, rounded to the storage format;
 *   weighted sum: each point decodes to u = x + off, the f32 expression runs
 *                 in mask order, and (result - off) is rounded to storage. */
float oracle_scaled_sum_offset_term(float coeff, int npoints, float off)
{
    return (float)(((double)coeff * (double)npoints - 1.0) * (double)off);
}

static double sweep3d_r16(const plan3d* p, int store, float off, const uint16_t* in,
                          uint16_t* out, size_t W, size_t H, size_t D, int th, int want_res,
                          float halo, int rounding, uint32_t seed, uint64_t sweep)
{
    const dec_fn dec = store == ORACLE_STORE_BF16 ? oracle_bf16_to_f32 : oracle_f16_to_f32;
    const enc_fn enc = store == ORACLE_STORE_BF16 ? oracle_f32_to_bf16 : oracle_f32_to_f16;
    const size_t r = p->r;
    const uint32_t key = oracle_sr_key(seed, sweep);
    const uint32_t cell_key = oracle_sr_cell_key(seed);
    const int scaled = p->kind == ORACLE_KIND_SCALED_SUM;
    const float c0 = oracle_scaled_sum_offset_term(p->cf, p->np, off);
    const float hx = scaled ? halo - off : halo;
    double* rows = want_res ? (double*)calloc(H, sizeof(double)) : NULL;
#pragma omp parallel for num_threads(th) schedule(static)
    for (size_t h = 0; h < H; ++h) {
        double rs = 0.0;
        const uint32_t plane_bits = oracle_sr_plane_bits(key, h);
        for (size_t w = 0; w < W; ++w) {
            const size_t base = (h * W + w) * D;
            for (size_t dd = 0; dd < D; ++dd) {
                uint16_t sv;
                int edge = (h < r || h >= H - r || w < r || w >= W - r || dd < r || dd >= D - r);
                if (edge && p->boundary == ORACLE_BOUNDARY_COPY) {
                    sv = in[base + dd];
                } else {
                    float acc = 0.0f;
                    for (int i = 0; i < p->np; ++i) {
                        size_t hh = (size_t)((ptrdiff_t)h + p->off[3 * i]);
                        size_t ww = (size_t)((ptrdiff_t)w + p->off[3 * i + 1]);
                        size_t de = (size_t)((ptrdiff_t)dd + p->off[3 * i + 2]);
                        float v = hx;
                        if (ww < W && hh < H && de < D) {
                            v = dec(in[(hh * W + ww) * D + de]);
                            if (!scaled) v = v + off;
                        }
                        if (scaled) acc = (i == 0) ? v : acc + v;
                        else acc = (i == 0) ? p->wf[i] * v : acc + p->wf[i] * v;
                    }
                    /* the value to store: the new x (scaled) or u - off (weighted) */
                    const float x = scaled ? fmaf(p->cf, acc, c0) : acc - off;
                    if (rounding == ORACLE_ROUND_STOCHASTIC) {
                        const uint32_t rb = oracle_sr_cell_bits(cell_key, w * D + dd) ^ plane_bits;
                        sv = store == ORACLE_STORE_BF16 ? oracle_f32_to_bf16_sr(x, rb)
                                                        : oracle_f32_to_f16_sr(x, rb);
                    } else {
                        sv = enc(x);
                    }
                }
                if (want_res) {
                    double dv = (double)dec(sv) - (double)dec(in[base + dd]);
                    rs += dv * dv;
                }
                out[base + dd] = sv;
            }
        }
        if (want_res) rows[h] = rs;
    }
    double s = 0.0;
    if (want_res) {
        for (size_t h = 0; h < H; ++h) s += rows[h];
        free(rows);
    }
    return sqrt(s);
}

size_t oracle_run_iterative_3d_f32(const oracle_desc* d, float* input, float* output,
                                   size_t width, size_t height, size_t depth, size_t iterations,
                                   int threads, size_t residual_every, double* residuals)
{
    plan3d p;
    plan3d_init(&p, d, width, depth);
    const int th = nthreads(threads);
    size_t nres = 0;
    double t0 = omp_get_wtime();
    for (size_t it = 0; it < iterations; ++it) {
        int want = residual_every && ((it + 1) % residual_every == 0);
        double rn;
        if (it % 2 == 0) rn = sweep3d_f32(&p, input, output, width, height, depth, th, want, d->halo_value);
        else rn = sweep3d_f32(&p, output, input, width, height, depth, th, want, d->halo_value);
        if (want && residuals) residuals[nres] = rn;
        if (want) nres++;
    }
    if (iterations % 2 == 0) memcpy(output, input, width * height * depth * sizeof(float));
    g_last_seconds = omp_get_wtime() - t0;
    return nres;
}

size_t oracle_run_iterative_3d_f64(const oracle_desc* d, double* input, double* output,
                                   size_t width, size_t height, size_t depth, size_t iterations,
                                   int threads, size_t residual_every, double* residuals)
{
    plan3d p;
    plan3d_init(&p, d, width, depth);
    const int th = nthreads(threads);
    size_t nres = 0;
    double t0 = omp_get_wtime();
    for (size_t it = 0; it < iterations; ++it) {
        int want = residual_every && ((it + 1) % residual_every == 0);
        double rn;
        if (it % 2 == 0) rn = sweep3d_f64(&p, input, output, width, height, depth, th, want, d->halo_value);
        else rn = sweep3d_f64(&p, output, input, width, height, depth, th, want, d->halo_value);
        if (want && residuals) residuals[nres] = rn;
        if (want) nres++;
    }
    if (iterations % 2 == 0) memcpy(output, input, width * height * depth * sizeof(double));
    g_last_seconds = omp_get_wtime() - t0;
    return nres;
}

size_t oracle_run_iterative_3d_r16_mode(const oracle_desc* d, int store, int rounding, uint32_t seed,
                                        uint64_t first_sweep, uint16_t* input, uint16_t* output,
                                        size_t width, size_t height, size_t depth, size_t iterations,
                                        int threads, size_t residual_every, double* residuals)
{
    plan3d p;
    plan3d_init(&p, d, width, depth);
    const int th = nthreads(threads);
    size_t nres = 0;
    double t0 = omp_get_wtime();
    for (size_t it = 0; it < iterations; ++it) {
        int want = residual_every && ((it + 1) % residual_every == 0);
        double rn;
        const uint64_t sw = first_sweep + it;
        if (it % 2 == 0)
            rn = sweep3d_r16(&p, store, d->store_offset, input, output, width, height, depth, th, want,
                             d->halo_value, rounding, seed, sw);
        else
            rn = sweep3d_r16(&p, store, d->store_offset, output, input, width, height, depth, th, want,
                             d->halo_value, rounding, seed, sw);
        if (want && residuals) residuals[nres] = rn;
        if (want) nres++;
    }
    if (iterations % 2 == 0) memcpy(output, input, width * height * depth * sizeof(uint16_t));
    g_last_seconds = omp_get_wtime() - t0;
    return nres;
}

size_t oracle_run_iterative_3d_r16(const oracle_desc* d, int store, uint16_t* input,
                                   uint16_t* output, size_t width, size_t height, size_t depth,
                                   size_t iterations, int threads, size_t residual_every,
                                   double* residuals)
{
    return oracle_run_iterative_3d_r16_mode(d, store, ORACLE_ROUND_NEAREST, 0, 0, input, output, width, height,
                                            depth, iterations, threads, residual_every, residuals);
}

/* ------------------------------------------------------------------------ */
/* FP64 MLP (the cost model's checker)                                      */
/* ------------------------------------------------------------------------ */
static size_t mlp_nparams(const int* sizes, int nlayers)
{
    size_t n = 0;
    for (int l = 0; l < nlayers; ++l) n += (size_t)sizes[l + 1] * sizes[l] + sizes[l + 1];
    return n;
}

/* Forward pass for one sample; acts[l] holds layer l's output (acts[0] = x). */
static void mlp_forward1(const int* sizes, int nlayers, const double* params, const double* x,
                         double** acts)
{
    memcpy(acts[0], x, sizeof(double) * sizes[0]);
    const double* p = params;
    for (int l = 0; l < nlayers; ++l) {
        const int ni = sizes[l], no = sizes[l + 1];
        const double* Wm = p;
        const double* b = p + (size_t)no * ni;
        for (int o = 0; o < no; ++o) {
            double z = b[o];
            for (int i = 0; i < ni; ++i) z += Wm[(size_t)o * ni + i] * acts[l][i];
            acts[l + 1][o] = (l + 1 < nlayers) ? tanh(z) : z;
        }
        p += (size_t)no * ni + no;
    }
}

static double** mlp_alloc_acts(const int* sizes, int nlayers)
{
    double** a = (double**)malloc(sizeof(double*) * (nlayers + 1));
    for (int l = 0; l <= nlayers; ++l) a[l] = (double*)calloc(sizes[l], sizeof(double));
    return a;
}

static void mlp_free_acts(double** a, int nlayers)
{
    for (int l = 0; l <= nlayers; ++l) free(a[l]);
    free(a);
}

void oracle_mlp_predict(const int* sizes, int nlayers, const double* params, const double* x,
                        size_t n, double* y)
{
    double** acts = mlp_alloc_acts(sizes, nlayers);
    for (size_t s = 0; s < n; ++s) {
        mlp_forward1(sizes, nlayers, params, x + s * sizes[0], acts);
        y[s] = acts[nlayers][0];
    }
    mlp_free_acts(acts, nlayers);
}

/* loss = mean over samples of (y_hat - y)^2 (one output). */
double oracle_mlp_loss_grad(const int* sizes, int nlayers, const double* params,
                            const double* x, const double* y, size_t n, double* grads)
{
    const size_t np = mlp_nparams(sizes, nlayers);
    if (grads) memset(grads, 0, sizeof(double) * np);
    double** acts = mlp_alloc_acts(sizes, nlayers);
    double** dl = mlp_alloc_acts(sizes, nlayers);
    size_t* poff = (size_t*)malloc(sizeof(size_t) * nlayers);
    {
        size_t o = 0;
        for (int l = 0; l < nlayers; ++l) { poff[l] = o; o += (size_t)sizes[l + 1] * sizes[l] + sizes[l + 1]; }
    }
    double loss = 0.0;
    for (size_t s = 0; s < n; ++s) {
        mlp_forward1(sizes, nlayers, params, x + s * sizes[0], acts);
        double e = acts[nlayers][0] - y[s];
        loss += e * e;
        if (!grads) continue;
        dl[nlayers][0] = 2.0 * e / (double)n; /* dLoss/dz of the linear output */
        for (int l = nlayers - 1; l >= 0; --l) {
            const int ni = sizes[l], no = sizes[l + 1];
            const double* Wm = params + poff[l];
            double* gW = grads + poff[l];
            double* gb = gW + (size_t)no * ni;
            for (int o = 0; o < no; ++o) {
                const double g = dl[l + 1][o];
                gb[o] += g;
                for (int i = 0; i < ni; ++i) gW[(size_t)o * ni + i] += g * acts[l][i];
            }
            if (l > 0) {
                for (int i = 0; i < ni; ++i) {
                    double a = 0.0;
                    for (int o = 0; o < no; ++o) a += Wm[(size_t)o * ni + i] * dl[l + 1][o];
                    const double t = acts[l][i]; /* tanh output */
                    dl[l][i] = a * (1.0 - t * t);
                }
            }
        }
    }
    free(poff);
    mlp_free_acts(acts, nlayers);
    mlp_free_acts(dl, nlayers);
    return loss / (double)n;
}
