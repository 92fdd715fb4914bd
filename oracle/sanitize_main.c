/*
 * Sanitizer driver for the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Runs every oracle runner on small ragged grids with several threads, so an
 * AddressSanitizer + UndefinedBehaviorSanitizer build (gcc) catches out-of-
 * bounds accesses and undefined arithmetic, and a ThreadSanitizer build (clang,
 * LLVM OpenMP with the Archer tool) catches data races in the OpenMP loops.
 * Built and run by `make -C oracle sanitize` (tests/test_oracle_sanitizers.py).
 * Exit status 0 when every run completes; the sanitizers abort otherwise.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pskel_oracle.h"

static uint64_t g_state = 0x9E3779B97F4A7C15ull;

static double urand(void)
{
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

static oracle_desc desc(int kind, int ndim, int np, const int* offs, float coeff, int boundary)
{
    oracle_desc d;
    memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.npoints = np;
    for (int i = 0; i < np * ndim; ++i) d.offsets[i] = offs[i];
    for (int i = 0; i < np; ++i) d.weights[i] = (float)(0.5 / np + 0.01 * i);
    d.coeff = coeff;
    d.boundary = boundary;
    d.halo_value = 0.25f;
    d.store_offset = 0.5f;
    return d;
}

int main(void)
{
    const int threads = 4;
    /* 2D: Jacobi, 3x3 weighted, life, CLOUDSIM on ragged grids */
    static const int j5[8] = {-1, 0, 1, 0, 0, -1, 0, 1};
    int box9[18], moore[16];
    for (int i = 0, k = 0; i < 9; ++i) {
        box9[2 * i] = i / 3 - 1;
        box9[2 * i + 1] = i % 3 - 1;
        if (i != 4) {
            moore[2 * k] = i / 3 - 1;
            moore[2 * k + 1] = i % 3 - 1;
            ++k;
        }
    }
    const size_t sizes2[][2] = {{1, 1}, {2, 3}, {3, 3}, {17, 9}, {64, 70}, {131, 97}};
    for (size_t t = 0; t < sizeof(sizes2) / sizeof(sizes2[0]); ++t) {
        const size_t W = sizes2[t][0], H = sizes2[t][1], n = W * H;
        float* a = malloc(n * sizeof(float));
        float* b = malloc(n * sizeof(float));
        double* a64 = malloc(n * sizeof(double));
        double* b64 = malloc(n * sizeof(double));
        uint8_t* la = malloc(n);
        uint8_t* lb = malloc(n);
        float* wx = malloc(n * sizeof(float));
        float* wy = malloc(n * sizeof(float));
        for (size_t i = 0; i < n; ++i) {
            a[i] = b[i] = (float)urand();
            a64[i] = b64[i] = urand();
            la[i] = lb[i] = urand() < 0.3;
            wx[i] = (float)(urand() - 0.5);
            wy[i] = (float)(urand() - 0.5);
        }
        oracle_desc dj = desc(1, 2, 4, j5, 0.25f, 0);
        oracle_desc db = desc(2, 2, 9, box9, 0.f, 0);
        oracle_desc dl = desc(3, 2, 8, moore, 0.f, 0);
        oracle_run_iterative_2d_f32(&dj, a, b, W, H, 5, threads);
        oracle_run_iterative_2d_f32(&db, a, b, W, H, 4, threads);
        oracle_run_iterative_2d_f64(&dj, a64, b64, W, H, 3, threads);
        oracle_run_iterative_2d_u8(&dl, la, lb, W, H, 7, threads);
        /* CLOUDSIM reads wind(x = h, y = w): fields of width H and height W */
        oracle_run_iterative_2d_cloud(a, b, W, H, wx, wy, H, 0.01f, 3, threads);
        free(a), free(b), free(a64), free(b64), free(la), free(lb), free(wx), free(wy);
    }
    /* 3D: 7-point COPY and HALO, f32 / f64 / 16-bit with both roundings, residuals */
    static const int j7[18] = {-1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1};
    const size_t sizes3[][3] = {{1, 1, 1}, {2, 3, 4}, {5, 9, 13}, {17, 6, 33}};
    for (size_t t = 0; t < sizeof(sizes3) / sizeof(sizes3[0]); ++t) {
        const size_t W = sizes3[t][0], H = sizes3[t][1], D = sizes3[t][2], n = W * H * D;
        float* a = malloc(n * sizeof(float));
        float* b = malloc(n * sizeof(float));
        double* a64 = malloc(n * sizeof(double));
        double* b64 = malloc(n * sizeof(double));
        uint16_t* wa = malloc(n * sizeof(uint16_t));
        uint16_t* wb = malloc(n * sizeof(uint16_t));
        double res[8];
        for (size_t i = 0; i < n; ++i) {
            a[i] = b[i] = (float)urand();
            a64[i] = b64[i] = urand();
        }
        for (int boundary = 0; boundary < 2; ++boundary) {
            oracle_desc d = desc(1, 3, 6, j7, 1.0f / 6.0f, boundary);
            oracle_desc dw = desc(2, 3, 6, j7, 0.f, boundary);
            oracle_run_iterative_3d_f32(&d, a, b, W, H, D, 6, threads, 2, res);
            oracle_run_iterative_3d_f32(&dw, a, b, W, H, D, 3, threads, 1, res);
            oracle_run_iterative_3d_f64(&d, a64, b64, W, H, D, 5, threads, 5, res);
            for (int store = 1; store <= 2; ++store)
                for (int rounding = 0; rounding < 2; ++rounding) {
                    for (size_t i = 0; i < n; ++i)
                        wa[i] = wb[i] = store == 1 ? oracle_f32_to_bf16(a[i] - 0.5f) : oracle_f32_to_f16(a[i] - 0.5f);
                    oracle_run_iterative_3d_r16_mode(&d, store, rounding, 7u, 3, wa, wb, W, H, D, 4, threads, 2, res);
                    oracle_run_iterative_3d_r16_mode(&dw, store, rounding, 7u, 0, wa, wb, W, H, D, 3, threads, 3, res);
                }
        }
        free(a), free(b), free(a64), free(b64), free(wa), free(wb);
    }
    /* conversions over special values */
    const float specials[] = {0.f, -0.f, 1e-45f, -1e-40f, 6e-8f, 65504.f, 65520.f, 3.4e38f, 1.f / 0.f, -1.f / 0.f};
    for (size_t i = 0; i < sizeof(specials) / sizeof(specials[0]); ++i) {
        volatile uint16_t s = oracle_f32_to_bf16(specials[i]) ^ oracle_f32_to_f16(specials[i]) ^
                              oracle_f32_to_bf16_sr(specials[i], 0xdeadbeefu) ^
                              oracle_f32_to_f16_sr(specials[i], 0x12345678u);
        (void)s;
    }
    /* the MLP checker */
    const int sizes[4] = {5, 7, 6, 1};
    size_t np = 0;
    for (int l = 0; l < 3; ++l) np += (size_t)sizes[l] * sizes[l + 1] + sizes[l + 1];
    double* p = malloc(np * sizeof(double));
    double* g = malloc(np * sizeof(double));
    double x[40], y[8], yp[8];
    for (size_t i = 0; i < np; ++i) p[i] = urand() - 0.5;
    for (int i = 0; i < 40; ++i) x[i] = urand();
    for (int i = 0; i < 8; ++i) y[i] = urand();
    const double loss = oracle_mlp_loss_grad(sizes, 3, p, x, y, 8, g);
    oracle_mlp_predict(sizes, 3, p, x, 8, yp);
    free(p), free(g);
    printf("sanitize: ok (loss %.6f, %d procs)\n", loss, oracle_num_procs());
    return 0;
}
