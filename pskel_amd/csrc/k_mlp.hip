// Cost-model MLP: forward, backward and Adam on the device, plus the C ABI.
//
// y = W_L . tanh(... tanh(W_1 x + b_1) ...) + b_L, one output (predicted log
// time per sweep).  Parameters live on the device as one flat f32 vector in
// layer order [W_1 (out x in, row-major), b_1, W_2, b_2, ...] -- the layout
// the CPU checker uses.  Inputs are standardised with per-feature mean/std and
// the target with (y_mean, y_std); both are part of the model.
//
// The layers are tiny (tens of units), so each kernel gives one lane to one
// output element and loops over the reduction in a fixed order: training is
// deterministic for a given seed.  All arithmetic is f32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "pskel_internal.hpp"

using namespace pskel;

namespace {

constexpr int kMaxLayers = 8;

// Z[s][o] = b[o] + sum_i W[o][i] * X[s][i];  A = tanh(Z) unless last layer.
__global__ void k_mlp_forward(const float* __restrict__ X, int nin, const float* __restrict__ Wb, int nout,
                              int n, int act, float* __restrict__ A)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nout) return;
    const int s = t / nout, o = t % nout;
    const float* w = Wb + (size_t)o * nin;
    const float* x = X + (size_t)s * nin;
    float z = Wb[(size_t)nout * nin + o];
    for (int i = 0; i < nin; ++i) z += w[i] * x[i];
    A[t] = act ? tanhf(z) : z;
}

// Standardise: Xn[s][i] = (X[s][i] - mean[i]) / std[i]
__global__ void k_mlp_standardise(const float* __restrict__ X, const float* __restrict__ mean,
                                  const float* __restrict__ sd, int nin, int n, float* __restrict__ Xn)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nin) return;
    const int i = t % nin;
    Xn[t] = (X[t] - mean[i]) / sd[i];
}

// Gather rows of X (nin) and y by index into contiguous batch buffers.
__global__ void k_mlp_gather(const float* __restrict__ X, const float* __restrict__ y,
                             const int* __restrict__ idx, int nin, int nb, float* __restrict__ Xb,
                             float* __restrict__ yb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * nin) return;
    const int s = t / nin, i = t % nin;
    Xb[t] = X[(size_t)idx[s] * nin + i];
    if (i == 0) yb[s] = y[idx[s]];
}

// dZ_out[s] = 2 (yhat[s] - y[s]) / n ; loss partial = sum (yhat - y)^2 (one block, fixed order)
__global__ void k_mlp_loss(const float* __restrict__ yhat, const float* __restrict__ y, int n,
                           float* __restrict__ dz, float* __restrict__ loss)
{
    __shared__ float red[256];
    float acc = 0.f;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        const float e = yhat[s] - y[s];
        dz[s] = 2.f * e / (float)n;
        acc += e * e;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / (float)n;
}

// gW[o][i] = sum_s dZ[s][o] * Ain[s][i];  gb[o] = sum_s dZ[s][o]
__global__ void k_mlp_grad_w(const float* __restrict__ dZ, const float* __restrict__ Ain, int nin, int nout,
                             int n, float* __restrict__ g)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nout * (nin + 1)) return;
    if (t < nout * nin) {
        const int o = t / nin, i = t % nin;
        float acc = 0.f;
        for (int s = 0; s < n; ++s) acc += dZ[(size_t)s * nout + o] * Ain[(size_t)s * nin + i];
        g[t] = acc;
    } else {
        const int o = t - nout * nin;
        float acc = 0.f;
        for (int s = 0; s < n; ++s) acc += dZ[(size_t)s * nout + o];
        g[t] = acc;
    }
}

// dZprev[s][i] = (sum_o dZ[s][o] * W[o][i]) * (1 - Aprev[s][i]^2)   (tanh')
__global__ void k_mlp_back(const float* __restrict__ dZ, const float* __restrict__ W, const float* __restrict__ Aprev,
                           int nin, int nout, int n, float* __restrict__ dZprev)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nin) return;
    const int s = t / nin, i = t % nin;
    float acc = 0.f;
    for (int o = 0; o < nout; ++o) acc += dZ[(size_t)s * nout + o] * W[(size_t)o * nin + i];
    const float a = Aprev[t];
    dZprev[t] = acc * (1.f - a * a);
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int np, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct pskel_mlp {
    int nl = 0;                 // number of layers
    int sizes[kMaxLayers + 1] = {0};
    size_t np = 0;
    size_t off[kMaxLayers] = {0};  // parameter offset of layer l
    int device = 0;
    hipStream_t stream = nullptr;
    float* d_p = nullptr;       // parameters
    float* d_g = nullptr;       // gradients
    float* d_m = nullptr;       // Adam moments
    float* d_v = nullptr;
    std::vector<float> mean, sd;  // input standardisation
    float y_mean = 0.f, y_sd = 1.f;
    float* d_mean = nullptr;
    float* d_sd = nullptr;
    // activation workspace for up to cap samples
    size_t cap = 0;
    float* d_x = nullptr;       // standardised inputs
    float* d_act[kMaxLayers] = {nullptr};  // layer outputs
    float* d_dz[kMaxLayers] = {nullptr};   // layer output gradients
    float* d_y = nullptr;
    float* d_loss = nullptr;
    int64_t step = 0;
};

namespace {

pskel_status_t ensure_cap(pskel_mlp_t* m, size_t n)
{
    if (n <= m->cap) return PSKEL_OK;
    auto fr = [](float*& p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    fr(m->d_x);
    fr(m->d_y);
    for (int l = 0; l < m->nl; ++l) {
        fr(m->d_act[l]);
        fr(m->d_dz[l]);
    }
    m->cap = 0;
    PSKEL_HIP(hipMalloc(&m->d_x, n * m->sizes[0] * sizeof(float)));
    PSKEL_HIP(hipMalloc(&m->d_y, n * sizeof(float)));
    for (int l = 0; l < m->nl; ++l) {
        PSKEL_HIP(hipMalloc(&m->d_act[l], n * m->sizes[l + 1] * sizeof(float)));
        PSKEL_HIP(hipMalloc(&m->d_dz[l], n * m->sizes[l + 1] * sizeof(float)));
    }
    m->cap = n;
    return PSKEL_OK;
}

pskel_status_t upload_norm(pskel_mlp_t* m)
{
    PSKEL_HIP(hipMemcpy(m->d_mean, m->mean.data(), m->mean.size() * sizeof(float), hipMemcpyHostToDevice));
    PSKEL_HIP(hipMemcpy(m->d_sd, m->sd.data(), m->sd.size() * sizeof(float), hipMemcpyHostToDevice));
    return PSKEL_OK;
}

// forward on n samples already standardised in d_x; output in d_act[nl-1]
hipError_t forward(pskel_mlp_t* m, int n, const float* x_std)
{
    const float* in = x_std;
    for (int l = 0; l < m->nl; ++l) {
        const int ni = m->sizes[l], no = m->sizes[l + 1];
        hipLaunchKernelGGL(k_mlp_forward, dim3(nblk((size_t)n * no)), dim3(256), 0, m->stream, in, ni,
                           m->d_p + m->off[l], no, n, l + 1 < m->nl ? 1 : 0, m->d_act[l]);
        in = m->d_act[l];
    }
    return hipGetLastError();
}

// backward from d_dz[nl-1] (already holding dLoss/dyhat); grads into d_g
hipError_t backward(pskel_mlp_t* m, int n, const float* x_std)
{
    for (int l = m->nl - 1; l >= 0; --l) {
        const int ni = m->sizes[l], no = m->sizes[l + 1];
        const float* ain = l > 0 ? m->d_act[l - 1] : x_std;
        hipLaunchKernelGGL(k_mlp_grad_w, dim3(nblk((size_t)no * (ni + 1))), dim3(256), 0, m->stream, m->d_dz[l],
                           ain, ni, no, n, m->d_g + m->off[l]);
        if (l > 0)
            hipLaunchKernelGGL(k_mlp_back, dim3(nblk((size_t)n * ni)), dim3(256), 0, m->stream, m->d_dz[l],
                               m->d_p + m->off[l], m->d_act[l - 1], ni, no, n, m->d_dz[l - 1]);
    }
    return hipGetLastError();
}

}  // namespace

namespace pskel {
pskel_status_t mlp_predict_host(const pskel_mlp_t* mc, const float* x, size_t n, float* y);
int mlp_inputs(const pskel_mlp_t* m) { return m ? m->sizes[0] : 0; }
}

extern "C" {

pskel_status_t pskel_mlp_create(const int32_t* sizes, int32_t nlayers, uint64_t seed, int32_t device,
                                pskel_mlp_t** out)
{
    if (!sizes || !out || nlayers < 1 || nlayers > kMaxLayers) return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad layer spec");
    if (sizes[nlayers] != 1) return fail(PSKEL_ERR_INVALID_ARGUMENT, "the model has one output");
    for (int l = 0; l <= nlayers; ++l)
        if (sizes[l] < 1 || sizes[l] > 4096) return fail(PSKEL_ERR_INVALID_ARGUMENT, "layer width");
    *out = nullptr;
    PSKEL_HIP(hipSetDevice(device));
    auto* m = new pskel_mlp_t();
    m->nl = nlayers;
    m->device = device;
    for (int l = 0; l <= nlayers; ++l) m->sizes[l] = sizes[l];
    for (int l = 0; l < nlayers; ++l) {
        m->off[l] = m->np;
        m->np += (size_t)sizes[l + 1] * sizes[l] + sizes[l + 1];
    }
    m->mean.assign(sizes[0], 0.f);
    m->sd.assign(sizes[0], 1.f);
    // Glorot-uniform weights, zero biases (host RNG: reproducible per seed)
    std::vector<float> p(m->np, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < nlayers; ++l) {
        const float a = std::sqrt(6.f / (float)(sizes[l] + sizes[l + 1]));
        std::uniform_real_distribution<float> u(-a, a);
        for (size_t k = 0; k < (size_t)sizes[l + 1] * sizes[l]; ++k) p[m->off[l] + k] = u(rng);
    }
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&m->d_p, m->np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->d_g, m->np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->d_m, m->np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->d_v, m->np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->d_mean, sizes[0] * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->d_sd, sizes[0] * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->d_loss, sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(m->d_p, p.data(), m->np * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->d_m, 0, m->np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->d_v, 0, m->np * sizeof(float));
    if (e != hipSuccess) {
        pskel_status_t st = hip_fail(e, "mlp allocation");
        pskel_mlp_destroy(m);
        return st;
    }
    pskel_status_t st = upload_norm(m);
    if (st != PSKEL_OK) {
        pskel_mlp_destroy(m);
        return st;
    }
    *out = m;
    return PSKEL_OK;
}

pskel_status_t pskel_mlp_destroy(pskel_mlp_t* m)
{
    if (!m) return PSKEL_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    float* ps[] = {m->d_p, m->d_g, m->d_m, m->d_v, m->d_mean, m->d_sd, m->d_x, m->d_y, m->d_loss};
    for (float* p : ps)
        if (p) (void)hipFree(p);
    for (int l = 0; l < m->nl; ++l) {
        if (m->d_act[l]) (void)hipFree(m->d_act[l]);
        if (m->d_dz[l]) (void)hipFree(m->d_dz[l]);
    }
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return PSKEL_OK;
}

pskel_status_t pskel_mlp_num_params(const pskel_mlp_t* m, size_t* n)
{
    if (!m || !n) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    *n = m->np;
    return PSKEL_OK;
}

pskel_status_t pskel_mlp_get_params(const pskel_mlp_t* m, float* p, size_t n)
{
    if (!m || !p || n != m->np) return fail(PSKEL_ERR_INVALID_ARGUMENT, "parameter count");
    PSKEL_HIP(hipSetDevice(m->device));
    PSKEL_HIP(hipStreamSynchronize(m->stream));
    PSKEL_HIP(hipMemcpy(p, m->d_p, n * sizeof(float), hipMemcpyDeviceToHost));
    return PSKEL_OK;
}

pskel_status_t pskel_mlp_set_params(pskel_mlp_t* m, const float* p, size_t n)
{
    if (!m || !p || n != m->np) return fail(PSKEL_ERR_INVALID_ARGUMENT, "parameter count");
    PSKEL_HIP(hipSetDevice(m->device));
    PSKEL_HIP(hipMemcpy(m->d_p, p, n * sizeof(float), hipMemcpyHostToDevice));
    return PSKEL_OK;
}

pskel_status_t pskel_mlp_set_normalization(pskel_mlp_t* m, const float* mean, const float* sd, float y_mean,
                                           float y_sd)
{
    if (!m || !mean || !sd || y_sd <= 0.f) return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad normalisation");
    for (int i = 0; i < m->sizes[0]; ++i)
        if (!(sd[i] > 0.f)) return fail(PSKEL_ERR_INVALID_ARGUMENT, "feature std must be positive");
    m->mean.assign(mean, mean + m->sizes[0]);
    m->sd.assign(sd, sd + m->sizes[0]);
    m->y_mean = y_mean;
    m->y_sd = y_sd;
    PSKEL_HIP(hipSetDevice(m->device));
    return upload_norm(m);
}

// One forward + backward on n raw samples: loss (standardised target units) and
// gradient w.r.t. the parameters (the CPU checker's dLoss/dparams).
pskel_status_t pskel_mlp_loss_grad(pskel_mlp_t* m, const float* x, const float* y, size_t n, float* loss,
                                   float* grads)
{
    if (!m || !x || !y || n == 0 || n > (1u << 24)) return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad batch");
    PSKEL_HIP(hipSetDevice(m->device));
    pskel_status_t st = ensure_cap(m, n);
    if (st != PSKEL_OK) return st;
    const int nin = m->sizes[0];
    std::vector<float> yn(n);
    for (size_t s = 0; s < n; ++s) yn[s] = (y[s] - m->y_mean) / m->y_sd;
    float* d_raw = nullptr;
    PSKEL_HIP(hipMalloc(&d_raw, n * nin * sizeof(float)));
    hipError_t e = hipMemcpyAsync(d_raw, x, n * nin * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(m->d_y, yn.data(), n * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_standardise, dim3(nblk(n * nin)), dim3(256), 0, m->stream, d_raw, m->d_mean,
                           m->d_sd, nin, (int)n, m->d_x);
        e = forward(m, (int)n, m->d_x);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(256), 0, m->stream, m->d_act[m->nl - 1], m->d_y, (int)n,
                           m->d_dz[m->nl - 1], m->d_loss);
        e = backward(m, (int)n, m->d_x);
    }
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess && grads)
        e = hipMemcpyAsync(grads, m->d_g, m->np * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_raw);
    if (e != hipSuccess) return hip_fail(e, "mlp loss/grad");
    return PSKEL_OK;
}

// Minibatch Adam over (x, y): `epochs` passes, batches of `batch` samples in an
// order shuffled per epoch by `seed`.  Returns the loss of the last epoch
// (mean over its batches, standardised units).
pskel_status_t pskel_mlp_train(pskel_mlp_t* m, const float* x, const float* y, size_t n, int32_t epochs,
                               int32_t batch, float lr, uint64_t seed, float* final_loss)
{
    if (!m || !x || !y || n == 0 || epochs < 1 || batch < 1 || !(lr > 0.f))
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad training arguments");
    PSKEL_HIP(hipSetDevice(m->device));
    const int nin = m->sizes[0];
    const size_t bsz = std::min<size_t>((size_t)batch, n);
    pskel_status_t st = ensure_cap(m, bsz);
    if (st != PSKEL_OK) return st;
    std::vector<float> yn(n);
    for (size_t s = 0; s < n; ++s) yn[s] = (y[s] - m->y_mean) / m->y_sd;
    float *d_raw = nullptr, *d_xs = nullptr, *d_ys = nullptr, *d_xb = nullptr;
    int* d_idx = nullptr;
    float* d_losses = nullptr;
    const size_t nb = (n + bsz - 1) / bsz;
    auto cleanup = [&]() {
        if (d_raw) (void)hipFree(d_raw);
        if (d_xs) (void)hipFree(d_xs);
        if (d_ys) (void)hipFree(d_ys);
        if (d_xb) (void)hipFree(d_xb);
        if (d_idx) (void)hipFree(d_idx);
        if (d_losses) (void)hipFree(d_losses);
    };
    hipError_t e = hipMalloc(&d_raw, n * nin * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_xs, n * nin * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_ys, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_xb, bsz * nin * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_idx, n * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_losses, nb * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(d_raw, x, n * nin * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ys, yn.data(), n * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "mlp training buffers");
    }
    hipLaunchKernelGGL(k_mlp_standardise, dim3(nblk(n * nin)), dim3(256), 0, m->stream, d_raw, m->d_mean, m->d_sd,
                       nin, (int)n, d_xs);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937_64 rng(seed);
    std::vector<float> losses(nb);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float last = 0.f;
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
        std::shuffle(idx.begin(), idx.end(), rng);
        e = hipMemcpyAsync(d_idx, idx.data(), n * sizeof(int), hipMemcpyHostToDevice, m->stream);
        // the host vector is reused next epoch: make sure the copy has consumed it
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        for (size_t b = 0; b < nb && e == hipSuccess; ++b) {
            const int cnt = (int)std::min(bsz, n - b * bsz);
            hipLaunchKernelGGL(k_mlp_gather, dim3(nblk((size_t)cnt * nin)), dim3(256), 0, m->stream, d_xs, d_ys,
                               d_idx + b * bsz, nin, cnt, d_xb, m->d_y);
            e = forward(m, cnt, d_xb);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(256), 0, m->stream, m->d_act[m->nl - 1], m->d_y, cnt,
                               m->d_dz[m->nl - 1], d_losses + b);
            e = backward(m, cnt, d_xb);
            if (e != hipSuccess) break;
            m->step++;
            const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
            hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(m->np)), dim3(256), 0, m->stream, m->d_p, m->d_g, m->d_m, m->d_v,
                               (int)m->np, lr, b1, b2, eps, c1, c2);
            e = hipGetLastError();
        }
        if (e == hipSuccess && ep == epochs - 1) {
            e = hipMemcpyAsync(losses.data(), d_losses, nb * sizeof(float), hipMemcpyDeviceToHost, m->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
            double acc = 0.0;
            for (float v : losses) acc += v;
            last = (float)(acc / (double)nb);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "mlp training");
    if (final_loss) *final_loss = last;
    return PSKEL_OK;
}

pskel_status_t pskel_mlp_predict(pskel_mlp_t* m, const float* x, size_t n, float* y)
{
    return mlp_predict_host(m, x, n, y);
}

// File: "PSKMLP01", int32 nlayers, int32 sizes[nlayers+1], f32 mean[in], f32 sd[in],
// f32 y_mean, f32 y_sd, f32 params[np]  (little endian).
pskel_status_t pskel_mlp_save(const pskel_mlp_t* m, const char* path)
{
    if (!m || !path) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    std::vector<float> p(m->np);
    pskel_status_t st = pskel_mlp_get_params(m, p.data(), p.size());
    if (st != PSKEL_OK) return st;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(PSKEL_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    const int32_t nl = m->nl;
    bool ok = std::fwrite("PSKMLP01", 1, 8, f) == 8 && std::fwrite(&nl, 4, 1, f) == 1 &&
              std::fwrite(m->sizes, 4, nl + 1, f) == (size_t)nl + 1 &&
              std::fwrite(m->mean.data(), 4, m->mean.size(), f) == m->mean.size() &&
              std::fwrite(m->sd.data(), 4, m->sd.size(), f) == m->sd.size() && std::fwrite(&m->y_mean, 4, 1, f) == 1 &&
              std::fwrite(&m->y_sd, 4, 1, f) == 1 && std::fwrite(p.data(), 4, p.size(), f) == p.size();
    std::fclose(f);
    return ok ? PSKEL_OK : fail(PSKEL_ERR_RUNTIME, "short write");
}

pskel_status_t pskel_mlp_load(const char* path, int32_t device, pskel_mlp_t** out)
{
    if (!path || !out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(PSKEL_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    char magic[8];
    int32_t nl = 0;
    int32_t sizes[kMaxLayers + 1] = {0};
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "PSKMLP01", 8) == 0 && std::fread(&nl, 4, 1, f) == 1 &&
              nl >= 1 && nl <= kMaxLayers && std::fread(sizes, 4, nl + 1, f) == (size_t)nl + 1;
    pskel_mlp_t* m = nullptr;
    pskel_status_t st = PSKEL_OK;
    if (ok) st = pskel_mlp_create(sizes, nl, 0, device, &m);
    if (ok && st == PSKEL_OK) {
        std::vector<float> mean(sizes[0]), sd(sizes[0]), p(m->np);
        float ym = 0.f, ys = 1.f;
        ok = std::fread(mean.data(), 4, mean.size(), f) == mean.size() && std::fread(sd.data(), 4, sd.size(), f) == sd.size() &&
             std::fread(&ym, 4, 1, f) == 1 && std::fread(&ys, 4, 1, f) == 1 && std::fread(p.data(), 4, p.size(), f) == p.size();
        if (ok) st = pskel_mlp_set_normalization(m, mean.data(), sd.data(), ym, ys);
        if (ok && st == PSKEL_OK) st = pskel_mlp_set_params(m, p.data(), p.size());
    }
    std::fclose(f);
    if (!ok || st != PSKEL_OK) {
        if (m) pskel_mlp_destroy(m);
        return st != PSKEL_OK ? st : fail(PSKEL_ERR_INVALID_ARGUMENT, "not a cost-model file");
    }
    *out = m;
    return PSKEL_OK;
}

}  // extern "C"

namespace pskel {
pskel_status_t mlp_predict_host(const pskel_mlp_t* mc, const float* x, size_t n, float* y)
{
    pskel_mlp_t* m = const_cast<pskel_mlp_t*>(mc);
    if (!m || !x || !y || n == 0 || n > (1u << 24)) return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad predict arguments");
    PSKEL_HIP(hipSetDevice(m->device));
    pskel_status_t st = ensure_cap(m, n);
    if (st != PSKEL_OK) return st;
    const int nin = m->sizes[0];
    float* d_raw = nullptr;
    PSKEL_HIP(hipMalloc(&d_raw, n * nin * sizeof(float)));
    hipError_t e = hipMemcpyAsync(d_raw, x, n * nin * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_standardise, dim3(nblk(n * nin)), dim3(256), 0, m->stream, d_raw, m->d_mean, m->d_sd,
                           nin, (int)n, m->d_x);
        e = forward(m, (int)n, m->d_x);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, m->d_act[m->nl - 1], n * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_raw);
    if (e != hipSuccess) return hip_fail(e, "mlp predict");
    for (size_t s = 0; s < n; ++s) y[s] = y[s] * m->y_sd + m->y_mean;
    return PSKEL_OK;
}
}  // namespace pskel
