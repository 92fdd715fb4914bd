// Field arguments (Args / Args2D / Args3D) resident in HBM, and the CLOUDSIM
// stencil that reads them.
//
// The reference keeps Args fields in zero-copy mapped host memory
// (include/PSkelArgs.hpp:49-161), so every sweep reads them over the host
// link.  Here a field is uploaded once per run and read from HBM.  Its host
// data is in the Args layout, x fastest: (z*height + y)*width + x
// (include/PSkelArgs.hpp:116-123, 155-160).  The reference's kernels call a
// field with the cell's coordinates, args.f(h, w) (include/PSkelStencil.hpp:
// 366-369), i.e. x = h and y = w; so the device copy reverses the axes --
// element (x, y) at x*pitch + y, (x, y, z) at (x*height + y)*pitch + z -- and
// a row of cells reads a contiguous row of the field.
//
// CLOUDSIM is the reference's CLOUDSIM_KERNEL formula (include/PSkelStencil.hpp:
// 349-387) on the global neighbours of every interior cell, in the same
// evaluation order, no contraction.
#include <algorithm>
#include <new>

#include "k_tiles2d.hpp"

struct pskel_field {
    int ndim = 2;
    size_t W = 0, H = 1, D = 1;  // Args extents: x, y, z
    size_t pitch = 0;            // elements of the contiguous device axis
    size_t elems = 0;            // device elements
    int device = 0;
    float* data = nullptr;
    float* stage = nullptr;      // host-layout staging buffer
};

namespace pskel {

namespace {

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// stage (host layout) -> data (reversed layout); one lane per device element,
// so the writes are coalesced and each element is read once
__global__ __launch_bounds__(256) void k_field_reverse(float* __restrict__ dst, const float* __restrict__ src,
                                                       int ndim, size_t W, size_t H, size_t D, size_t pitch,
                                                       size_t total)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (ndim == 2) {
        const size_t x = i / pitch, y = i % pitch;
        dst[i] = y < H ? src[y * W + x] : 0.f;
    } else {  // 3D
        const size_t z = i % pitch, xy = i / pitch;
        const size_t y = xy % H, x = xy / H;
        dst[i] = z < D ? src[(z * H + y) * W + x] : 0.f;
    }
}

// data (reversed layout) -> stage (host layout)
__global__ __launch_bounds__(256) void k_field_restore(float* __restrict__ dst, const float* __restrict__ src,
                                                       int ndim, size_t W, size_t H, size_t D, size_t pitch,
                                                       size_t total)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (ndim == 2) {
        const size_t x = i % W, y = i / W;
        dst[i] = src[x * pitch + y];
    } else {
        const size_t x = i % W, yz = i / W;
        const size_t y = yz % H, z = yz / H;
        dst[i] = src[(x * H + y) * pitch + z];
    }
}

// ---------------------------------------------------------------------------
// CLOUDSIM
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cloud_cell(float c, float n, float e, float ww, float s, float xw, float yw,
                                            float dt)
{
    const float sum = 4.0f * c - (((n + e) + ww) + s);
    const int xf = xw > 0.f ? 1 : -1;
    const int yf = yw > 0.f ? 1 : -1;
    const float cy = (float)yf * yw;
    const float cx = (float)xf * xw;
    const float tx = xf > 0 ? e : ww;  // in(h, w + xf)
    const float ty = yf > 0 ? s : n;   // in(h + yf, w)
    const float wind = (-cx * ((c - tx) * 10.0f)) - (cy * ((c - ty) * 10.0f));
    const float cond = (-0.0243f * (sum * 0.25f)) * dt;
    const float result = c + cond;
    return result + wind * dt;
}

__global__ __launch_bounds__(1024) void k2d_cloud_plain(Sweep2D s, CloudArgs a)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y * blockDim.y + threadIdx.y;
    if (w < 1 || w >= s.W - 1 || h < 1 || h >= s.H - 1) return;
    const float* __restrict__ in = (const float*)s.src;
    const size_t p = s.pitch;
    const size_t i = (size_t)h * p + w;
    const size_t f = (size_t)h * a.fpitch + w;
    ((float*)s.dst)[i] = cloud_cell(in[i], in[i - p], in[i + 1], in[i - 1], in[i + p], a.wx[f], a.wy[f], a.dt);
}

// 64 lanes x float4 = 256 columns per wave, RY rows per lane, waves stacked
// along h; the window and its cross-lane halo as in the tuned 2D kernels, plus
// one float4 of each wind field per output row.
template <int RY>
__global__ __launch_bounds__(256) void k2d_cloud_tuned(Sweep2D s, CloudArgs a)
{
    const int lane = threadIdx.x;
    const int w0 = blockIdx.x * 256;
    const int w = w0 + lane * 4;
    const int hb = ((int)blockIdx.y * (int)blockDim.y + (int)threadIdx.y) * RY;
    const int H = s.H, Wd = s.W;
    const float* __restrict__ src = (const float*)s.src;
    float* __restrict__ dst = (float*)s.dst;
    const size_t pitch = s.pitch;
    const bool col_live = w < Wd;

    float4 v[RY + 2];
    float lfl[RY + 2], rtl[RY + 2];
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) {
        const int h = hb - 1 + i;
        v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        lfl[i] = 0.f;
        rtl[i] = 0.f;
        if (h >= 0 && h < H) {
            const float* row = src + (size_t)h * pitch;
            if (col_live) v[i] = *reinterpret_cast<const float4*>(row + w);
            if (lane == 0 && w0 > 0) lfl[i] = row[w0 - 1];
            if (lane == kWave - 1 && w0 + 256 < Wd) rtl[i] = row[w0 + 256];
        }
    }
    // wind of the output rows (field rows are padded to whole tiles)
    float4 fx[RY], fy[RY];
#pragma unroll
    for (int i = 0; i < RY; ++i) {
        const int h = hb + i;
        fx[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        fy[i] = fx[i];
        if (col_live && h >= 1 && h < H - 1) {
            fx[i] = *reinterpret_cast<const float4*>(a.wx + (size_t)h * a.fpitch + w);
            fy[i] = *reinterpret_cast<const float4*>(a.wy + (size_t)h * a.fpitch + w);
        }
    }
    float lf[RY + 2], rt[RY + 2];
#pragma unroll
    for (int i = 0; i < RY + 2; ++i) {
        lf[i] = from_lower_lane(v[i].w, lfl[i]);  // lane 0: the cell loaded for it
        rt[i] = from_upper_lane(v[i].x, rtl[i]);  // lane 63: likewise
    }
    if (!col_live) return;
    const bool full = (w >= 1) && (w + 3 < Wd - 1);
#pragma unroll
    for (int i = 1; i <= RY; ++i) {
        const int h = hb - 1 + i;
        if (h < 1 || h >= H - 1) continue;
        const Win win{&v[i - 1], &lf[i - 1], &rt[i - 1]};
        const float xw[4] = {fx[i - 1].x, fx[i - 1].y, fx[i - 1].z, fx[i - 1].w};
        const float yw[4] = {fy[i - 1].x, fy[i - 1].y, fy[i - 1].z, fy[i - 1].w};
        float o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            o[j] = cloud_cell(win.at(1, j), win.at(0, j), win.at(1, j + 1), win.at(1, j - 1), win.at(2, j), xw[j],
                              yw[j], a.dt);
        float* row = dst + (size_t)h * pitch + w;
        if (full) {
            *reinterpret_cast<float4*>(row) = make_float4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (w + j >= 1 && w + j < Wd - 1) row[j] = o[j];
        }
    }
}

}  // namespace

hipError_t launch_cloud_plain(const Sweep2D& s, const CloudArgs& a, int bx, int by, hipStream_t st)
{
    const dim3 block(bx, by);
    const dim3 grid((s.W + bx - 1) / bx, (s.H + by - 1) / by);
    hipLaunchKernelGGL(k2d_cloud_plain, grid, block, 0, st, s, a);
    return hipGetLastError();
}

hipError_t launch_cloud_tuned(const Sweep2D& s, const CloudArgs& a, const pskel_launch_cfg_t& cfg, hipStream_t st)
{
    const int ry = cfg.tile_y > 0 ? cfg.tile_y : 4;
    const int nwy = cfg.tile_z > 0 ? cfg.tile_z : 4;
    if (nwy < 1 || nwy > 4) return hipErrorInvalidValue;
    const dim3 block(kWave, nwy);
#define PSKEL_CLOUD_CASE(RYV)                                                                   \
    case RYV: {                                                                                 \
        const dim3 grid((s.W + 255) / 256, (s.H + nwy * RYV - 1) / (nwy * RYV));                \
        hipLaunchKernelGGL((k2d_cloud_tuned<RYV>), grid, block, 0, st, s, a);                   \
        break;                                                                                  \
    }
    switch (ry) {
        PSKEL_CLOUD_CASE(1)
        PSKEL_CLOUD_CASE(2)
        PSKEL_CLOUD_CASE(4)
        PSKEL_CLOUD_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef PSKEL_CLOUD_CASE
    return hipGetLastError();
}

bool field_info(const pskel_field_t* f, int* ndim, size_t* W, size_t* H, size_t* D, const float** data,
                size_t* pitch)
{
    if (!f) return false;
    *ndim = f->ndim;
    *W = f->W;
    *H = f->H;
    *D = f->D;
    *data = f->data;
    *pitch = f->pitch;
    return true;
}

}  // namespace pskel

using namespace pskel;

extern "C" {

pskel_status_t pskel_field_create(const pskel_field_desc_t* desc, pskel_field_t** out)
{
    if (!desc || !out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (desc->ndim < 1 || desc->ndim > 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "field ndim must be 1, 2 or 3");
    if (desc->dtype != PSKEL_DTYPE_F32) return fail(PSKEL_ERR_UNSUPPORTED, "fields hold f32");
    const size_t W = desc->width, H = desc->ndim >= 2 ? desc->height : 1, D = desc->ndim == 3 ? desc->depth : 1;
    if (W == 0 || H == 0 || D == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "empty field");
    if (W > (1u << 30) || H > (1u << 30) || D > (1u << 30)) return fail(PSKEL_ERR_INVALID_ARGUMENT, "field too large");
    PSKEL_HIP(hipSetDevice(desc->device));
    auto* f = new (std::nothrow) pskel_field_t();
    if (!f) return fail(PSKEL_ERR_OUT_OF_MEMORY, "host allocation");
    f->ndim = desc->ndim;
    f->W = W;
    f->H = H;
    f->D = D;
    f->device = desc->device;
    // the contiguous device axis is padded to whole 1 KiB so tile loads stay inside
    if (f->ndim == 1) {
        f->pitch = round_up(W, 256);
        f->elems = f->pitch;
    } else if (f->ndim == 2) {
        f->pitch = round_up(H, 256);
        f->elems = W * f->pitch;
    } else {
        f->pitch = round_up(D, 256);
        f->elems = W * H * f->pitch;
    }
    hipError_t e = hipMalloc(&f->data, f->elems * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&f->stage, W * H * D * sizeof(float));
    if (e == hipSuccess) e = hipMemset(f->data, 0, f->elems * sizeof(float));
    if (e != hipSuccess) {
        const pskel_status_t st = hip_fail(e, "field allocation");
        pskel_field_destroy(f);
        return st;
    }
    *out = f;
    return PSKEL_OK;
}

pskel_status_t pskel_field_destroy(pskel_field_t* f)
{
    if (!f) return PSKEL_OK;
    (void)hipSetDevice(f->device);
    (void)hipDeviceSynchronize();
    if (f->data) (void)hipFree(f->data);
    if (f->stage) (void)hipFree(f->stage);
    delete f;
    return PSKEL_OK;
}

pskel_status_t pskel_field_upload(pskel_field_t* f, const void* host)
{
    if (!f || !host) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    PSKEL_HIP(hipSetDevice(f->device));
    const size_t n = f->W * f->H * f->D;
    if (f->ndim == 1) {
        PSKEL_HIP(hipMemcpy(f->data, host, n * sizeof(float), hipMemcpyHostToDevice));
        return PSKEL_OK;
    }
    PSKEL_HIP(hipMemcpy(f->stage, host, n * sizeof(float), hipMemcpyHostToDevice));
    const size_t total = f->elems;
    hipLaunchKernelGGL(k_field_reverse, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, nullptr, f->data,
                       f->stage, f->ndim, f->W, f->H, f->D, f->pitch, total);
    PSKEL_HIP(hipGetLastError());
    PSKEL_HIP(hipDeviceSynchronize());
    return PSKEL_OK;
}

pskel_status_t pskel_field_download(const pskel_field_t* f, void* host)
{
    if (!f || !host) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    PSKEL_HIP(hipSetDevice(f->device));
    const size_t n = f->W * f->H * f->D;
    if (f->ndim == 1) {
        PSKEL_HIP(hipMemcpy(host, f->data, n * sizeof(float), hipMemcpyDeviceToHost));
        return PSKEL_OK;
    }
    hipLaunchKernelGGL(k_field_restore, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, f->stage, f->data,
                       f->ndim, f->W, f->H, f->D, f->pitch, n);
    PSKEL_HIP(hipGetLastError());
    PSKEL_HIP(hipMemcpy(host, f->stage, n * sizeof(float), hipMemcpyDeviceToHost));
    return PSKEL_OK;
}

pskel_status_t pskel_field_info(const pskel_field_t* f, void** device_ptr, size_t* pitch)
{
    if (!f) return fail(PSKEL_ERR_INVALID_ARGUMENT, "field is NULL");
    if (device_ptr) *device_ptr = f->data;
    if (pitch) *pitch = f->pitch;
    return PSKEL_OK;
}

}  // extern "C"
