// Device-side helpers shared by the kernel translation units (gfx950, wave64).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pskel_internal.hpp"

namespace pskel {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Stochastic rounding words.  mix32 is a murmur3-style avalanche.  The 32-bit
// random word of cell (h, w, d) in sweep t is A(w*D + d) ^ B(t, h), two
// independent hash streams: A of the in-plane index under a key of the seed
// alone, B of the plane under the sweep's key (seed and sweep index).  Two
// distinct (cell, sweep) pairs get pairwise independent words (distinct
// in-plane index: independent A terms; the same one: the B terms differ), so
// the rounding errors of a linear sweep are uncorrelated.  A kernel marching
// along h hashes A once per cell it owns, for all its planes and sweeps, and B
// once per plane (a scalar).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

__host__ __device__ __forceinline__ uint32_t sweep_key(uint32_t seed, uint64_t sweep)
{
    return mix32(seed + 0x9e3779b9u * ((uint32_t)sweep + 1u));
}

__host__ __device__ __forceinline__ uint32_t sr_cell_key(uint32_t seed)
{
    return mix32(mix32(seed) ^ 0xa511e9b3u);
}

// B(t, h): `key` is sweep_key(seed, t)
__host__ __device__ __forceinline__ uint32_t sr_plane_bits(uint32_t key, uint32_t h)
{
    return mix32(mix32(key + 0x632be5abu) ^ h);
}

// A(wd): `cell_key` is sr_cell_key(seed)
__host__ __device__ __forceinline__ uint32_t sr_cell_bits(uint32_t cell_key, uint32_t wd)
{
    return mix32(cell_key ^ wd);
}

// ---------------------------------------------------------------------------
// Storage formats.  Arithmetic is always f32; 16-bit formats store (u - off),
// rounded to nearest even (enc) or stochastically (enc_sr) on the way out.
// dec / enc / enc_sr add and subtract the offset (weighted sums); dec_raw /
// enc_raw / enc_sr_raw work on the stored values themselves (scaled sums,
// whose offset term is folded into one fused multiply-add).
// Stochastic rounding is gfx950's v_cvt_sr_{bf16,f16}_f32: the top 16 (bf16)
// or 13 (f16) bits of the random word are added below the kept bits of the
// magnitude, which is then truncated, so it rounds up with probability equal
// to the dropped fraction of an ulp.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
template <int ST>
struct Stor;

template <>
struct Stor<PSKEL_DTYPE_F32> {
    using T = float;
    __device__ __forceinline__ static float dec(T v, float) { return v; }
    __device__ __forceinline__ static float dec_raw(T v) { return v; }
    __device__ __forceinline__ static T enc(float v, float) { return v; }
    __device__ __forceinline__ static T enc_sr(float v, float, uint32_t) { return v; }
    __device__ __forceinline__ static T enc_raw(float v) { return v; }
    __device__ __forceinline__ static T enc_sr_raw(float v, uint32_t) { return v; }
};

template <>
struct Stor<PSKEL_DTYPE_BF16> {
    using T = uint16_t;
    __device__ __forceinline__ static float dec_raw(T v) { return __uint_as_float((uint32_t)v << 16); }
    __device__ __forceinline__ static float dec(T v, float off) { return dec_raw(v) + off; }
    // gfx950 converts natively (v_cvt_pk_bf16_f32, round to nearest even)
    __device__ __forceinline__ static T enc(float v, float off)
    {
        const __bf16 b = (__bf16)(v - off);
        return __builtin_bit_cast(T, b);
    }
    __device__ __forceinline__ static T enc_sr(float v, float off, uint32_t r) { return enc_sr_raw(v - off, r); }
    __device__ __forceinline__ static T enc_raw(float v)
    {
        const __bf16 b = (__bf16)v;
        return __builtin_bit_cast(T, b);
    }
    __device__ __forceinline__ static T enc_sr_raw(float v, uint32_t r)
    {
        // the converted half only: the other half is left unspecified (no
        // register initialisation)
        bf16x2_t o = __builtin_nondeterministic_value(o);
        o = __builtin_amdgcn_cvt_sr_bf16_f32(o, v, r, false);
        return __builtin_bit_cast(T, o[0]);
    }
};

// The value as rounded to f32, hidden from the optimiser: gfx9 has a
// mixed-precision FMA that rounds an f32 product-sum straight to f16
// (v_fma_mix*_f16), one rounding where the storage expression rounds twice
// (to f32, then to storage).  An empty asm with the value in a VGPR keeps the
// conversion on the f32 result; it emits no instruction.
__device__ __forceinline__ float as_rounded_f32(float x)
{
    asm("" : "+v"(x));
    return x;
}

template <>
struct Stor<PSKEL_DTYPE_F16> {
    using T = uint16_t;
    __device__ __forceinline__ static float dec_raw(T v) { return __half2float(__ushort_as_half(v)); }
    __device__ __forceinline__ static float dec(T v, float off) { return dec_raw(v) + off; }
    __device__ __forceinline__ static T enc(float v, float off)
    {
        return __half_as_ushort(__float2half_rn(as_rounded_f32(v - off)));
    }
    __device__ __forceinline__ static T enc_sr(float v, float off, uint32_t r) { return enc_sr_raw(v - off, r); }
    __device__ __forceinline__ static T enc_raw(float v) { return __half_as_ushort(__float2half_rn(as_rounded_f32(v))); }
    __device__ __forceinline__ static T enc_sr_raw(float v, uint32_t r)
    {
        f16x2_t o = __builtin_nondeterministic_value(o);
        o = __builtin_amdgcn_cvt_sr_f16_f32(o, v, r, false);
        return __builtin_bit_cast(T, o[0]);
    }
};

// Raw buffer resources.  The hardware range check turns a load past the
// resource's size into zeros and drops such a store, so kernels can issue the
// same memory operations on every path (exact wait counts) and mask with an
// offset past the data, or a resource of 0 bytes.
using u32x4_t = __attribute__((ext_vector_type(4))) unsigned int;
using u32x2_t = __attribute__((ext_vector_type(2))) unsigned int;

// a raw buffer over `bytes` bytes from p (0 bytes: every access out of range)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}

constexpr uint32_t kNoAccess = 0x80000000u;  // an offset past any buffer (all buffers stay under 2 GiB)

// Vector load/store of N consecutive elements of T (N * sizeof(T) in {2,4,8,16} bytes).
template <typename T, int N>
struct Vec {
    T v[N];
};

template <typename T, int N>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, T (&out)[N])
{
    constexpr int B = N * (int)sizeof(T);
    if constexpr (B == 16) {
        uint4 x = *reinterpret_cast<const uint4*>(p);
        __builtin_memcpy(out, &x, 16);
    } else if constexpr (B == 8) {
        uint2 x = *reinterpret_cast<const uint2*>(p);
        __builtin_memcpy(out, &x, 8);
    } else if constexpr (B == 4) {
        uint32_t x = *reinterpret_cast<const uint32_t*>(p);
        __builtin_memcpy(out, &x, 4);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = p[i];
    }
}

template <typename T, int N>
__device__ __forceinline__ void store_vec(T* __restrict__ p, const T (&in)[N])
{
    // non-temporal: a sweep's output grid is not read again within its launch
    constexpr int B = N * (int)sizeof(T);
    if constexpr (B == 16) {
        u32x4_t x;
        __builtin_memcpy(&x, in, 16);
        __builtin_nontemporal_store(x, reinterpret_cast<u32x4_t*>(p));
    } else if constexpr (B == 8) {
        u32x2_t x;
        __builtin_memcpy(&x, in, 8);
        __builtin_nontemporal_store(x, reinterpret_cast<u32x2_t*>(p));
    } else if constexpr (B == 4) {
        uint32_t x;
        __builtin_memcpy(&x, in, 4);
        __builtin_nontemporal_store(x, reinterpret_cast<uint32_t*>(p));
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) p[i] = in[i];
    }
}

// Neighbour exchange across lanes: lane l receives lane l-1's value
// (wave_shr:1) or lane l+1's (wave_shl:1); the lane with no source (lane 0,
// respectively 63) keeps `edge`.  DPP controls of the row-shift family on
// wave64 GFX9 parts (gfx950 included): a VALU modifier, no LDS round trip.
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v, uint32_t edge)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, kDppWaveShr1, 0xf, 0xf, false);
}

__device__ __forceinline__ uint32_t from_upper_lane(uint32_t v, uint32_t edge)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, kDppWaveShl1, 0xf, 0xf, false);
}

__device__ __forceinline__ float from_lower_lane(float v, float edge)
{
    return __uint_as_float(from_lower_lane(__float_as_uint(v), __float_as_uint(edge)));
}

__device__ __forceinline__ float from_upper_lane(float v, float edge)
{
    return __uint_as_float(from_upper_lane(__float_as_uint(v), __float_as_uint(edge)));
}

// ---------------------------------------------------------------------------
// Deterministic FP64 block reduction: wave butterfly, then waves in order.
// Every thread of the block must call it.  Returns the block sum in thread 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <int NWAVES>
__device__ __forceinline__ double block_sum(double v, double* scratch /* NWAVES entries */)
{
    const int lane = threadIdx.x % kWave;
    const int wv = threadIdx.x / kWave;
    v = wave_sum(v);
    if (lane == 0) scratch[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NWAVES; ++i) s += scratch[i];
    }
    return s;
}

// Map a linear block id so that consecutive logical ids land on the same XCD:
// the dispatcher hands blocks to the 8 XCDs round-robin, so logical neighbours
// (which share halo lines) would otherwise sit in 8 different L2s.
// XCD x receives blocks x, x+8, x+16, ...: ceil((n - x) / 8) of them, the
// first n % 8 XCDs one more than the rest.  It gets the logical ids
// [start(x), start(x) + count(x)), a bijection for any block count.
__host__ __device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t x = bid % kXcd, k = bid / kXcd;
    const uint32_t q = nblocks / kXcd, rem = nblocks % kXcd;
    return x * q + (x < rem ? x : rem) + k;
}

}  // namespace pskel
