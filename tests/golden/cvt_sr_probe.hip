// Records what gfx950's stochastic-rounding conversions v_cvt_sr_bf16_f32 and
// v_cvt_sr_f16_f32 return for chosen (value, random word) pairs: special
// values crossed with structured random words, then random values spread over
// all bit patterns, the f16 range and the f16 subnormal band.
//
// Build: hipcc -O2 --offload-arch=gfx950 -o cvt_sr_probe cvt_sr_probe.hip
// Run (on a gfx950 device): ./cvt_sr_probe out.bin
// Output: int32 n, then n f32 values, n u32 random words, n x 3 u32 results
// (bf16 in the low half, f16 in the low half, bf16 written to the high half).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

__global__ void k_probe(const float* x, const unsigned* r, unsigned* out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bf16x2 lo = {};
    lo = __builtin_amdgcn_cvt_sr_bf16_f32(lo, x[i], r[i], false);
    f16x2 h = {};
    h = __builtin_amdgcn_cvt_sr_f16_f32(h, x[i], r[i], false);
    bf16x2 hi = {};
    hi = __builtin_amdgcn_cvt_sr_bf16_f32(hi, x[i], r[i], true);
    out[3 * i] = __builtin_bit_cast(unsigned, lo);
    out[3 * i + 1] = __builtin_bit_cast(unsigned, h);
    out[3 * i + 2] = __builtin_bit_cast(unsigned, hi);
}

static uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

#define CHECK(call)                                                              \
    do {                                                                         \
        if ((call) != hipSuccess) {                                              \
            fprintf(stderr, "%s failed\n", #call);                               \
            return 1;                                                            \
        }                                                                        \
    } while (0)

int main(int argc, char** argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s out.bin\n", argv[0]);
        return 2;
    }
    std::vector<float> xs;
    std::vector<unsigned> rs;
    const float specials[] = {0.f, -0.f, 1.f, -1.f, 1.00390625f, 0.5f, 65504.f, 65519.f, 65520.f, 70000.f,
                              3.0e38f, 1e-30f, 6.1e-5f, 6.0e-8f, 3.0e-8f, 1e-7f, 2.9e-8f, -5e-6f,
                              1.17549435e-38f, 1e-39f};
    const unsigned rspec[] = {0u, 1u, 0xffffu, 0x10000u, 0xffff0000u, 0xffffffffu,
                              0x8000u, 0x80000000u, 0x1fffu, 0x2000u, 0x7fffu, 0x00ff00ffu};
    for (float f : specials)
        for (unsigned r : rspec) {
            xs.push_back(f);
            rs.push_back(r);
        }
    const uint32_t bits[] = {0x7f800000u, 0xff800000u, 0x7fc00000u, 0x7fffffffu, 0x7f800001u,
                             0xffc00001u, 0x7f7fffffu, 0x3f80ffffu, 0x477fe000u, 0x477fefffu};
    for (uint32_t b : bits)
        for (unsigned r : rspec) {
            float f;
            memcpy(&f, &b, 4);
            xs.push_back(f);
            rs.push_back(r);
        }
    uint64_t st = 12345;
    for (int i = 0; i < 200000; ++i) {
        const uint64_t a = splitmix(st), b = splitmix(st);
        float f;
        uint32_t u;
        switch (i % 4) {
        case 0: u = (uint32_t)a; memcpy(&f, &u, 4); break;                                   // any bits
        case 1: u = 0x30000000u + (uint32_t)(a % 0x18000000u); memcpy(&f, &u, 4); break;    // f16 range
        case 2: u = 0x33000000u + (uint32_t)(a % 0x06000000u); memcpy(&f, &u, 4); break;    // f16 subnormals
        default: f = (float)((int64_t)(a >> 11) - (int64_t)(1ll << 52)) / (float)(1ll << 52) * 2.0f; break;
        }
        xs.push_back(f);
        rs.push_back((unsigned)b);
    }
    const int n = (int)xs.size();
    float* dx = nullptr;
    unsigned *dr = nullptr, *dout = nullptr;
    CHECK(hipMalloc(&dx, n * sizeof(float)));
    CHECK(hipMalloc(&dr, n * sizeof(unsigned)));
    CHECK(hipMalloc(&dout, 3 * (size_t)n * sizeof(unsigned)));
    CHECK(hipMemcpy(dx, xs.data(), n * sizeof(float), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dr, rs.data(), n * sizeof(unsigned), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dr, dout, n);
    CHECK(hipGetLastError());
    std::vector<unsigned> out(3 * (size_t)n);
    CHECK(hipMemcpy(out.data(), dout, out.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
    FILE* f = fopen(argv[1], "wb");
    if (!f) return 1;
    fwrite(&n, 4, 1, f);
    fwrite(xs.data(), 4, n, f);
    fwrite(rs.data(), 4, n, f);
    fwrite(out.data(), 4, out.size(), f);
    fclose(f);
    printf("wrote %d cases\n", n);
    return 0;
}
