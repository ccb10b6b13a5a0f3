// rms_math_check.hip -- exhaustive equivalence checks of two device-side
// shortcuts the walk kernel takes in FP32 (run as a program on the GPU):
//   sqrt_rn_mid(x) == sqrtf(x) (correctly rounded) for every float x in
//     [2^-25, 1/4], the range of u(1-u) in the direction draw;
//   sincos_turns_word(w) == sincos_turns(u01_f32(w)) for all 2^24 distinct
//     mantissas m = (w >> 8) | 1.
// Prints the mismatch counts; exit status 0 when both are zero.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "rms_device.h"

using namespace rms;

__global__ void check_sqrt(uint32_t first, uint32_t n, unsigned long long *bad)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const float x = __uint_as_float(first + t);
        if (__float_as_uint(sqrt_rn_mid(x)) != __float_as_uint(sqrtf(x))) atomicAdd(bad, 1ull);
    }
}

__global__ void check_sincos(unsigned long long *bad)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < (1u << 24); t += gridDim.x * blockDim.x) {
        const uint32_t w = t << 8;
        float s1, c1, s2, c2;
        sincos_turns_word(w, s1, c1);
        sincos_turns(u01_f32(w), s2, c2);
        if (__float_as_uint(s1) != __float_as_uint(s2) || __float_as_uint(c1) != __float_as_uint(c2))
            atomicAdd(bad, 1ull);
    }
}

int main()
{
    unsigned long long *bad = nullptr;
    if (hipMalloc(&bad, 2 * sizeof(unsigned long long)) != hipSuccess) return 1;
    if (hipMemset(bad, 0, 2 * sizeof(unsigned long long)) != hipSuccess) return 1;
    const float lo = 2.98023223876953125e-08f, hi = 0.25f;  // 2^-25, 1/4
    uint32_t a = 0, b = 0;
    memcpy(&a, &lo, 4);
    memcpy(&b, &hi, 4);
    check_sqrt<<<4096, 256>>>(a, b - a + 1, bad);
    check_sincos<<<4096, 256>>>(bad + 1);
    unsigned long long h[2] = {1, 1};
    if (hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    hipFree(bad);
    printf("sqrt_rn_mid mismatches over %u floats: %llu\n", b - a + 1, h[0]);
    printf("sincos_turns_word mismatches over %u words: %llu\n", 1u << 24, h[1]);
    return (h[0] || h[1]) ? 2 : 0;
}
