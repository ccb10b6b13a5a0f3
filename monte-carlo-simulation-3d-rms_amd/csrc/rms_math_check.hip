// rms_math_check.hip -- exhaustive checks of the walk kernel's FP32 direction
// math (run as a program on the GPU):
//   sqrt_rn_mid(x) == sqrtf(x) (correctly rounded) for every float x in
//     [2^-25, 1/4], the range of u(1-u) in the direction draw;
//   sincos_turns_table(w) against double sin/cos(2 pi m 2^-24) for all 2^24
//     distinct mantissas m = (w >> 8) | 1: largest absolute error (in units
//     of 2^-24) and largest deviation of s^2 + c^2 from 1.
// Prints the results; exit status 0 when the square root never differs, the
// sin/cos error is below 2.5 units and the unit-length error below 3e-7.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "rms_device.h"
#include "rms_turn_table.h"

using namespace rms;

__global__ void check_sqrt(uint32_t first, uint32_t n, unsigned long long *bad)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const float x = __uint_as_float(first + t);
        if (__float_as_uint(sqrt_rn_mid(x)) != __float_as_uint(sqrtf(x))) atomicAdd(bad, 1ull);
    }
}

__global__ void check_sincos(unsigned long long *err)
{
    __shared__ float2 tab[256];
    for (int k = (int)threadIdx.x; k < 256; k += (int)blockDim.x) tab[k] = turn_table[k];
    __syncthreads();
    unsigned long long worst = 0, worst_n = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < (1u << 24); t += gridDim.x * blockDim.x) {
        const uint32_t w = t << 8;
        float s, c;
        sincos_turns_table(w, tab, s, c);
        const double a = 6.283185307179586 * (double)((w >> 8) | 1u) * 5.9604644775390625e-08;
        const double e = fmax(fabs((double)s - sin(a)), fabs((double)c - cos(a)));
        const double n = fabs((double)s * s + (double)c * c - 1.0);
        // error scaled to units of 2^-34 (2^-24 / 1024) and of 1e-12 as integers
        const unsigned long long ei = (unsigned long long)(e * 17179869184.0);
        const unsigned long long ni = (unsigned long long)(n * 1e12);
        worst = ei > worst ? ei : worst;
        worst_n = ni > worst_n ? ni : worst_n;
    }
    atomicMax(err, worst);
    atomicMax(err + 1, worst_n);
}

int main()
{
    unsigned long long *bad = nullptr;
    if (hipMalloc(&bad, 3 * sizeof(unsigned long long)) != hipSuccess) return 1;
    if (hipMemset(bad, 0, 3 * sizeof(unsigned long long)) != hipSuccess) return 1;
    const float lo = 2.98023223876953125e-08f, hi = 0.25f;  // 2^-25, 1/4
    uint32_t a = 0, b = 0;
    memcpy(&a, &lo, 4);
    memcpy(&b, &hi, 4);
    check_sqrt<<<4096, 256>>>(a, b - a + 1, bad);
    check_sincos<<<4096, 256>>>(bad + 1);
    unsigned long long h[3] = {1, 1, 1};
    if (hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    (void)hipFree(bad);
    const double sc_err = (double)h[1] / 1024.0, n_err = (double)h[2] * 1e-12;
    printf("sqrt_rn_mid mismatches over %u floats: %llu\n", b - a + 1, h[0]);
    printf("sincos_turns_table max abs error over %u words: %.3f x 2^-24\n", 1u << 24, sc_err);
    printf("sincos_turns_table max |s^2 + c^2 - 1|: %.3e\n", n_err);
    return (h[0] || sc_err >= 2.5 || n_err >= 3e-7) ? 2 : 0;
}
