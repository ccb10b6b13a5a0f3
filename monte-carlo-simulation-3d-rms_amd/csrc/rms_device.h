// rms_device.h -- device-side random numbers and elementary functions for the
// RMS walker kernels (gfx950).
//
// Every routine here is a fixed sequence of IEEE-754 basic operations (add,
// mul, fma, floor, correctly rounded div/sqrt) so that the host restatement
// reproduces it bit for bit.  The kernels are compiled with
// -ffp-contract=off: no a*b+c is fused unless written as fma().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rms {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy
// as 1, 2, 3", SC'11).  Replaces the per-thread cuRAND state of
// hpc_code/lib/main.cu:58-61,66: counters are (index, stream, walker lo, hi).
struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                     uint32_t k1)
{
    // the key is wave-uniform: keep it (and its per-round bumps) in SGPRs
    k0 = __builtin_amdgcn_readfirstlane(k0);
    k1 = __builtin_amdgcn_readfirstlane(k1);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // one 32x32->64 product per multiplier (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        // three-input xor in one v_bitop3_b32 (truth table 0x96)
        const uint32_t n0 = r == 0 ? hi1 ^ (c1 ^ k0) : __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = r == 0 ? hi0 ^ (c3 ^ k1) : __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// Streams of the counter layout.
// Streams 5..7: reservoir decisions of the adjoint's exchange-face events.
enum : uint32_t {
    STREAM_DIR = 0,
    STREAM_EXCH_A = 1,
    STREAM_EXCH_B = 2,
    STREAM_INIT_XY = 3,
    STREAM_INIT_Z = 4,
    STREAM_JUMP = 5,     // 5..7: exchange-face and lattice-edge event reservoir
    STREAM_CORNER = 8    // 8..13: voxel-edge event reservoir
};

__device__ __forceinline__ U4 rng_block(uint32_t key0, uint32_t key1, uint32_t idx, uint32_t stream, uint64_t wid)
{
    return philox(idx, stream, (uint32_t)wid, (uint32_t)(wid >> 32), key0, key1);
}

// Open-interval uniforms: (2k+1)*2^-53 with a 52-bit k, (2k+1)*2^-24 with 23 bits.
__device__ __forceinline__ double u01_f64(uint32_t a, uint32_t b)
{
    double hi = (double)(a >> 6);
    double lo = (double)(b >> 6);
    double k = hi * 67108864.0 + lo;
    return (2.0 * k + 1.0) * 1.1102230246251565e-16;
}

__device__ __forceinline__ float u01_f32(uint32_t w)
{
    float k = (float)(w >> 9);
    return (2.0f * k + 1.0f) * 5.9604644775390625e-08f;
}

// ---------------------------------------------------------------- sin / cos
// Taylor polynomials on |x| <= pi/4 (terms to x^16 in FP64, x^10 in FP32).
__device__ __forceinline__ void sc_poly(double x, double &s, double &c)
{
    double z = x * x;
    double p = -7.647163731819816e-13;
    p = __builtin_fma(p, z, 1.6059043836821613e-10);
    p = __builtin_fma(p, z, -2.505210838544172e-08);
    p = __builtin_fma(p, z, 2.7557319223985893e-06);
    p = __builtin_fma(p, z, -0.0001984126984126984);
    p = __builtin_fma(p, z, 0.008333333333333333);
    p = __builtin_fma(p, z, -0.16666666666666666);
    s = __builtin_fma(x * z, p, x);
    double q = 4.779477332387385e-14;
    q = __builtin_fma(q, z, -1.1470745597729725e-11);
    q = __builtin_fma(q, z, 2.08767569878681e-09);
    q = __builtin_fma(q, z, -2.755731922398589e-07);
    q = __builtin_fma(q, z, 2.48015873015873e-05);
    q = __builtin_fma(q, z, -0.001388888888888889);
    q = __builtin_fma(q, z, 0.041666666666666664);
    q = __builtin_fma(q, z, -0.5);
    c = __builtin_fma(z, q, 1.0);
}

__device__ __forceinline__ void sc_poly(float x, float &s, float &c)
{
    float z = x * x;
    float p = 2.7557318844628753e-06f;
    p = __builtin_fmaf(p, z, -0.00019841270113829523f);
    p = __builtin_fmaf(p, z, 0.008333333767950535f);
    p = __builtin_fmaf(p, z, -0.1666666716337204f);
    s = __builtin_fmaf(x * z, p, x);
    float q = -2.755731998149713e-07f;
    q = __builtin_fmaf(q, z, 2.4801587642286904e-05f);
    q = __builtin_fmaf(q, z, -0.0013888889225199819f);
    q = __builtin_fmaf(q, z, 0.0416666679084301f);
    q = __builtin_fmaf(q, z, -0.5f);
    c = __builtin_fmaf(z, q, 1.0f);
}

template <typename R>
__device__ __forceinline__ void quadrant(int q, R s, R c, R &so, R &co)
{
    q &= 3;
    R a = (q & 1) ? c : s;   // sin(x + q*pi/2)
    R b = (q & 1) ? s : c;   // cos(x + q*pi/2) up to sign
    so = (q >= 2) ? -a : a;
    co = (q == 1 || q == 2) ? -b : b;
}

// sin(2*pi*t), cos(2*pi*t), t in [0,1].
__device__ __forceinline__ void sincos_turns(double t, double &s, double &c)
{
    double k = __builtin_floor(4.0 * t + 0.5);
    double r = t - k * 0.25;
    double x = r * 6.283185307179586;
    double sx, cx;
    sc_poly(x, sx, cx);
    quadrant<double>((int)k, sx, cx, s, c);
}

__device__ __forceinline__ void sincos_turns(float t, float &s, float &c)
{
    float k = __builtin_floorf(4.0f * t + 0.5f);
    float r = t - k * 0.25f;
    float x = r * 6.2831854820251465f;
    float sx, cx;
    sc_poly(x, sx, cx);
    quadrant<float>((int)k, sx, cx, s, c);
}

// sin(2*pi*t), cos(2*pi*t) for t = u01_f32(w) = m * 2^-24, m = (w >> 8) | 1,
// from a table of 256 angles: k = round(256 t) (as the integer (m + 2^15) >>
// 16), x = 2*pi*(t - k/256) = 2*pi*(m - k*2^16)*2^-24 with |x| <= pi/256, and
// the angle sum with cos x = 1 - x^2/2, sin x = x - x^3/6 (truncation below
// 1e-9).  tab = {sin, cos} of 2*pi*k/256 (rms_turn_table.h, staged in LDS).
__device__ __forceinline__ void sincos_turns_table(uint32_t w, const float2 *tab, float &s, float &c)
{
    const uint32_t m = (w >> 8) | 1u;
    const uint32_t k = (m + (1u << 15)) >> 16;
    const int32_t num = (int32_t)m - (int32_t)(k << 16);
    const float x = (float)num * (6.2831854820251465f * 5.9604644775390625e-08f);
    const float2 a = tab[k & 255u];
    const float x2 = x * x;
    const float cr = __builtin_fmaf(x2, -0.5f, 1.0f);
    const float sr = __builtin_fmaf(x * x2, -0.16666667163372040f, x);
    s = __builtin_fmaf(a.x, cr, a.y * sr);
    c = __builtin_fmaf(a.y, cr, -(a.x * sr));
}

// Correctly rounded square root of a normal x far from the ends of the float
// range (the direction draw takes it of u(1-u) in [2^-25, 1/4]): the hardware
// estimate, within one ulp, moves to the neighbour that the square residuals
// x - s_down*s and x - s_up*s show to be the rounded root.
__device__ __forceinline__ float sqrt_rn_mid(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    const float r = rd <= 0.0f ? sd : s;
    return ru > 0.0f ? su : r;
}

// exp(x), x <= 0: Cody-Waite reduction by ln 2, Taylor to r^13.
__device__ __forceinline__ double exp_neg(double x)
{
    if (!(x > -708.0)) return 0.0;
    double k = __builtin_floor(x * 1.4426950408889634 + 0.5);
    double r = __builtin_fma(-k, 0.6931471803691238, x);
    r = __builtin_fma(-k, 1.9082149292705877e-10, r);
    double p = 1.6059043836821613e-10;
    p = __builtin_fma(p, r, 2.08767569878681e-09);
    p = __builtin_fma(p, r, 2.505210838544172e-08);
    p = __builtin_fma(p, r, 2.755731922398589e-07);
    p = __builtin_fma(p, r, 2.7557319223985893e-06);
    p = __builtin_fma(p, r, 2.48015873015873e-05);
    p = __builtin_fma(p, r, 0.0001984126984126984);
    p = __builtin_fma(p, r, 0.001388888888888889);
    p = __builtin_fma(p, r, 0.008333333333333333);
    p = __builtin_fma(p, r, 0.041666666666666664);
    p = __builtin_fma(p, r, 0.16666666666666666);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    long long e = (long long)k + 1023;
    double scale = __longlong_as_double(e << 52);
    return p * scale;
}

// sin and cos of an arbitrary argument: 3-part Cody-Waite reduction by pi/2.
__device__ __forceinline__ void sincos_any(double x, double &s, double &c)
{
    double k = __builtin_floor(x * 0.6366197723675814 + 0.5);
    double r = __builtin_fma(-k, 1.5707963267341256, x);
    r = __builtin_fma(-k, 6.077100506303966e-11, r);
    r = __builtin_fma(-k, 2.0222662487959506e-21, r);
    double sx, cx;
    sc_poly(r, sx, cx);
    quadrant<double>((int)(((long long)k) & 3), sx, cx, s, c);
}

__device__ __forceinline__ double cos_any(double x)
{
    double s, c;
    sincos_any(x, s, c);
    return c;
}

// ---------------------------------------------------------------- wave sums
// Butterfly over the 64 lanes: every lane ends with the same pairwise tree
// ((x0+x1)+(x2+x3))+... (a+b == b+a bit for bit), which is the order the host
// restatement sums a batch in.
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x = x + __shfl_xor(x, m, 64);
    return x;
}

// The value of lane (lane ^ M), for a full wavefront: DPP quad permutations
// for M = 1, 2 (VALU), ds_swizzle's xor mode inside each 32-lane half for
// M = 4, 8, 16, ds_bpermute across the halves for M = 32.
template <int M>
__device__ __forceinline__ int xlane_i32(int v)
{
    static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32, "lane distance");
    if constexpr (M == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    else if constexpr (M == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // 2,3,0,1
    else if constexpr (M == 32) return __shfl_xor(v, 32, 64);
    else return __builtin_amdgcn_ds_swizzle(v, 0x1F | (M << 10));  // and 0x1F, or 0, xor M
}

template <int M>
__device__ __forceinline__ double xlane(double x)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)xlane_i32<M>((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)xlane_i32<M>((int)(uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x = x + (uint32_t)__shfl_xor((int)x, m, 64);
    return x;
}

}  // namespace rms
