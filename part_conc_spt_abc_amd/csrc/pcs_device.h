// pcs_device.h -- gfx950 device primitives of the walk kernel: Philox4x32-10,
// the project's FP32 log / sqrt / sincos-in-turns, and the per-step walk,
// wrap, detect and track logic.
//
// Every function here fixes its FP32 operation order: the library is built
// with -ffp-contract=off and correctly rounded division/sqrt, fmaf() is the
// single-rounding fused multiply-add, so results are reproducible bit for bit
// on any IEEE-754 host that evaluates the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcs {

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): multipliers and Weyl
// key increments of the paper's Philox4x32 parameter set.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct U4 {
    uint32_t x, y, z, w;
};

// Three-way XOR in one gfx950 VALU instruction (v_bitop3_b32, truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ uint32_t mulhi(uint32_t m, uint32_t x) { return (uint32_t)(((uint64_t)m * x) >> 32); }

// Counter words block (c0), video (c2) and simulation (c3) and the key are
// wave-uniform; only the particle word c1 varies per lane. Rounds 1-3 are
// written out so that the uniform halves stay scalar (plain ^, s_mul) and
// only lane-varying words use the fused v_bitop3 XOR; from round 4 on every
// word varies per lane. Same function as the textbook round loop.
__device__ __forceinline__ U4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1)
{
    // round 1: c1 varies
    uint32_t x0 = xor3(mulhi(kPhiloxM1, c2), c1, k0);  // varying
    uint32_t x1 = kPhiloxM1 * c2;                       // uniform
    uint32_t x2 = mulhi(kPhiloxM0, c0) ^ c3 ^ k1;       // uniform
    uint32_t x3 = kPhiloxM0 * c0;                       // uniform
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 2: x0 varies
    uint32_t y0 = mulhi(kPhiloxM1, x2) ^ x1 ^ k0;       // uniform
    uint32_t y1 = kPhiloxM1 * x2;                       // uniform
    uint32_t y2 = xor3(mulhi(kPhiloxM0, x0), x3, k1);   // varying
    uint32_t y3 = kPhiloxM0 * x0;                       // varying
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 3: y2, y3 vary
    c0 = xor3(mulhi(kPhiloxM1, y2), y1, k0);
    c1 = kPhiloxM1 * y2;
    c2 = xor3(mulhi(kPhiloxM0, y0), y3, k1);
    c3 = kPhiloxM0 * y0;  // uniform
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return U4{c0, c1, c2, c3};
}

// --------------------------------------------------------------- FP32 math
// ln(1+f) = f + f^2 P(f) on [sqrt(1/2)-1, sqrt(2)-1]; coefficients from
// fit_math_coeffs.py (reweighted least squares towards minimax), rounded to
// FP32. ln 2 is split (Cody & Waite) so E * kLn2Hi is exact for |E| < 256.
constexpr float kLogP0 = -0x1.fffff4p-2f, kLogP1 = 0x1.5556e8p-2f, kLogP2 = -0x1.0006c6p-2f,
                kLogP3 = 0x1.98da1cp-3f, kLogP4 = -0x1.52fb0cp-3f, kLogP5 = 0x1.30d166p-3f,
                kLogP6 = -0x1.277654p-3f, kLogP7 = 0x1.6fccd4p-4f;
constexpr float kLn2Hi = 0x1.62e4p-1f;
constexpr float kLn2Lo = 0x1.7f7d1cp-20f;
constexpr float kSqrtHalf = 0x1.6a09e6p-1f;
// sin(2 pi t) = t S(t^2), cos(2 pi t) = C(t^2) for |t| <= 1/8 turn.
constexpr float kSin0 = 0x1.921fb6p+2f, kSin1 = -0x1.4abbbap+5f, kSin2 = 0x1.465ec6p+6f,
                kSin3 = -0x1.2d9bcep+6f;
constexpr float kCos1 = -0x1.3bd3ccp+4f, kCos2 = 0x1.03c1dep+6f, kCos3 = -0x1.55c664p+6f,
                kCos4 = 0x1.d9f7bcp+5f;

// -2 ln(u) for u = ((w >> 8) + 1) 2^-24 in (0, 1].
__device__ __forceinline__ float neg2log_u24(uint32_t w)
{
    const uint32_t j = (w >> 8) + 1u;
    const float xf = (float)j;  // exact: j <= 2^24
    int e;
    float m = __builtin_frexpf(xf, &e);  // v_frexp_mant / v_frexp_exp
    const bool lo = m < kSqrtHalf;
    m = lo ? m * 2.0f : m;
    e = lo ? e - 1 : e;
    const float f = m - 1.0f;  // exact (Sterbenz)
    const float E = (float)(e - 24);
    const float f2 = f * f;
    float P = kLogP7;
    P = __builtin_fmaf(P, f, kLogP6);
    P = __builtin_fmaf(P, f, kLogP5);
    P = __builtin_fmaf(P, f, kLogP4);
    P = __builtin_fmaf(P, f, kLogP3);
    P = __builtin_fmaf(P, f, kLogP2);
    P = __builtin_fmaf(P, f, kLogP1);
    P = __builtin_fmaf(P, f, kLogP0);
    const float l1p = __builtin_fmaf(f2, P, f);
    const float L = __builtin_fmaf(E, kLn2Hi, __builtin_fmaf(E, kLn2Lo, l1p));
    return __builtin_fmaf(-2.0f, L, 0.0f);
}

// Correctly rounded sqrt for y = 0 or 2^-96 <= y < inf: the hardware
// square root (v_sqrt_f32, within 1 ulp) followed by the exact test of its
// two neighbours, y - s'*s against zero with one fused multiply-add each,
// which selects the correctly rounded result. The general lowering also
// rescales tiny inputs and passes infinities/NaNs through; -2 ln u is either
// 0 or at least 1.19e-7 and finite, so those steps are not needed here.
__device__ __forceinline__ float sqrt_rn_domain(float y)
{
    const float s = __builtin_amdgcn_sqrtf(y);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, y);
    const float ru = __builtin_fmaf(-su, s, y);
    float r = (rd <= 0.0f) ? sd : s;
    r = (ru > 0.0f) ? su : r;
    return r;
}

// Box-Muller radius sqrt(-2 ln u), correctly rounded.
__device__ __forceinline__ float bm_radius(uint32_t w) { return sqrt_rn_domain(neg2log_u24(w)); }

// sin and cos of 2 pi t, t = a 2^-24 for a 24-bit integer a: exact reduction
// to the nearest quarter turn in integers, polynomial on |t| <= 1/8, quadrant
// rotation by swap and sign flips.
__device__ __forceinline__ void sincos_turn24(uint32_t a, float& s_out, float& c_out)
{
    const uint32_t j = a + (1u << 21);
    const uint32_t q = (j >> 22) & 3u;
    const int32_t r = (int32_t)(j & ((1u << 22) - 1u)) - (1 << 21);
    const float t = (float)r * 0x1p-24f;  // exact
    const float t2 = t * t;
    float sp = kSin3;
    sp = __builtin_fmaf(sp, t2, kSin2);
    sp = __builtin_fmaf(sp, t2, kSin1);
    sp = __builtin_fmaf(sp, t2, kSin0);
    const float s = t * sp;
    float c = kCos4;
    c = __builtin_fmaf(c, t2, kCos3);
    c = __builtin_fmaf(c, t2, kCos2);
    c = __builtin_fmaf(c, t2, kCos1);
    c = __builtin_fmaf(c, t2, 1.0f);
    const bool swap = (q & 1u) != 0u;
    const float so = swap ? c : s;
    const float co = swap ? s : c;
    const uint32_t sb = (q & 2u) << 30;
    const uint32_t cb = ((q + 1u) & 2u) << 30;
    s_out = __uint_as_float(__float_as_uint(so) ^ sb);
    c_out = __uint_as_float(__float_as_uint(co) ^ cb);
}

// The same for t = (w >> 8) 2^-24 (the top 24 bits of a word).
__device__ __forceinline__ void sincos_turn_u24(uint32_t w, float& s_out, float& c_out)
{
    sincos_turn24(w >> 8, s_out, c_out);
}

// Angle of Box-Muller pair p of a chunk, 16 bits as a 24-bit turn fraction:
// the low byte of the pair's radius word (whose top 24 bits give the radius
// uniform) above byte p % 4 of the chunk's angle word.
__device__ __forceinline__ uint32_t bm_angle24(uint32_t w_radius, uint32_t w_angle, uint32_t p)
{
    return ((w_radius << 16) & 0xff0000u) | ((w_angle >> (8u * (p & 3u)) << 8) & 0xff00u);
}

__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

// Single periodic wrap of a centred coordinate (the reference's
// periodic_bounds_: src/gpu/cusim/kernel-simulate-unroll.inl:24-32).
__device__ __forceinline__ float wrap_periodic(float v, float h, float L)
{
    return (__builtin_fabsf(v) > h) ? v - __builtin_copysignf(L, v) : v;
}
// Reflecting wall: v' = +-L - v beyond +-L/2.
__device__ __forceinline__ float wrap_reflect(float v, float h, float L)
{
    return (__builtin_fabsf(v) > h) ? __builtin_copysignf(L, v) - v : v;
}

}  // namespace pcs
