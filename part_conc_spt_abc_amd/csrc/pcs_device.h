// pcs_device.h -- gfx950 device primitives of the walk kernel: Philox4x32-10
// and the project's FP32 math: the Box-Muller log, sqrt and table sin/cos, and
// the bound of the Gaussian detection draw.
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

// Philox4x32-10 of the counter (video, particle, sim, blk). video, sim and the
// key are fixed for a tile and wave-uniform, the particle word varies per lane
// and is fixed for a tile, blk is wave-uniform and changes from chunk to chunk.
// Rounds 1-4 are written out: rounds 1 and 2 mix the lane word with tile
// constants only (the compiler hoists them out of the chunk loop, so a chunk's
// block pays for them with scalar work), round 3 joins the lane halves with the
// block halves by XOR, round 4 still takes one uniform word, and from round 5
// on every word varies per lane. Uniform XOR operands are grouped so they fold
// on the scalar unit. Same function as the textbook round loop.
__device__ __forceinline__ U4 philox10(uint32_t blk, uint32_t particle, uint32_t video, uint32_t sim,
                                       uint32_t k0, uint32_t k1)
{
    // round 1: counter (video, particle, sim, blk)
    const uint64_t a0 = (uint64_t)kPhiloxM0 * video;  // uniform
    const uint64_t a1 = (uint64_t)kPhiloxM1 * sim;    // uniform
    const uint32_t x0 = particle ^ ((uint32_t)(a1 >> 32) ^ k0);  // lane, tile constant
    const uint32_t x1 = (uint32_t)a1;                            // uniform
    const uint32_t x2 = (uint32_t)(a0 >> 32) ^ blk ^ k1;         // uniform, per chunk
    const uint32_t x3 = (uint32_t)a0;                            // uniform
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 2
    const uint64_t b0 = (uint64_t)kPhiloxM0 * x0;                // lane, tile constant
    const uint64_t b1 = (uint64_t)kPhiloxM1 * x2;                // uniform, per chunk
    const uint32_t y0 = (uint32_t)(b1 >> 32) ^ x1 ^ k0;          // uniform, per chunk
    const uint32_t y1 = (uint32_t)b1;                            // uniform, per chunk
    const uint32_t y2 = (uint32_t)(b0 >> 32) ^ (x3 ^ k1);        // lane, tile constant
    const uint32_t y3 = (uint32_t)b0;                            // lane, tile constant
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 3
    const uint64_t c0p = (uint64_t)kPhiloxM0 * y0;               // uniform, per chunk
    const uint64_t c1p = (uint64_t)kPhiloxM1 * y2;               // lane, tile constant
    uint32_t c0 = (uint32_t)(c1p >> 32) ^ (y1 ^ k0);
    uint32_t c1 = (uint32_t)c1p;
    uint32_t c2 = y3 ^ ((uint32_t)(c0p >> 32) ^ k1);
    uint32_t c3 = (uint32_t)c0p;                                 // uniform, per chunk
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 4: c3 is still uniform, so its XOR with the key is scalar work
    {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ (c3 ^ k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
#pragma unroll
    for (int r = 4; r < 10; ++r) {
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
// FP32. -2 ln 2 rounded to FP32 is within 2.8e-9 relative of the exact value,
// so E (-2 ln 2) for |E| <= 24 is within 0.03 ulp of any result it enters.
constexpr float kLogP0 = -0x1.000060p-1f, kLogP1 = 0x1.5559f0p-2f, kLogP2 = -0x1.ff7476p-3f,
                kLogP3 = 0x1.976dd6p-3f, kLogP4 = -0x1.6049bep-3f, kLogP5 = 0x1.4d64c0p-3f,
                kLogP6 = -0x1.a7da34p-4f;
constexpr float kNeg2Ln2 = -0x1.62e43p+0f;
constexpr float kSqrtHalf = 0x1.6a09e6p-1f;
constexpr uint32_t kSqrtHalfBits = 0x3f3504f3u;  // bits of kSqrtHalf
static_assert(__builtin_bit_cast(uint32_t, kSqrtHalf) == kSqrtHalfBits, "bits of sqrt(1/2)");
// The series in g = -2 (m - 1): -2 ln(1+f) = g + g^2 Q(g) with Q(g) = -P(-g/2)/2,
// whose coefficients are P's times powers of two (exact in FP32).
constexpr float kLogQ0 = kLogP0 * -0.5f, kLogQ1 = kLogP1 * 0.25f, kLogQ2 = kLogP2 * -0.125f,
                kLogQ3 = kLogP3 * 0.0625f, kLogQ4 = kLogP4 * -0.03125f, kLogQ5 = kLogP5 * 0.015625f,
                kLogQ6 = kLogP6 * -0.0078125f;
constexpr float kLogTiny = 0x1p-100f;

// -2 ln(u) for the uniform u = x 2^-32 of the top 24 bits of w, with
// x = (float)(w | 255) (rounded to nearest even): x = 256 J + 255 exactly for
// J = w >> 8 < 2^16 and x = 256 (J + 1) above, so u is in [255 2^-32, 1] and
// positive for every w; the low byte of w does not enter. With m, E as below,
// g = -2 (m - 1) = fma(m, -2, 2) is exact and every Horner step in g is the
// corresponding step in f = m - 1 scaled by a power of two, so the result is
// -2 (E ln 2 + f + f^2 P(f)) with the same roundings (within 2.4 ulp of the
// exact value; the radius too). g^2 carries an added 2^-100, which changes the
// result only where u = 1 (then 2^-100 Q(0), about 2^-102, instead of 0, so
// the square root below never sees 0) and where g^2 is exactly halfway
// between two floats (rounded up instead of to even).
__device__ __forceinline__ float neg2log_u24(uint32_t w)
{
    // x = m 2^e with m in [sqrt(1/2), sqrt(2)), in integer steps on the bits
    // of x: t = bits(x) - bits(sqrt(1/2)) - (32 << 23) carries e - 32 in its
    // top 9 bits (arithmetic shift) and m's bits are the low 23 bits of t plus
    // bits(sqrt(1/2)) (a borrow from the mantissa lands m in [1, sqrt 2)).
    // Same m and e as frexp followed by the sqrt(1/2) fold-up.
    const int32_t t = (int32_t)(__float_as_uint((float)(w | 0xffu)) - (kSqrtHalfBits + (32u << 23)));
    const float m = __uint_as_float(((uint32_t)t & 0x7fffffu) + kSqrtHalfBits);
    const float g = __builtin_fmaf(m, -2.0f, 2.0f);  // exact (Sterbenz)
    const float E = (float)(t >> 23);
    const float g2 = __builtin_fmaf(g, g, kLogTiny);
    float Q = kLogQ6;
    Q = __builtin_fmaf(Q, g, kLogQ5);
    Q = __builtin_fmaf(Q, g, kLogQ4);
    Q = __builtin_fmaf(Q, g, kLogQ3);
    Q = __builtin_fmaf(Q, g, kLogQ2);
    Q = __builtin_fmaf(Q, g, kLogQ1);
    Q = __builtin_fmaf(Q, g, kLogQ0);
    return __builtin_fmaf(E, kNeg2Ln2, __builtin_fmaf(g2, Q, g));
}

// 2^f = 1 + f Q(f) on [0, 1); Q from fit_math_coeffs.py (within 1.07 ulp of 2^f,
// and 2^0 = 1 exactly).
constexpr float kExp2Q0 = 0x1.62e42cp-1f, kExp2Q1 = 0x1.ebfd58p-3f, kExp2Q2 = 0x1.c68912p-5f,
                kExp2Q3 = 0x1.3d0b92p-7f, kExp2Q4 = 0x1.46d64cp-10f, kExp2Q5 = 0x1.c54176p-13f;
constexpr float kLog2E = 0x1.715476p0f;

// Gaussian detection in z (the model's extension of the slab): a particle at
// z (from the centre, in the unit of s) is detected with probability
// exp(-z^2 / (2 s^2)), drawn with the uniform u = x 2^-32 of the top 24 bits
// of w, x = (float)(w | 255) (as for the Box-Muller radius), as
// x <= 2^32 exp(-z^2 / (2 s^2)) = 2^(32 - c z^2), c = log2(e) / (2 s^2)
// (gauss_c). 32 - c z^2 = n + f with n = floor (so f in [0, 1) is exact), 2^f
// by the polynomial, scaled by 2^n; at z = 0 the bound is 2^32 and detection
// is certain. Below -128 the exponent is clamped: 2^-128 is already under
// every x, and n stays in the range of the conversion to int.
__device__ __forceinline__ float gauss_c(float sig2) { return kLog2E / (sig2 + sig2); }

__device__ __forceinline__ bool gauss_detect(float z, float c, uint32_t w)
{
    const float v = __builtin_fmaxf(__builtin_fmaf(-c, z * z, 32.0f), -128.0f);
    const float n = __builtin_floorf(v);
    const float f = v - n;
    float Q = kExp2Q5;
    Q = __builtin_fmaf(Q, f, kExp2Q4);
    Q = __builtin_fmaf(Q, f, kExp2Q3);
    Q = __builtin_fmaf(Q, f, kExp2Q2);
    Q = __builtin_fmaf(Q, f, kExp2Q1);
    Q = __builtin_fmaf(Q, f, kExp2Q0);
    const float t = __builtin_ldexpf(__builtin_fmaf(f, Q, 1.0f), (int)n);
    return (float)(w | 0xffu) <= t;
}

// Correctly rounded sqrt for the Box-Muller radius y = -2 ln u (2^-100 Q(0) =
// 0x1.00006p-102 at u = 1, else 1.19e-7 <= y <= 33.3): s = y rsq(y) with one
// correction by the residual y - s^2 (Markstein). It equals the correctly
// rounded square root on every float in [2^-102, 64) (checked exhaustively on
// the GPU, each result against the midpoints to its neighbours squared exactly
// in FP64; far below, where the residual term is subnormal, it misses); the
// hardware square root alone misses on 15 % of the radius inputs.
__device__ __forceinline__ float sqrt_rn_domain(float y)
{
    const float q = __builtin_amdgcn_rsqf(y);
    const float s = y * q;
    const float h = 0.5f * q;
    const float r = __builtin_fmaf(-s, s, y);
    return __builtin_fmaf(r, h, s);
}

// Box-Muller radius sqrt(-2 ln u), correctly rounded.
__device__ __forceinline__ float bm_radius(uint32_t w) { return sqrt_rn_domain(neg2log_u24(w)); }

// sin and cos of the Box-Muller angle 2 pi (a + 1/2) 2^-16 of the code
// a = i << 8 | r (i, r bytes), from the angle tables in LDS: tab[i] = (sin, cos)
// of the sector centre 2 pi (i + 1/2) / 256, tab[256 + r] = (sin beta, cos beta - 1)
// of the rest beta = 2 pi (r - 127.5) 2^-16. sin(a + beta) = (sa + ca sb) + sa (cb - 1):
// the fused sa + ca sb cancels exactly near zero before its one rounding, then
// the small term is added (cos likewise).
__device__ __forceinline__ void sincos_tab(const float2* tab, uint32_t i, uint32_t r, float& s_out, float& c_out)
{
    const float2 ta = tab[i];
    const float2 tb = tab[256u + r];
    s_out = __builtin_fmaf(ta.x, tb.y, __builtin_fmaf(ta.y, tb.x, ta.x));
    c_out = __builtin_fmaf(ta.y, tb.y, __builtin_fmaf(-ta.x, tb.x, ta.y));
}

__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

// Reflecting walls at +-h: v' = copysign(2h, v) - v beyond them (single
// reflection), as 2 med3(v, -h, h) - v in one fma: v itself inside (2v - v is
// exact) and +-2h - v, rounded once, outside - the same value as the
// reflection from L = 2h.
__device__ __forceinline__ float wrap_reflect(float v, float h)
{
    return __builtin_fmaf(2.0f, __builtin_amdgcn_fmed3f(v, -h, h), -v);
}

// ---------------------------------------------------------------- FP64 10^x
// 10^x = 2^k 10^f, k = rint(x log2 10), f = x - k log10 2 with log10 2 split
// into a 42-bit head (k * head exact for |k| < 2^11) and a tail; 10^f = e^y,
// y = f ln 10 (|y| <= 0.35), by its Taylor series to degree 16 (remainder below
// 1e-23 relative) in fma Horner form; within 2 ulp. Fixed operation order, so
// the same on any IEEE-754 host.
constexpr double kExp10Log2_10 = 0x1.a934f0979a371p+1;
constexpr double kExp10Log10_2Hi = 0x1.34413509f7800p-2;
constexpr double kExp10Log10_2Lo = 0x1.fef311f12b358p-46;
constexpr double kExp10Ln10 = 0x1.26bb1bbb55516p+1;

__device__ __forceinline__ double exp10_d(double x)
{
    constexpr double T[17] = {
        0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
        0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
        0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
        0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41,
        0x1.ae7f3e733b81fp-45,
    };
    const double k = __builtin_rint(x * kExp10Log2_10);
    double f = __builtin_fma(-k, kExp10Log10_2Hi, x);
    f = __builtin_fma(-k, kExp10Log10_2Lo, f);
    const double y = f * kExp10Ln10;
    double p = T[16];
#pragma unroll
    for (int i = 15; i >= 0; --i) p = __builtin_fma(p, y, T[i]);
    return __builtin_ldexp(p, (int)k);
}

}  // namespace pcs
