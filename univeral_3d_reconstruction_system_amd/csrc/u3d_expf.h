// Single-precision exp that is bit-identical, on every finite and special input, to the host
// C library's expf (glibc >= 2.28 on x86-64 with FMA). The reference trainer calls std::exp on
// floats (engine/src/splat/Rasterizer.cpp:71, :196, :309, :389; GaussianModel.h:11); using the
// same function on the device keeps radii, alpha cut-offs, early termination and therefore
// whole rendered images bit-exact with the reference.
//
// Method: the table-driven expf published in ARM's optimized-routines (Szabolcs Nagy, 2017),
// which glibc adopted: x·32/ln2 = k + r with integer k and |r| <= 1/2, then
// exp(x) = 2^(k/32) · p(r), where 2^(k/32) comes from a 32-entry table of 2^(i/32) with the
// exponent field pre-biased, and p is a cubic evaluated in double. glibc's FMA build contracts
// the reduction and the polynomial into fused multiply-adds; the fma() calls below do the same.
// The final double -> float conversion is the only rounding that reaches the caller.
// Verified by the test suite against the host expf over all 2^32 bit patterns.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define U3D_HD __host__ __device__
#else
#define U3D_HD
#endif

namespace u3dmath {

// Bits of 2^(i/32) rounded to double, minus (i << 47), so that adding k << 47 for the
// integer k = 32·e + i yields the bits of 2^(k/32).
#define U3D_EXP2_TABLE                                                                         \
    {0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull, \
     0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, \
     0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull, \
     0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull, \
     0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull, \
     0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull, \
     0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull, \
     0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}

U3D_HD inline double bits_to_double(uint64_t u)
{
    union { uint64_t u; double d; } c;
    c.u = u;
    return c.d;
}
U3D_HD inline uint64_t double_to_bits(double d)
{
    union { uint64_t u; double d; } c;
    c.d = d;
    return c.u;
}
U3D_HD inline uint32_t float_to_bits(float f)
{
    union { uint32_t u; float f; } c;
    c.f = f;
    return c.u;
}

// Core of expf for |x| < 88 (and the sub-normal tail down to -103.97); `table` holds the 32
// biased entries above. Kept separate so device code can supply the table from LDS.
U3D_HD inline float expf_core(float x, const uint64_t *table)
{
    const double kInvLn2N = 0x1.71547652b82fep+0 * 32.0;
    const double kShift = 0x1.8p+52;
    const double kP0 = 0x1.c6af84b912394p-5 / (32.0 * 32.0 * 32.0);
    const double kP1 = 0x1.ebfce50fac4f3p-3 / (32.0 * 32.0);
    const double kP2 = 0x1.62e42ff0c52d6p-1 / 32.0;
    const double xd = (double)x;
    double kd = __builtin_fma(kInvLn2N, xd, kShift);
    const uint64_t ki = double_to_bits(kd);
    kd -= kShift;
    const double r = __builtin_fma(kInvLn2N, xd, -kd);
    const uint64_t t = table[ki & 31u] + (ki << 47);
    const double s = bits_to_double(t);
    const double p = __builtin_fma(kP0, r, kP1);
    const double r2 = r * r;
    double y = __builtin_fma(kP2, r, 1.0);
    y = __builtin_fma(p, r2, y);
    return (float)(y * s);
}

// Full expf including the special inputs, with the same results as the host library:
// -inf -> 0, +inf/nan -> x + x, overflow -> +inf, underflow below -103.97 -> +0.
U3D_HD inline float expf_special(float x, bool *handled)
{
    const uint32_t top = (float_to_bits(x) >> 20) & 0x7ffu;
    *handled = false;
    if (top >= 0x42bu) { // |x| >= 88 or nan
        *handled = true;
        if (float_to_bits(x) == 0xff800000u)
            return 0.0f;
        if (top >= 0x7f8u)
            return x + x;
        if (x > 0x1.62e42ep6f)
            return __builtin_inff();
        if (x < -0x1.9fe368p6f)
            return 0.0f;
        *handled = false;
    }
    return 0.0f;
}

} // namespace u3dmath
