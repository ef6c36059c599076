// tt_math.h -- deterministic scalar math shared by the gfx950 kernels and the
// CPU restatement of the terrain step.
//
// Every function here is built from IEEE-754 basic operations only (+, -, *, /,
// sqrt, fused multiply-add, float<->double conversion, bit manipulation).  Both compilers are run
// with -ffp-contract=off and without fast-math, so a call produces the same bits
// on the host (g++, SSE2) and on the device (hipcc, gfx950).  That is what lets
// float heights stay bit-identical between the GPU path and the CPU oracle over
// hundreds of steps, even though the step uses expf / log10f / sinf.
//
// The reference calls CUDA's device libm (expf at plate_tectonics_kernel.cu:1654,
// log10 at :1649, sinf at :2547, fmod at :1654,1712).  Those are not reproduced
// here: each function below is written from a published method and is checked
// in tests against the host libm within a stated error bound.
//   * exp   : Cody-Waite reduction by ln2 (fdlibm split constants), Taylor
//             polynomial of degree 13 on |r| <= ln2/2, evaluated in double.
//   * log   : x = m * 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with
//             s = (m-1)/(m+1), odd series to s^23, evaluated in double.
//   * sin   : Cody-Waite reduction by pi/2 (three 33-bit fdlibm pieces, exact
//             for |x| < 2^19 * pi/2), Taylor polynomials on |r| <= pi/4
//             (sine to r^13, cosine to r^14) evaluated with explicit FMAs.
//   * atan2 : octant reduction, atan(u) = pi/4 + atan((u-1)/(u+1)) for
//             u > tan(pi/8), odd Taylor series to u^41.
//   * fmod  : exact repeated subtraction of b*2^k (every subtraction is exact
//             by Sterbenz' lemma), so the result is the exact remainder.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TT_HD __host__ __device__ __forceinline__
#else
#define TT_HD static inline
#endif

// ---------------------------------------------------------------------------
// bit casts and exact helpers
// ---------------------------------------------------------------------------
TT_HD uint32_t tt_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
TT_HD float tt_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
TT_HD uint64_t tt_d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
TT_HD double tt_u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

TT_HD float tt_fabsf(float x) { return tt_u2f(tt_f2u(x) & 0x7fffffffu); }
TT_HD double tt_fabs(double x) { return tt_u2d(tt_d2u(x) & 0x7fffffffffffffffull); }
TT_HD int tt_isnanf(float x) { return (tt_f2u(x) & 0x7fffffffu) > 0x7f800000u; }

// fminf / fmaxf with C99 NaN semantics (a NaN operand yields the other one),
// which is what CUDA's min/max/fminf/fmaxf do (U4: min(NaN, 0) == 0).
// Written with float comparisons only (a NaN compares false), so that a
// constant operand folds to one compare and one select:
//   a NaN, b not  -> b;   b NaN, a not -> a;   both NaN -> b;
//   otherwise the smaller (larger); equal operands (+0 / -0) give b.
TT_HD float tt_fminf(float a, float b)
{
    return (a < b || (b != b && a == a)) ? a : b;
}
TT_HD float tt_fmaxf(float a, float b)
{
    return (a > b || (b != b && a == a)) ? a : b;
}

// round-to-nearest-even of a double to an integer-valued double, |x| < 2^51.
TT_HD double tt_rint_small(double x)
{
    const double magic = 6755399441055744.0;  // 1.5 * 2^52
    const double t = x + magic;  // not folded: (x + c) - c is not an identity in IEEE
    return t - magic;
}

// 2^n as a double for -1022 <= n <= 1023.
TT_HD double tt_pow2i(int n) { return tt_u2d((uint64_t)(n + 1023) << 52); }

// floor of a float, exact (no library call, identical on host and device).
TT_HD float tt_floorf(float x)
{
    const uint32_t u = tt_f2u(x);
    const int e = (int)((u >> 23) & 0xff) - 127;
    if (e >= 23) return x;  // integral, inf or nan
    if (e < 0) {            // |x| < 1
        if ((u & 0x7fffffffu) == 0) return x;
        return (u >> 31) ? -1.0f : 0.0f;
    }
    const uint32_t frac = 0x007fffffu >> e;
    if ((u & frac) == 0) return x;
    uint32_t t = u & ~frac;
    float r = tt_u2f(t);
    if (u >> 31) r = r - 1.0f;
    return r;
}

// ---------------------------------------------------------------------------
// exp
// ---------------------------------------------------------------------------
TT_HD double tt_exp_core(double x)  // x in [-745, 709], result in double
{
    const double log2e = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;  // fdlibm split of ln 2
    const double ln2_lo = 1.90821492927058770002e-10;
    const double k = tt_rint_small(x * log2e);
    const double r = (x - k * ln2_hi) - k * ln2_lo;
    // Taylor coefficients 1/n!, folded at compile time.
    double p = 1.0 / 6227020800.0;           // 1/13!
    p = p * r + 1.0 / 479001600.0;           // 1/12!
    p = p * r + 1.0 / 39916800.0;            // 1/11!
    p = p * r + 1.0 / 3628800.0;             // 1/10!
    p = p * r + 1.0 / 362880.0;              // 1/9!
    p = p * r + 1.0 / 40320.0;               // 1/8!
    p = p * r + 1.0 / 5040.0;                // 1/7!
    p = p * r + 1.0 / 720.0;                 // 1/6!
    p = p * r + 1.0 / 120.0;                 // 1/5!
    p = p * r + 1.0 / 24.0;                  // 1/4!
    p = p * r + 1.0 / 6.0;                   // 1/3!
    p = p * r + 0.5;                         // 1/2!
    p = p * r + 1.0;
    p = p * r + 1.0;
    int n = (int)k;
    // split the scale so that each factor stays a normal double
    if (n < -1000) { p = p * tt_pow2i(-1000); n += 1000; }
    if (n > 1000) { p = p * tt_pow2i(1000); n -= 1000; }
    return p * tt_pow2i(n);
}

TT_HD float tt_expf(float x)
{
    if (tt_isnanf(x)) return x;
    if (x > 89.0f) return tt_u2f(0x7f800000u);
    if (x < -110.0f) return 0.0f;
    return (float)tt_exp_core((double)x);
}

// ---------------------------------------------------------------------------
// log / log10
// ---------------------------------------------------------------------------
TT_HD double tt_log_core(double m, int e)  // ln(m * 2^e), m in [1, 2)
{
    const double sqrt2 = 1.41421356237309504880;
    if (m > sqrt2) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    const double lnm = 2.0 * s * p;
    const double ln2 = 6.93147180559945286227e-01;
    return (double)e * ln2 + lnm;
}

TT_HD float tt_log10f(float x)
{
    const uint32_t u = tt_f2u(x);
    if (tt_isnanf(x)) return x;
    if (u >> 31) {
        if ((u & 0x7fffffffu) == 0) return tt_u2f(0xff800000u);  // -0 -> -inf
        return tt_u2f(0x7fc00000u);                               // negative -> NaN
    }
    if (u == 0) return tt_u2f(0xff800000u);
    if (u == 0x7f800000u) return x;
    int e = (int)((u >> 23) & 0xff);
    uint32_t mant = u & 0x007fffffu;
    if (e == 0) {  // denormal: normalise
        int shift = 0;
        while ((mant & 0x00800000u) == 0) { mant <<= 1; ++shift; }
        mant &= 0x007fffffu;
        e = 1 - shift;
    }
    const double m = tt_u2d(((uint64_t)0x3ff << 52) | ((uint64_t)mant << 29));
    const double ln = tt_log_core(m, e - 127);
    const double inv_ln10 = 4.34294481903251816668e-01;
    return (float)(ln * inv_ln10);
}

// ---------------------------------------------------------------------------
// sin / cos
// ---------------------------------------------------------------------------
// Fused multiply-add, correctly rounded on both sides (v_fma_f64 on the
// device; the FMA instruction or libm's fma on the host).  Contraction is off,
// so an FMA appears only where it is written.
TT_HD double tt_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

TT_HD double tt_sin_poly(double r)  // |r| <= pi/4; truncation error < 2^-45 relative
{
    // sin r = r + r z P(z), P(z) = -1/3! + z/5! - z^2/7! + z^3/9! - z^4/11! + z^5/13!
    const double z = r * r;
    double p = 1.0 / 6227020800.0;           // +1/13!
    p = tt_fma(p, z, -1.0 / 39916800.0);     // -1/11!
    p = tt_fma(p, z, 1.0 / 362880.0);        // +1/9!
    p = tt_fma(p, z, -1.0 / 5040.0);         // -1/7!
    p = tt_fma(p, z, 1.0 / 120.0);           // +1/5!
    p = tt_fma(p, z, -1.0 / 6.0);            // -1/3!
    return tt_fma(r * z, p, r);
}

TT_HD double tt_cos_poly(double r)  // |r| <= pi/4; truncation error < 2^-49
{
    // cos r = 1 - z (1/2 - z/4! + z^2/6! - ... + z^6/14!)
    const double z = r * r;
    double p = -1.0 / 87178291200.0;         // -1/14!
    p = tt_fma(p, z, 1.0 / 479001600.0);     // +1/12!
    p = tt_fma(p, z, -1.0 / 3628800.0);      // -1/10!
    p = tt_fma(p, z, 1.0 / 40320.0);         // +1/8!
    p = tt_fma(p, z, -1.0 / 720.0);          // -1/6!
    p = tt_fma(p, z, 1.0 / 24.0);            // +1/4!
    p = tt_fma(p, z, -0.5);                  // -1/2
    return tt_fma(z, p, 1.0);
}

// returns quadrant in *q and the reduced argument
TT_HD double tt_reduce_pio2(double x, int* q)
{
    const double two_over_pi = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;   // first 33 bits of pi/2
    const double pio2_2 = 6.07710050630396597660e-11;   // next 33 bits
    const double pio2_3 = 2.02226624871116645580e-21;   // next 33 bits
    const double k = tt_rint_small(x * two_over_pi);
    const double r = ((x - k * pio2_1) - k * pio2_2) - k * pio2_3;
    const long long ki = (long long)k;
    *q = (int)(ki & 3);
    return r;
}

// one polynomial per call: the sine polynomial for even quadrants, the
// cosine polynomial for odd ones, then the sign of the quadrant
TT_HD double tt_sin_d(double x)
{
    int q;
    const double r = tt_reduce_pio2(x, &q);
    const double v = (q & 1) ? tt_cos_poly(r) : tt_sin_poly(r);
    return (q & 2) ? -v : v;
}

TT_HD double tt_cos_d(double x)
{
    int q;
    const double r = tt_reduce_pio2(x, &q);
    const double v = (q & 1) ? tt_sin_poly(r) : tt_cos_poly(r);
    return ((q + 1) & 2) ? -v : v;
}

TT_HD float tt_sinf(float x)
{
    if (tt_isnanf(x)) return x;
    if ((tt_f2u(x) & 0x7fffffffu) == 0x7f800000u) return tt_u2f(0x7fc00000u);
    return (float)tt_sin_d((double)x);
}

TT_HD float tt_cosf(float x)
{
    if (tt_isnanf(x)) return x;
    if ((tt_f2u(x) & 0x7fffffffu) == 0x7f800000u) return tt_u2f(0x7fc00000u);
    return (float)tt_cos_d((double)x);
}

// ---------------------------------------------------------------------------
// atan2 (double)
// ---------------------------------------------------------------------------
TT_HD double tt_atan_small(double u)  // |u| <= tan(pi/8)
{
    const double z = u * u;
    double p = 1.0 / 41.0;
    p = -(p * z) + 1.0 / 39.0;
    p = -(p * z) + 1.0 / 37.0;
    p = -(p * z) + 1.0 / 35.0;
    p = -(p * z) + 1.0 / 33.0;
    p = -(p * z) + 1.0 / 31.0;
    p = -(p * z) + 1.0 / 29.0;
    p = -(p * z) + 1.0 / 27.0;
    p = -(p * z) + 1.0 / 25.0;
    p = -(p * z) + 1.0 / 23.0;
    p = -(p * z) + 1.0 / 21.0;
    p = -(p * z) + 1.0 / 19.0;
    p = -(p * z) + 1.0 / 17.0;
    p = -(p * z) + 1.0 / 15.0;
    p = -(p * z) + 1.0 / 13.0;
    p = -(p * z) + 1.0 / 11.0;
    p = -(p * z) + 1.0 / 9.0;
    p = -(p * z) + 1.0 / 7.0;
    p = -(p * z) + 1.0 / 5.0;
    p = -(p * z) + 1.0 / 3.0;
    p = -(p * z) + 1.0;
    return u * p;
}

TT_HD double tt_atan_unit(double t)  // t in [0, 1]
{
    const double tan_pi_8 = 4.14213562373095034e-01;
    const double pi_4 = 7.85398163397448278999e-01;
    if (t > tan_pi_8) return pi_4 + tt_atan_small((t - 1.0) / (t + 1.0));
    return tt_atan_small(t);
}

TT_HD double tt_atan2(double y, double x)
{
    const double pi = 3.14159265358979311600e+00;
    const double pi_2 = 1.57079632679489655800e+00;
    const double ax = tt_fabs(x), ay = tt_fabs(y);
    if (ax == 0.0 && ay == 0.0) return 0.0;
    double a;
    if (ay <= ax) a = tt_atan_unit(ay / ax);
    else a = pi_2 - tt_atan_unit(ax / ay);
    if (x < 0.0) a = pi - a;
    if (y < 0.0) a = -a;
    return a;
}

// ---------------------------------------------------------------------------
// fmod (exact)
// ---------------------------------------------------------------------------
TT_HD float tt_fmodf(float a, float b)
{
    if (tt_isnanf(a) || tt_isnanf(b)) return tt_u2f(0x7fc00000u);
    const float aa = tt_fabsf(a), bb = tt_fabsf(b);
    if (bb == 0.0f || aa == tt_u2f(0x7f800000u)) return tt_u2f(0x7fc00000u);
    if (aa < bb) return a;
    // r := |a| mod |b| by subtracting the largest |b| * 2^k <= r, exactly.
    float r = aa;
    while (r >= bb) {
        float s = bb;
        while (s <= r * 0.5f) s = s * 2.0f;  // s * 2 <= r, no overflow since r finite
        r = r - s;
    }
    return (tt_f2u(a) >> 31) ? -r : r;
}

// ---------------------------------------------------------------------------
// C++20 std::lerp (libc++ / libcu++ algorithm, used at
// plate_tectonics_kernel.cu:356): exact at the ends, monotone.
// ---------------------------------------------------------------------------
TT_HD float tt_lerp(float a, float b, float t)
{
    if ((a <= 0.0f && b >= 0.0f) || (a >= 0.0f && b <= 0.0f)) return t * b + (1.0f - t) * a;
    if (t == 1.0f) return b;
    const float x = a + t * (b - a);
    if ((t > 1.0f) == (b > a)) return b < x ? x : b;
    return x < b ? x : b;
}

// float -> unsigned int conversion with the saturating behaviour of the GPU
// instruction (negative and NaN -> 0, >= 2^32 -> 0xffffffff); the C cast is
// undefined outside [0, 2^32).
TT_HD uint32_t tt_f2u32_sat(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

// ---------------------------------------------------------------------------
// Exact order-free accumulation: a float is turned into a fixed-point int64
// (round to nearest even of v * 2^bits); integer sums are associative, so a
// parallel sum in any order gives the same bits as a serial loop.  Inputs
// outside [-2^(62-bits), 2^(62-bits)] saturate and NaN counts as 0.
// ---------------------------------------------------------------------------
TT_HD long long tt_fix(float v, int bits)
{
    if (tt_isnanf(v)) return 0;
    const double d = (double)v * tt_pow2i(bits);
    const double lim = 4611686018427387904.0;  // 2^62
    if (d >= lim) return 4611686018427387904LL;
    if (d <= -lim) return -4611686018427387904LL;
    // round half to even (rint in the default rounding mode) of a double that
    // already has all bits of v; |d| < 2^62 converts exactly
    return (long long)__builtin_rint(d);
}

TT_HD long long tt_fix_d(double v, int bits)
{
    if (v != v) return 0;
    const double d = v * tt_pow2i(bits);
    const double lim = 4611686018427387904.0;  // 2^62
    if (d >= lim) return 4611686018427387904LL;
    if (d <= -lim) return -4611686018427387904LL;
    return (long long)__builtin_rint(d);
}

TT_HD float tt_unfix(long long q, int bits) { return (float)((double)q * tt_pow2i(-bits)); }
TT_HD double tt_unfix_d(long long q, int bits) { return (double)q * tt_pow2i(-bits); }
