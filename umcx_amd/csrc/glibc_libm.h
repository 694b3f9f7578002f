// glibc_libm.h -- bit-exact device replicas of the host C library's single-precision
// logf, sinf, cosf, acosf and expf, plus nextafterf.
//
// Why: umcx's photon trajectories go through libm (src/umcx.cpp:213 logf, :437-438 sinf/cosf,
// :232,:372,:374 acosf, :31 nextafterf). glibc 2.35 is not correctly rounded on 0.7-7.7% of the
// arguments the photon loop can reach, so a correctly rounded GPU libm would send almost every
// photon down a different path. These functions reproduce the host library instead, from the
// published designs it implements:
//   * logf, sinf, cosf, expf: the ARM optimized-routines algorithms (Szabolcs Nagy, 2017-2018)
//     that glibc uses since 2.28 -- a 16-entry (invc, log c) table with a cubic in double for
//     logf; double range reduction by pi/2 and even/odd polynomials for sinf/cosf; a 32-entry
//     2^(i/32) table and a cubic for expf. The x86-64 glibc selects its FMA-compiled variants
//     on FMA-capable hosts; the fused forms are used here.
//   * acosf: the fdlibm single-precision algorithm (Sun Microsystems, 1993) that glibc's
//     sysdeps/ieee754/flt-32 keeps: rational approximation of asin on [0, 0.5] and the
//     half-angle reduction above it, in float.
// Verified bit for bit against the host libm on the GPU over every positive normal float for
// logf, every float for expf, sinf and cosf, and every float of [-1, 1] for acosf
// (tests/test_gpu_libm_exhaustive.py), and on the CPU over every argument the photon loop can
// form and sampled floats (tests/test_host_contract.py).
//
// Implemented ranges: logf for positive normal inputs (the loop's u + FLT_EPSILON >= 2^-23),
// sinf/cosf for every argument (sincosf_host, the fast form for |x| < 120 used by scattering;
// sincosf_any for all floats, used where a user-set angle enters), acosf on [-1, 1], expf for
// every float.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define UMCX_HD __host__ __device__ inline
#else
#define UMCX_HD inline
#include <math.h>
#include <string.h>
#endif

namespace umcx {
namespace libm {

UMCX_HD uint32_t f2u(float f) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(uint32_t, f);
#else
    uint32_t u; memcpy(&u, &f, 4); return u;
#endif
}
UMCX_HD float u2f(uint32_t u) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(float, u);
#else
    float f; memcpy(&f, &u, 4); return f;
#endif
}
UMCX_HD double dfma(double a, double b, double c) {
#if defined(__HIPCC__)
    return __builtin_fma(a, b, c);
#else
    return fma(a, b, c);
#endif
}

// ---- logf: log(x) = log1p(z/c - 1) + log(c) + k ln2, with x = 2^k z, z near c ----
struct LogfEntry { double invc, logc; };
#if defined(__HIPCC__)
__device__ __constant__
#endif
static const LogfEntry kLogfTab[16] = {
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2}, {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2},
    {0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2},  {0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3},
    {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3}, {0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3},
    {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4}, {0x1.12358f08ae5bap+0, -0x1.1973c5a611ccp-4},
    {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5}, {0x1p+0, 0x0p+0},
    {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5},  {0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4},
    {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3}, {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3},
    {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2}, {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2},
};

UMCX_HD float logf_tab(float x, const LogfEntry* tab) {
    const double ln2 = 0x1.62e42fefa39efp-1;
    const double a0 = -0x1.00ea348b88334p-2, a1 = 0x1.5575b0be00b6ap-2, a2 = -0x1.ffffef20a4123p-2;
    uint32_t ix = f2u(x);
    if (ix == 0x3f800000u) return 0.f;
    uint32_t t = ix - 0x3f330000u;            // subintervals start at 0x3f330000 (~0.7)
    int idx = (int)((t >> 19) & 15u);
    int k = (int32_t)t >> 23;
    float zf = u2f(ix - (t & 0xff800000u));
    const LogfEntry e = tab[idx];
    double r = dfma((double)zf, e.invc, -1.0);
    double y0 = dfma((double)k, ln2, e.logc);
    double r2 = r * r;
    double y = dfma(a1, r, a2);
    y = dfma(a0, r2, y);
    y = dfma(y, r2, y0 + r);
    return (float)y;
}

UMCX_HD float logf_host(float x) { return logf_tab(x, kLogfTab); }

// ---- sinf/cosf: x = n pi/2 + r, |r| <= pi/4, then odd/even polynomials in double ----
UMCX_HD uint32_t abstop12(float x) { return (f2u(x) >> 20) & 0x7ffu; }

// odd polynomial S(r) ~ sin r; S(-r) == -S(r) exactly (every fused step is sign-symmetric)
UMCX_HD float sin_poly(double x, double x2) {
    const double s1 = -0x1.555545995a603p-3, s2 = 0x1.1107605230bc4p-7, s3 = -0x1.994eb3774cf24p-13;
    double x3 = x * x2;
    double p = dfma(x2, s3, s2);
    double x7 = x3 * x2;
    double s = dfma(x3, s1, x);
    return (float)dfma(x7, p, s);
}
// even polynomial C(r) ~ cos r; the host library's negated-coefficient table gives exactly -C(r)
UMCX_HD float cos_poly(double x2) {
    const double c0 = 0x1p0, c1 = -0x1.ffffffd0c621cp-2, c2 = 0x1.55553e1068f19p-5,
                 c3 = -0x1.6c087e89a359dp-10, c4 = 0x1.99343027bf8c3p-16;
    double x4 = x2 * x2;
    double q2 = dfma(x2, c4, c3);
    double q1 = dfma(x2, c1, c0);
    double x6 = x4 * x2;
    double c = dfma(x4, c2, q1);
    return (float)dfma(x6, q2, c);
}

// Quadrant reduction shared by sinf and cosf of one argument; |y| < 120.
UMCX_HD double reduce_quadrant(double x, int* n) {
    const double hpi_inv_scaled = 0x1.45F306DC9C883p+23;   // 2/pi * 2^24
    const double hpi = 0x1.921FB54442D18p0;
    double r = x * hpi_inv_scaled;
    int q = ((int32_t)r + 0x800000) >> 24;
    *n = q;
    return dfma(-(double)q, hpi, x);
}

// sinf(y) and cosf(y) of one argument, each bitwise equal to the separate host calls. Both
// polynomials are evaluated once and the quadrant picks them with signs:
//   n%4 = 0: (S, C)   1: (C, -S)   2: (-S, -C)   3: (-C, S)
UMCX_HD void sincosf_host(float y, float* s, float* c) {
    const uint32_t top = abstop12(y);
    double r = (double)y;
    int n = 0;
    if (top >= abstop12(0x1.921FB6p-1f)) r = reduce_quadrant(r, &n);   // |y| >= pi/4
    const double r2 = r * r;
    const float S = sin_poly(r, r2), C = cos_poly(r2);
    const int q = n & 3;
    float sv = (q & 1) ? C : S;
    float cv = (q & 1) ? S : C;
    sv = (q & 2) ? -sv : sv;
    cv = (q == 1 || q == 2) ? -cv : cv;
    const bool tiny = top < abstop12(0x1p-12f);
    *s = tiny ? y : sv;
    *c = tiny ? 1.0f : cv;
}
UMCX_HD float sinf_host(float y) { float s, c; sincosf_host(y, &s, &c); return s; }
UMCX_HD float cosf_host(float y) { float s, c; sincosf_host(y, &s, &c); return c; }

// The binary fraction of 2/pi in 32-bit windows: entry i >= 3 holds fraction bits
// [8(i-3), 8(i-3)+32); entries 0-2 hold its first 8, 16 and 24 bits. Computed from pi (Machin's
// formula in integer arithmetic), not copied from any binary.
#if defined(__HIPCC__)
__device__ __constant__
#endif
static const uint32_t kInvPio4[24] = {
    0x000000a2u, 0x0000a2f9u, 0x00a2f983u, 0xa2f9836eu, 0xf9836e4eu, 0x836e4e44u, 0x6e4e4415u, 0x4e441529u,
    0x441529fcu, 0x1529fc27u, 0x29fc2757u, 0xfc2757d1u, 0x2757d1f5u, 0x57d1f534u, 0xd1f534ddu, 0xf534ddc0u,
    0x34ddc0dbu, 0xddc0db62u, 0xc0db6295u, 0xdb629599u, 0x6295993cu, 0x95993c43u, 0x993c4390u, 0x3c439041u,
};

// Reduction of |y| >= 120 by pi/2, in integers: the 24-bit mantissa (shifted by the low three
// exponent bits) times the 96 bits of 2/pi that matter at this exponent gives the fraction of
// |y| 2/pi in 2.62 fixed point. Returns r in [-pi/4, pi/4] (accurate to about 33 bits) and the
// quadrant in *np. Only the low 32 bits of the first product are kept.
UMCX_HD double reduce_large(uint32_t xi, int* np, const uint32_t* tab) {
    const double pi63 = 0x1.921FB54442D18p-62;
    const uint32_t* arr = tab + ((xi >> 26) & 15u);
    const int shift = (int)((xi >> 23) & 7u);
    const uint32_t m = ((xi & 0xffffffu) | 0x800000u) << shift;
    uint64_t res0 = (uint64_t)(uint32_t)(m * arr[0]);
    const uint64_t res1 = (uint64_t)m * arr[4];
    const uint64_t res2 = (uint64_t)m * arr[8];
    res0 = (res2 >> 32) | (res0 << 32);
    res0 += res1;
    const uint64_t n = (res0 + (1ull << 61)) >> 62;
    res0 -= n << 62;
    *np = (int)n;
    return (double)(int64_t)res0 * pi63;
}

// sinf and cosf of any argument, each bitwise equal to the host library: |y| < 120 as
// sincosf_host; larger finite arguments through reduce_large, where the argument's sign enters
// the quadrant as (n + sign); infinities and NaN give NaN. For source sampling, whose angles the
// user sets (cone half-angle, fourier phase).
UMCX_HD void sincosf_any(float y, float* s, float* c, const uint32_t* tab) {
    const uint32_t top = abstop12(y);
    if (top < abstop12(120.0f)) {
        sincosf_host(y, s, c);
        return;
    }
    if (top >= 0x7f8u) {
        *s = *c = (y - y) / (y - y);
        return;
    }
    const uint32_t xi = f2u(y);
    int n = 0;
    const double r = reduce_large(xi, &n, tab);
    const int q = (n + (int)(xi >> 31)) & 3;
    const double sg = (q == 1 || q == 2) ? -1.0 : 1.0;
    const double r2 = r * r;
    const float S = sin_poly(r * sg, r2), C = cos_poly(r2);
    const float Cq = (q & 2) ? -C : C;
    *s = (n & 1) ? Cq : S;
    *c = (n & 1) ? S : Cq;
}
UMCX_HD float sinf_any(float y) { float s, c; sincosf_any(y, &s, &c, kInvPio4); return s; }
UMCX_HD float cosf_any(float y) { float s, c; sincosf_any(y, &s, &c, kInvPio4); return c; }

UMCX_HD float sqrtf_cr(float x) {
#if defined(__HIPCC__)
    return __builtin_sqrtf(x);      // correctly rounded under -fhip-fp32-correctly-rounded-divide-sqrt
#else
    return sqrtf(x);
#endif
}

// Correctly rounded sqrt, reciprocal and division for operands away from the special ranges,
// shorter than the compiler's correctly rounded sequences, which scale their operands and fix up
// special cases:
//  * sqrt_ns(x): x finite and >= 2^-96 (the scaling threshold; zero, inf and NaN excluded): the
//    hardware estimate, then the neighbour below or above when its residual says so;
//  * rcp_ns(b): b normal with |b| < 2^126 (1/b normal): the hardware estimate refined by one
//    Newton step (three instructions); the estimate alone is not correctly rounded, the refined
//    value is on every float of that range (checked on the GPU);
//  * div_ns(a, b): 2^-120 <= |b| <= 2^120, and a == +0 or (2^-100 <= |a| and 2^-120 <= |a/b| <=
//    2^120): q = a * rcp_ns(b), then one residual correction q + (a - b q) * rcp_ns(b), the
//    residual exact by FMA. With the reciprocal correctly rounded this gives the correctly rounded
//    quotient (Markstein's theorem; P. Markstein, IA-64 and Elementary Functions, 2000).
// Bitwise equal to sqrtf_cr, 1 / b and a / b there (exhaustive and sampled GPU checks in the tests).
UMCX_HD float sqrt_ns(float x) {
#if defined(__HIPCC__)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float below = u2f(f2u(s) - 1u), above = u2f(f2u(s) + 1u);
    const float t = __builtin_fmaf(-below, s, x) <= 0.f ? below : s;
    return __builtin_fmaf(-above, s, x) > 0.f ? above : t;
#else
    return sqrtf(x);
#endif
}
UMCX_HD float rcp_ns(float b) {
#if defined(__HIPCC__)
    const float r = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
#else
    return 1.0f / b;
#endif
}
UMCX_HD float div_ns(float a, float b) {
#if defined(__HIPCC__)
    const float y = rcp_ns(b);
    const float q = a * y;
    return __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
#else
    return a / b;
#endif
}
// 1 / x for a direction component (|x| <= 1 up to rounding): the short sequence for
// 2^-60 <= |x| <= 2^60, the full correctly rounded division otherwise (zero, tiny or huge x).
UMCX_HD float recip_dir(float x) {
    return (fabsf(x) >= 0x1p-60f && fabsf(x) <= 0x1p60f) ? rcp_ns(x) : 1.0f / x;
}
// ---- acosf: fdlibm single precision ----
UMCX_HD float acos_rational(float z) {
    const float pS0 = 1.6666667163e-01f, pS1 = -3.2556581497e-01f, pS2 = 2.0121252537e-01f,
                pS3 = -4.0055535734e-02f, pS4 = 7.9153501429e-04f, pS5 = 3.4793309169e-05f,
                qS1 = -2.4033949375e+00f, qS2 = 2.0209457874e+00f, qS3 = -6.8828397989e-01f,
                qS4 = 7.7038154006e-02f;
    float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    float q = 1.0f + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    // q in [0.22, 1] and p near z / 6: the short division for z >= 2^-80 (acosf's small-argument
    // range reaches z = 2^-114)
    return z >= 0x1p-80f ? div_ns(p, q) : p / q;
}
UMCX_HD float acosf_host(float x) {
    const float pi = 3.1415925026e+00f, pio2_hi = 1.5707962513e+00f, pio2_lo = 7.5497894159e-08f;
    const int32_t hx = (int32_t)f2u(x);
    const int32_t ix = hx & 0x7fffffff;
    if (ix >= 0x3f800000) {                                  // |x| >= 1
        if (ix == 0x3f800000) return hx > 0 ? 0.0f : pi + 2.0f * pio2_lo;
        return (x - x) / (x - x);
    }
    // the three ranges share one rational approximation R(z) = p(z)/q(z)
    const bool small = ix < 0x3f000000;                      // |x| < 0.5: z = x^2
    const bool neg = hx < 0;                                 // x <= -0.5: z = (1+x)/2, else (1-x)/2
    const float z = small ? x * x : (neg ? (1.0f + x) * 0.5f : (1.0f - x) * 0.5f);
    const float r = acos_rational(z);
    if (small) {
        if (ix <= 0x23000000) return pio2_hi + pio2_lo;
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    const float s = sqrt_ns(z);                 // z in [2^-25, 0.25]
    if (neg) {
        const float w = r * s - pio2_lo;
        return pi - 2.0f * (s + w);
    }
    const float df = u2f(f2u(s) & 0xfffff000u);
    const float c = div_ns(z - df * df, s + df);   // s + df in [2^-13, 1]; numerator +0 or >= 2^-50
    const float w = r * s + c;
    return 2.0f * (df + w);
}

// ---- expf: 2^(k/32) * 2^(r/32) with a cubic for 2^(r/32), in double ----
// kExp2Tab[i] = bits(2^(i/32) rounded to double) - (i << 47): the integer k/32 is added to the
// exponent field afterwards. Values computed from 2^(i/32) at 80 digits (not copied from any
// binary). The FMA build of the host library fuses InvLn2N*x into both of its uses.
#if defined(__HIPCC__)
__device__ __constant__
#endif
static const uint64_t kExp2Tab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
};

UMCX_HD double u642d(uint64_t u) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(double, u);
#else
    double d; memcpy(&d, &u, 8); return d;
#endif
}
UMCX_HD uint64_t d2u64(double d) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(uint64_t, d);
#else
    uint64_t u; memcpy(&u, &d, 8); return u;
#endif
}

// exp(x) for x in [-103.97, 88.72] (the photon loop only passes -mua*d <= 0); outside that
// range it saturates to 0 / +inf like the host library.
UMCX_HD float expf_tab(float x, const uint64_t* tab) {
    const double inv_ln2_n = 0x1.71547652b82fep+0 * 32, shift = 0x1.8p+52;
    const double c0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, c1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
                 c2 = 0x1.62e42ff0c52d6p-1 / 32;
    if (x != x) return x + x;
    if (x < -0x1.9fe368p6f) return 0.0f;
    if (x > 0x1.62e42ep6f) return u2f(0x7f800000u);
    double xd = (double)x;
    double kd = dfma(inv_ln2_n, xd, shift);
    uint64_t ki = d2u64(kd);
    kd -= shift;
    double r = dfma(inv_ln2_n, xd, -kd);
    uint64_t t = tab[ki % 32] + (ki << 47);
    double s = u642d(t);
    double z = dfma(c0, r, c1);
    double r2 = r * r;
    double y = dfma(c2, r, 1.0);
    y = dfma(z, r2, y);
    return (float)(y * s);
}

UMCX_HD float expf_host(float x) { return expf_tab(x, kExp2Tab); }

// ---- nextafterf(x, y) by bit stepping (IEEE 754 semantics, denormals kept) ----
UMCX_HD float nextafterf_host(float x, float y) {
    if (x != x || y != y) return x + y;
    if (x == y) return y;
    uint32_t ux = f2u(x);
    if ((ux & 0x7fffffffu) == 0) return u2f((f2u(y) & 0x80000000u) | 1u);
    bool up = (x < y) == (x > 0.f);      // move away from zero when heading outward
    return u2f(up ? ux + 1u : ux - 1u);
}

}  // namespace libm
}  // namespace umcx
