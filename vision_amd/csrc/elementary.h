// Single-precision elementary functions written for this library: sin/cos,
// exp, log, log1p, atanh, cosh, pow, asin/acos and atan2.
//
// The path tracer's sampling decisions (reflect or refract, which texel,
// which wavelength) depend on the last bits of these functions, so the HIP
// kernels and every host-side restatement of them must round identically.
// Vendor math libraries do not (the device and host libraries differ by an
// ulp or two), so both sides call these instead.  Every operation below is an
// IEEE single-precision add, multiply, divide, sqrt, fused multiply-add or an
// exact bit manipulation, each correctly rounded on gfx950 and on x86-64
// (HIP divides and square roots are correctly rounded by default), so the
// results are bit-identical on both.
//
// Methods:
//  * argument reduction by Cody & Waite (Software Manual for the Elementary
//    Functions, 1980): x - k*c with c split into pieces whose products with k
//    are exact (here through fused multiply-adds);
//  * kernels are polynomials in the reduced argument, with coefficients
//    fitted by vision_amd/gen/fit_elementary.py (Chebyshev fits in mpmath);
//  * log follows the atanh series form log(1+f) = 2s + s R(s^2),
//    s = f / (2 + f), with the f - (hfsq - s (hfsq + R)) evaluation order;
//  * log1p uses Goldberg's correction log(u) * x / (u - 1), u = 1 + x
//    ("What every computer scientist should know about floating-point
//    arithmetic", 1991, theorem 4);
//  * acos(x) = 2 asin(sqrt((1 - |x|) / 2)) for |x| > 1/2;
//  * atan2 reduces a = min/max > tan(pi/8) by atan a = pi/4 + atan((a - 1) /
//    (a + 1)), the quotient formed from the operands as (min - max) /
//    (min + max).
//
// Accuracy (tests/test_elementary.py, against float64 references): at most 2
// ulp for sin/cos, exp and log, 2.5 for atan2, 3 for acos, atanh, cosh and
// log1p, and 8 for pow on the domain the film's sRGB encoding uses.  sin/cos take |x| <= 1e5
// (larger arguments return NaN); every call site passes angles in
// [-2 pi, 2 pi].
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "vmath.h"

namespace vs {
namespace em {

// ---- coefficients (printed by vision_amd/gen/fit_elementary.py)
// sin remainder S(z): degree 3 on [0, 0.61685], fit error 2.81e-11
// cos remainder C(z): degree 3 on [0, 0.61685], fit error 2.34e-12
// exp remainder E(r): degree 5 on [-0.346574, 0.346574], fit error 1.4e-9
// log remainder L(z): degree 3 on [0, 0.0294373], fit error 1.15e-9
// asin remainder A(z): degree 5 on [0, 0.25], fit error 4.17e-9
// atan remainder T(z): degree 5 on [0, 0.171573], fit error 5.35e-10
constexpr float kPio2_1 = 0x1.921fb6p+0f, kPio2_2 = -0x1.777a5cp-25f, kPio2_3 = -0x1.ee59dap-50f;
constexpr float kLn2Hi = 0x1.62ep-1f, kLn2Lo = 0x1.0bfbe8p-15f;
constexpr float kPio2Hi = 0x1.921fb6p+0f, kPio2Lo = -0x1.777a5cp-25f;
constexpr float kPio4Hi = 0x1.921fb6p-1f, kPio4Lo = -0x1.777a5cp-26f;
constexpr float kPiHi = 0x1.921fb6p+1f, kPiLo = -0x1.777a5cp-24f;
constexpr float kTanPi8 = 0x1.a8279ap-2f;
constexpr float kS0 = -0x1.555556p-3f, kS1 = 0x1.11110ep-7f, kS2 = -0x1.a013a8p-13f, kS3 = 0x1.6dbe08p-19f;
constexpr float kC0 = 0x1.555556p-5f, kC1 = -0x1.6c16cp-10f, kC2 = 0x1.a015c4p-16f, kC3 = -0x1.25244ep-22f;
constexpr float kE0 = 0x1p-1f, kE1 = 0x1.555556p-3f, kE2 = 0x1.5554eap-5f, kE3 = 0x1.1110ep-7f,
                kE4 = 0x1.6d4316p-10f, kE5 = 0x1.a124e4p-13f;
constexpr float kL0 = 0x1.555556p-1f, kL1 = 0x1.9999ecp-2f, kL2 = 0x1.245c44p-2f, kL3 = 0x1.ddced8p-3f;
constexpr float kA0 = 0x1.555554p-3f, kA1 = 0x1.33343p-4f, kA2 = 0x1.6d5bbap-5f, kA3 = 0x1.fd8da2p-6f,
                kA4 = 0x1.18f91ep-6f, kA5 = 0x1.13fed4p-5f;
constexpr float kT0 = -0x1.555556p-2f, kT1 = 0x1.99997cp-3f, kT2 = -0x1.248a18p-3f, kT3 = 0x1.c57baep-4f,
                kT4 = -0x1.614448p-4f, kT5 = 0x1.9d8b24p-5f;

VS_HD uint32_t bits_of(float f) { return __builtin_bit_cast(uint32_t, f); }
VS_HD float float_of(uint32_t u) { return __builtin_bit_cast(float, u); }
VS_HD float quiet_nan() { return float_of(0x7fc00000u); }
VS_HD float infinity() { return float_of(0x7f800000u); }
// 2^e for e in [-126, 127]
VS_HD float pow2i(int e) { return float_of(uint32_t(e + 127) << 23); }

// sin and cos of x, |x| <= 1e5
VS_HD void sincos(float x, float *s, float *c) {
    if (!(fabsf(x) <= 1e5f)) {
        *s = *c = quiet_nan();
        return;
    }
    const float k = rintf(x * 0x1.45f306p-1f);  // x * 2/pi, rounded half to even
    float r = fmaf(k, -kPio2_1, x);
    r = fmaf(k, -kPio2_2, r);
    r = fmaf(k, -kPio2_3, r);
    const float z = r * r;
    const float ps = fmaf(r * z, fmaf(fmaf(fmaf(kS3, z, kS2), z, kS1), z, kS0), r);
    const float pc = fmaf(z * z, fmaf(fmaf(fmaf(kC3, z, kC2), z, kC1), z, kC0), fmaf(-0.5f, z, 1.f));
    const int q = int(k - 4.f * floorf(k * 0.25f));  // k mod 4, exact for |k| < 2^22
    const float sv = (q & 1) ? pc : ps;
    const float cv = (q & 1) ? ps : pc;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}
VS_HD float sin(float x) {
    float s, c;
    sincos(x, &s, &c);
    return s;
}
VS_HD float cos(float x) {
    float s, c;
    sincos(x, &s, &c);
    return c;
}

VS_HD float exp(float x) {
    if (x != x) return x;
    if (x > 0x1.62e42ep+6f) return infinity();  // above ln(FLT_MAX)
    if (x < -0x1.9fe368p+6f) return 0.f;        // below ln of half the least subnormal
    const float k = rintf(x * 0x1.715476p+0f);  // x / ln 2
    float r = fmaf(k, -kLn2Hi, x);
    r = fmaf(k, -kLn2Lo, r);
    const float e = fmaf(fmaf(fmaf(fmaf(kE5, r, kE4), r, kE3), r, kE2), r, kE1);
    const float p = 1.f + fmaf(r * r, fmaf(e, r, kE0), r);
    const int ki = int(k);  // in [-150, 128]
    const int h = ki / 2;
    return p * pow2i(h) * pow2i(ki - h);
}

VS_HD float log(float x) {
    if (!(x > 0.f)) return x == 0.f ? -infinity() : quiet_nan();
    if (x == infinity()) return x;
    int e = 0;
    uint32_t b = bits_of(x);
    if (b < 0x00800000u) {  // subnormal
        b = bits_of(x * 0x1p23f);
        e = -23;
    }
    e += int(b >> 23) - 127;
    float m = float_of((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 0x1.6a09e6p+0f) {                            // above sqrt 2
        m *= 0.5f;
        e += 1;
    }
    const float f = m - 1.f;  // exact: m in (1/sqrt 2, sqrt 2]
    const float s = f / (2.f + f);
    const float z = s * s;
    const float R = z * fmaf(fmaf(fmaf(kL3, z, kL2), z, kL1), z, kL0);
    const float hfsq = 0.5f * f * f;
    const float lm = f - (hfsq - s * (hfsq + R));
    const float fe = float(e);
    return fmaf(fe, kLn2Hi, fmaf(fe, kLn2Lo, lm));
}

VS_HD float log1p(float x) {
    const float u = 1.f + x;
    if (u == 1.f) return x;
    if (u == infinity()) return u;
    return log(u) * (x / (u - 1.f));
}

// odd: atanh(-x) = -atanh(x), so the log1p argument 2|x| / (1 - |x|) is never
// near -1, where 1 + y would cancel
VS_HD float atanh(float x) {
    const float ax = fabsf(x);
    return copysignf(0.5f * log1p((ax + ax) / (1.f - ax)), x);
}

VS_HD float cosh(float x) {
    const float e = exp(fabsf(x));
    return 0.5f * (e + 1.f / e);
}

// x^y for x >= 0 (the film's sRGB encoding); negative x gives NaN
VS_HD float pow(float x, float y) {
    if (x == 0.f) return y > 0.f ? 0.f : (y == 0.f ? 1.f : infinity());
    return exp(y * log(x));
}

// asin on |x| <= 1/2: x + x^3 A(x^2)
VS_HD float asin_kernel(float x) {
    const float z = x * x;
    const float a = fmaf(fmaf(fmaf(fmaf(fmaf(kA5, z, kA4), z, kA3), z, kA2), z, kA1), z, kA0);
    return fmaf(x * z, a, x);
}

VS_HD float acos(float x) {
    const float ax = fabsf(x);
    if (ax <= 0.5f) return kPio2Hi - (asin_kernel(x) - kPio2Lo);
    const float t = (1.f - ax) * 0.5f;  // NaN for |x| > 1 through sqrt
    const float r = 2.f * asin_kernel(sqrtf(t));
    return x > 0.f ? r : kPiHi - (r - kPiLo);
}

// atan on |a| <= tan(pi/8): a + a^3 T(a^2)
VS_HD float atan_kernel(float a) {
    const float z = a * a;
    const float t = fmaf(fmaf(fmaf(fmaf(fmaf(kT5, z, kT4), z, kT3), z, kT2), z, kT1), z, kT0);
    return fmaf(a * z, t, a);
}

VS_HD float atan2(float y, float x) {
    if (x != x || y != y) return x + y;
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == infinity() || ay == infinity()) {
        if (ax == infinity() && ay == infinity()) r = x > 0.f ? kPio4Hi : 3.f * kPio4Hi;
        else if (ay == infinity()) r = kPio2Hi;
        else r = x > 0.f ? 0.f : kPiHi;
        return copysignf(r, y);
    }
    if (ay == 0.f) {
        r = (bits_of(x) >> 31) ? kPiHi : 0.f;  // atan2(+-0, -0 or x < 0) = +-pi
        return copysignf(r, y);
    }
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mn / mx;
    if (a > kTanPi8) r = kPio4Hi + (atan_kernel((mn - mx) / (mn + mx)) + kPio4Lo);
    else r = atan_kernel(a);
    if (ay > ax) r = kPio2Hi - (r - kPio2Lo);
    if (x < 0.f) r = kPiHi - (r - kPiLo);
    return copysignf(r, y);
}

}// namespace em
}// namespace vs
