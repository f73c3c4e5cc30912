// CPU oracle: deterministic float transcendentals.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// The reference evaluates exp/cos/sin in these paths through libm or
// OpenImageIO's fast_* approximations (OIIO source is not part of the
// reference tree). The oracle fixes one explicit algorithm so that its
// results are reproducible on any IEEE machine: Cody-Waite range reduction
// followed by a Horner polynomial evaluated with fmaf. Accuracy against
// libm is checked by tests/test_oracle_math.py (<= 2 ulp on the tested
// ranges).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oref {

#ifdef OREF_F64
// FP64 evaluation build (f64/f64_main.cpp): `float` is binary64 there and
// each approximation below is replaced by the exact function it stands
// for, so that build evaluates the shading path's mathematics, not its
// binary32 rounding or its polynomial truncation.
inline float ofast_expf(float x) { return std::exp(x); }
inline void ofast_sincosf(float x, float* s, float* c)
{
    *s = std::sin(x);
    *c = std::cos(x);
}
inline float ofast_expm1f(float x) { return std::expm1(x); }
inline float ofast_log2f(float x) { return std::log2(x); }
inline float ofast_exp2f(float t) { return std::exp2(t); }
inline float ofast_safe_powf(float x, float y)
{
    if (y == 0.0f)
        return 1.0f;
    if (x == 0.0f)
        return 0.0f;
    if (y == 1.0f)
        return x;
    if (y == 2.0f) {
        const float x2   = x * x;
        const float fmax = 3.40282346638528859812e+38f;
        return x2 < fmax ? x2 : fmax;
    }
    if (x < 0.0f)
        return 0.0f;
    return std::pow(x, y);
}
inline float ofast_cbrtf(float x) { return std::cbrt(x); }
inline float ofast_logf(float x) { return std::log(x); }
inline float ofast_erff(float x) { return std::erf(x); }
// the inverse of erf by Newton steps on std::erf from a bisection start
inline float ofast_ierff(float x)
{
    if (!(x > -1.0f && x < 1.0f))
        return x * HUGE_VAL;
    float lo = -6.0f, hi = 6.0f;
    for (int i = 0; i < 60; ++i) {
        const float mid = 0.5f * (lo + hi);
        if (std::erf(mid) < x)
            lo = mid;
        else
            hi = mid;
    }
    float y = 0.5f * (lo + hi);
    for (int i = 0; i < 3; ++i)
        y -= (std::erf(y) - x) / (1.1283791670955126 * std::exp(-y * y));
    return y;
}
inline float osafe_powf(float x, float y)
{
    if (y == 0.0f)
        return 1.0f;
    if (x == 0.0f || (x < 0.0f && y != std::floor(y)))
        return 0.0f;
    return std::pow(x, y);
}
#else

inline float bits_f(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// e^x. Range reduction x = k*ln2 + r, |r| <= ln2/2; e^r by the Taylor
// polynomial to degree 7 (truncation error < 6e-9 relative); 2^k applied
// in two exact steps so results down to the subnormal range are formed.
inline float ofast_expf(float x)
{
    if (!(x == x))
        return x;
    if (x > 88.72283935546875f)
        return bits_f(0x7f800000u);
    if (x < -103.97208404541016f)
        return 0.0f;
    const float log2e  = 1.44269502162933349609375f;
    const float ln2_hi = 0.693145751953125f;          // 0x3f317200
    const float ln2_lo = 1.428606765330187045e-06f;   // ln2 - ln2_hi
    float k            = rintf(x * log2e);
    float r            = fmaf(-k, ln2_hi, x);
    r                  = fmaf(-k, ln2_lo, r);
    float p            = 1.0f / 5040.0f;
    p                  = fmaf(p, r, 1.0f / 720.0f);
    p                  = fmaf(p, r, 1.0f / 120.0f);
    p                  = fmaf(p, r, 1.0f / 24.0f);
    p                  = fmaf(p, r, 1.0f / 6.0f);
    p                  = fmaf(p, r, 0.5f);
    p                  = fmaf(p, r, 1.0f);
    p                  = fmaf(p, r, 1.0f);
    int ki             = (int)k;
    int k1             = ki / 2;
    int k2             = ki - k1;
    float s1 = bits_f((uint32_t)(k1 + 127) << 23);
    float s2 = bits_f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// sin and cos of x. Reduction by multiples of pi/2 with a three-part
// constant (exact for |x| < ~1e5), then Taylor polynomials on
// [-pi/4, pi/4]: sin to r^9 and cos to r^10.
inline void ofast_sincosf(float x, float* s, float* c)
{
    const float two_over_pi = 0.636619746685028076171875f;
    const float pio2_1      = 1.5703125f;                  // 0x3fc90000
    const float pio2_2      = 4.837512969970703125e-04f;   // 0x39fda000
    const float pio2_3      = 7.549790126404332e-08f;      // pi/2 - (1 + 2)
    float q                 = rintf(x * two_over_pi);
    float r                 = fmaf(-q, pio2_1, x);
    r                       = fmaf(-q, pio2_2, r);
    r                       = fmaf(-q, pio2_3, r);
    float r2                = r * r;
    float ps                = 1.0f / 362880.0f;
    ps                      = fmaf(ps, r2, -1.0f / 5040.0f);
    ps                      = fmaf(ps, r2, 1.0f / 120.0f);
    ps                      = fmaf(ps, r2, -1.0f / 6.0f);
    float sn                = fmaf(ps * r2, r, r);
    float pc                = -1.0f / 3628800.0f;
    pc                      = fmaf(pc, r2, 1.0f / 40320.0f);
    pc                      = fmaf(pc, r2, -1.0f / 720.0f);
    pc                      = fmaf(pc, r2, 1.0f / 24.0f);
    pc                      = fmaf(pc, r2, -0.5f);
    float cs                = fmaf(pc, r2, 1.0f);
    int qi                  = ((int)q) & 3;
    float sv, cv;
    switch (qi) {
    case 0: sv = sn; cv = cs; break;
    case 1: sv = cs; cv = -sn; break;
    case 2: sv = -sn; cv = -cs; break;
    default: sv = -cs; cv = sn; break;
    }
    *s = sv;
    *c = cv;
}

// e^x - 1: Taylor series to x^9 for |x| < 0.5, exp(x) - 1 otherwise.
inline float ofast_expm1f(float x)
{
    if (fabsf(x) < 0.5f) {
        float p = 1.0f / 362880.0f;
        p       = fmaf(p, x, 1.0f / 40320.0f);
        p       = fmaf(p, x, 1.0f / 5040.0f);
        p       = fmaf(p, x, 1.0f / 720.0f);
        p       = fmaf(p, x, 1.0f / 120.0f);
        p       = fmaf(p, x, 1.0f / 24.0f);
        p       = fmaf(p, x, 1.0f / 6.0f);
        p       = fmaf(p, x, 0.5f);
        p       = fmaf(p, x, 1.0f);
        return p * x;
    }
    return ofast_expf(x) - 1.0f;
}

// log2(x) for finite x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, series to s^9
// (truncation < 3e-9 relative), then e + ln(m) log2(e).
inline float ofast_log2f(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    int e = 0;
    if (u < 0x00800000u) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;
        std::memcpy(&u, &x, 4);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = bits_f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s  = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p  = 1.0f / 9.0f;
    p        = fmaf(p, s2, 1.0f / 7.0f);
    p        = fmaf(p, s2, 1.0f / 5.0f);
    p        = fmaf(p, s2, 1.0f / 3.0f);
    p        = fmaf(p, s2, 1.0f);
    float ln = 2.0f * s * p;
    return fmaf(ln, 1.44269502162933349609375f, (float)e);
}

// 2^t: t = k + f with k = rint(t) (f exact, |f| <= 1/2), 2^f by the Taylor
// series of e^(f ln2) to degree 7 (truncation < 6e-9), 2^k applied in two
// exact scalings.
inline float ofast_exp2f(float t)
{
    if (!(t == t))
        return t;
    if (t >= 128.0f)
        return bits_f(0x7f800000u);
    if (t < -150.0f)
        return 0.0f;
    float k = rintf(t);
    float f = t - k;
    float p = 1.5252733804059840e-05f;  // ln2^7 / 7!
    p       = fmaf(p, f, 1.5403530393381608e-04f);  // ln2^6 / 6!
    p       = fmaf(p, f, 1.3333558146428443e-03f);  // ln2^5 / 5!
    p       = fmaf(p, f, 9.6181291076284772e-03f);  // ln2^4 / 4!
    p       = fmaf(p, f, 5.5504108664821580e-02f);  // ln2^3 / 3!
    p       = fmaf(p, f, 2.4022650695910071e-01f);  // ln2^2 / 2!
    p       = fmaf(p, f, 6.9314718055994531e-01f);  // ln2
    p       = fmaf(p, f, 1.0f);
    int ki  = (int)k;
    int k1  = ki / 2;
    int k2  = ki - k1;
    float s1 = bits_f((uint32_t)(k1 + 127) << 23);
    float s2 = bits_f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// x^y with the special cases of OpenImageIO's fast_safe_pow (y == 0 -> 1,
// x == 0 -> 0, y == 1 -> x, y == 2 -> min(x*x, FLT_MAX)); x < 0 returns 0
// (the closure code calls it with x >= 0 only); otherwise
// 2^(log2(x) * y).
inline float ofast_safe_powf(float x, float y)
{
    if (y == 0.0f)
        return 1.0f;
    if (x == 0.0f)
        return 0.0f;
    if (y == 1.0f)
        return x;
    if (y == 2.0f) {
        const float x2   = x * x;
        const float fmax = 3.40282346638528859812e+38f;
        return x2 < fmax ? x2 : fmax;
    }
    if (x < 0.0f)
        return 0.0f;
    return ofast_exp2f(ofast_log2f(x) * y);
}

// x^y with the special cases of OpenImageIO's safe_pow (y == 0 -> 1, x == 0
// or a negative x with a non-integer y -> 0, the result clamped to the
// finite range), as 2^(log2|x| * y), negated for a negative x and an odd
// integer y.
inline float osafe_powf(float x, float y)
{
    if (y == 0.0f)
        return 1.0f;
    if (x == 0.0f || (x < 0.0f && y != std::floor(y)))
        return 0.0f;
    float r = ofast_exp2f(ofast_log2f(std::fabs(x)) * y);
    if (x < 0.0f && std::fmod(y, 2.0f) != 0.0f)
        r = -r;
    const float big = 3.40282346638528859812e+38f;
    return r < -big ? -big : (r > big ? big : r);
}

// Cube root: bit-level initial guess, then three Newton steps
// y <- y - (y^3 - x) / (3 y^2) for x > 0 (sign restored for x < 0).
inline float ofast_cbrtf(float x)
{
    if (x == 0.0f || !(x == x))
        return x;
    float ax = fabsf(x);
    uint32_t u;
    std::memcpy(&u, &ax, 4);
    uint32_t g = u / 3u + 709921077u;
    float y    = bits_f(g);
    for (int i = 0; i < 3; ++i) {
        float y2 = y * y;
        y        = y - (y2 * y - ax) / (3.0f * y2);
    }
    return x < 0.0f ? -y : y;
}

// ln x for finite x > 0: the reduction of ofast_log2f, e ln2 added last.
inline float ofast_logf(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    int e = 0;
    if (u < 0x00800000u) {
        x = x * 8388608.0f;
        std::memcpy(&u, &x, 4);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = bits_f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s  = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p  = 1.0f / 9.0f;
    p        = fmaf(p, s2, 1.0f / 7.0f);
    p        = fmaf(p, s2, 1.0f / 5.0f);
    p        = fmaf(p, s2, 1.0f / 3.0f);
    p        = fmaf(p, s2, 1.0f);
    float ln = 2.0f * s * p;
    return fmaf((float)e, 0.693147182464599609375f, ln);
}

// erf by Abramowitz and Stegun 7.1.28: 1 - (1 + a1|x| + ... + a6|x|^6)^-16
// with the sign of x (absolute error <= 3e-7 in exact arithmetic).
inline float ofast_erff(float x)
{
    const float ax = std::fabs(x);
    float p        = 0.0000430638f;
    p              = fmaf(p, ax, 0.0002765672f);
    p              = fmaf(p, ax, 0.0001520143f);
    p              = fmaf(p, ax, 0.0092705272f);
    p              = fmaf(p, ax, 0.0422820123f);
    p              = fmaf(p, ax, 0.0705230784f);
    p              = fmaf(p, ax, 1.0f);
    const float p2 = p * p, p4 = p2 * p2, p8 = p4 * p4, p16 = p8 * p8;
    return std::copysign(1.0f - 1.0f / p16, x);
}

// erf^-1 by M. Giles, "Approximating the erfinv function" (GPU Computing
// Gems, 2011), single-precision branch pair in w = -ln((1-x)(1+x)); x = +-1
// (or beyond) gives an infinity of x's sign.
inline float ofast_ierff(float x)
{
    const float q = (1.0f - x) * (1.0f + x);
    if (!(q > 0.0f))
        return x * bits_f(0x7f800000u);
    float w = -ofast_logf(q);
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = fmaf(p, w, 3.43273939e-07f);
        p = fmaf(p, w, -3.5233877e-06f);
        p = fmaf(p, w, -4.39150654e-06f);
        p = fmaf(p, w, 0.00021858087f);
        p = fmaf(p, w, -0.00125372503f);
        p = fmaf(p, w, -0.00417768164f);
        p = fmaf(p, w, 0.246640727f);
        p = fmaf(p, w, 1.50140941f);
    } else {
        w = std::sqrt(w) - 3.0f;
        p = -0.000200214257f;
        p = fmaf(p, w, 0.000100950558f);
        p = fmaf(p, w, 0.00134934322f);
        p = fmaf(p, w, -0.00367342844f);
        p = fmaf(p, w, 0.00573950773f);
        p = fmaf(p, w, -0.0076224613f);
        p = fmaf(p, w, 0.00943887047f);
        p = fmaf(p, w, 1.00167406f);
        p = fmaf(p, w, 2.83297682f);
    }
    return p * x;
}
#endif  // OREF_F64

}  // namespace oref
