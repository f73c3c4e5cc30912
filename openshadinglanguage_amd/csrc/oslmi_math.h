// Device transcendentals with a fixed, fully specified evaluation: range
// reduction plus a Horner polynomial in fmaf. Results are a pure function
// of the IEEE operations written here, so they are reproducible on any
// IEEE-754 binary32 machine that provides a correctly rounded fma.
//
//   exp:    x = k ln2 + r, |r| <= ln2/2 (two-part ln2), Taylor to r^7,
//           2^k applied in two exact scalings (subnormal results formed).
//   sincos: q = rint(x * 2/pi), three-part pi/2 reduction, Taylor sin to
//           r^9 and cos to r^10 on [-pi/4, pi/4], quadrant swap.
//   cbrt:   bit-level seed, three Newton steps.
#pragma once
#include "oslmi_types.h"

namespace oslmi {

OSLMI_DEV float fexp(float x)
{
    if (!(x == x))
        return x;
    if (x > 88.72283935546875f)
        return __uint_as_float(0x7f800000u);
    if (x < -103.97208404541016f)
        return 0.0f;
    const float log2e  = 1.44269502162933349609375f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.428606765330187045e-06f;
    float k            = rintf(x * log2e);
    float r            = __fmaf_rn(-k, ln2_hi, x);
    r                  = __fmaf_rn(-k, ln2_lo, r);
    float p            = 1.0f / 5040.0f;
    p                  = __fmaf_rn(p, r, 1.0f / 720.0f);
    p                  = __fmaf_rn(p, r, 1.0f / 120.0f);
    p                  = __fmaf_rn(p, r, 1.0f / 24.0f);
    p                  = __fmaf_rn(p, r, 1.0f / 6.0f);
    p                  = __fmaf_rn(p, r, 0.5f);
    p                  = __fmaf_rn(p, r, 1.0f);
    p                  = __fmaf_rn(p, r, 1.0f);
    int ki             = (int)k;
    int k1             = ki / 2;
    int k2             = ki - k1;
    float s1           = __uint_as_float((uint32_t)(k1 + 127) << 23);
    float s2           = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

OSLMI_DEV void fsincos(float x, float* s, float* c)
{
    const float two_over_pi = 0.636619746685028076171875f;
    const float pio2_1      = 1.5703125f;
    const float pio2_2      = 4.837512969970703125e-04f;
    const float pio2_3      = 7.549790126404332e-08f;
    float q                 = rintf(x * two_over_pi);
    float r                 = __fmaf_rn(-q, pio2_1, x);
    r                       = __fmaf_rn(-q, pio2_2, r);
    r                       = __fmaf_rn(-q, pio2_3, r);
    float r2                = r * r;
    float ps                = 1.0f / 362880.0f;
    ps                      = __fmaf_rn(ps, r2, -1.0f / 5040.0f);
    ps                      = __fmaf_rn(ps, r2, 1.0f / 120.0f);
    ps                      = __fmaf_rn(ps, r2, -1.0f / 6.0f);
    float sn                = __fmaf_rn(ps * r2, r, r);
    float pc                = -1.0f / 3628800.0f;
    pc                      = __fmaf_rn(pc, r2, 1.0f / 40320.0f);
    pc                      = __fmaf_rn(pc, r2, -1.0f / 720.0f);
    pc                      = __fmaf_rn(pc, r2, 1.0f / 24.0f);
    pc                      = __fmaf_rn(pc, r2, -0.5f);
    float cs                = __fmaf_rn(pc, r2, 1.0f);
    int qi                  = ((int)q) & 3;
    float sv = (qi == 0) ? sn : ((qi == 1) ? cs : ((qi == 2) ? -sn : -cs));
    float cv = (qi == 0) ? cs : ((qi == 1) ? -sn : ((qi == 2) ? -cs : sn));
    *s       = sv;
    *c       = cv;
}

OSLMI_DEV float fcbrt(float x)
{
    if (x == 0.0f || !(x == x))
        return x;
    float ax   = fabsf(x);
    uint32_t g = __float_as_uint(ax) / 3u + 709921077u;
    float y    = __uint_as_float(g);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float y2 = y * y;
        y        = y - (y2 * y - ax) / (3.0f * y2);
    }
    return x < 0.0f ? -y : y;
}

// log2(x), finite x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), series to s^9.
OSLMI_DEV float flog2(float x)
{
    uint32_t u = __float_as_uint(x);
    int e      = 0;
    if (u < 0x00800000u) {
        x = x * 8388608.0f;
        u = __float_as_uint(x);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s  = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p  = 1.0f / 9.0f;
    p        = __fmaf_rn(p, s2, 1.0f / 7.0f);
    p        = __fmaf_rn(p, s2, 1.0f / 5.0f);
    p        = __fmaf_rn(p, s2, 1.0f / 3.0f);
    p        = __fmaf_rn(p, s2, 1.0f);
    float ln = 2.0f * s * p;
    return __fmaf_rn(ln, 1.44269502162933349609375f, (float)e);
}

// ln x for finite x > 0: the reduction of flog2, e ln2 added last.
OSLMI_DEV float flog(float x)
{
    uint32_t u = __float_as_uint(x);
    int e      = 0;
    if (u < 0x00800000u) {
        x = x * 8388608.0f;
        u = __float_as_uint(x);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s  = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p  = 1.0f / 9.0f;
    p        = __fmaf_rn(p, s2, 1.0f / 7.0f);
    p        = __fmaf_rn(p, s2, 1.0f / 5.0f);
    p        = __fmaf_rn(p, s2, 1.0f / 3.0f);
    p        = __fmaf_rn(p, s2, 1.0f);
    float ln = 2.0f * s * p;
    return __fmaf_rn((float)e, 0.693147182464599609375f, ln);
}

// erf by Abramowitz and Stegun 7.1.28: 1 - (1 + a1|x| + ... + a6|x|^6)^-16
// with the sign of x.
OSLMI_DEV float ferf(float x)
{
    const float ax = fabsf(x);
    float p        = 0.0000430638f;
    p              = __fmaf_rn(p, ax, 0.0002765672f);
    p              = __fmaf_rn(p, ax, 0.0001520143f);
    p              = __fmaf_rn(p, ax, 0.0092705272f);
    p              = __fmaf_rn(p, ax, 0.0422820123f);
    p              = __fmaf_rn(p, ax, 0.0705230784f);
    p              = __fmaf_rn(p, ax, 1.0f);
    const float p2 = p * p, p4 = p2 * p2, p8 = p4 * p4, p16 = p8 * p8;
    return copysignf(1.0f - 1.0f / p16, x);
}

// erf^-1 by M. Giles, "Approximating the erfinv function" (GPU Computing
// Gems, 2011), single-precision branch pair in w = -ln((1-x)(1+x)); x = +-1
// (or beyond) gives an infinity of x's sign.
OSLMI_DEV float fierf(float x)
{
    const float q = (1.0f - x) * (1.0f + x);
    if (!(q > 0.0f))
        return x * __uint_as_float(0x7f800000u);
    float w = -flog(q);
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = __fmaf_rn(p, w, 3.43273939e-07f);
        p = __fmaf_rn(p, w, -3.5233877e-06f);
        p = __fmaf_rn(p, w, -4.39150654e-06f);
        p = __fmaf_rn(p, w, 0.00021858087f);
        p = __fmaf_rn(p, w, -0.00125372503f);
        p = __fmaf_rn(p, w, -0.00417768164f);
        p = __fmaf_rn(p, w, 0.246640727f);
        p = __fmaf_rn(p, w, 1.50140941f);
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = __fmaf_rn(p, w, 0.000100950558f);
        p = __fmaf_rn(p, w, 0.00134934322f);
        p = __fmaf_rn(p, w, -0.00367342844f);
        p = __fmaf_rn(p, w, 0.00573950773f);
        p = __fmaf_rn(p, w, -0.0076224613f);
        p = __fmaf_rn(p, w, 0.00943887047f);
        p = __fmaf_rn(p, w, 1.00167406f);
        p = __fmaf_rn(p, w, 2.83297682f);
    }
    return p * x;
}

// 2^t: t = k + f, k = rint(t), 2^f by the Taylor series of e^(f ln2) to
// degree 7, 2^k in two exact scalings.
OSLMI_DEV float fexp2(float t)
{
    if (!(t == t))
        return t;
    if (t >= 128.0f)
        return __uint_as_float(0x7f800000u);
    if (t < -150.0f)
        return 0.0f;
    float k  = rintf(t);
    float f  = t - k;
    float p  = 1.5252733804059840e-05f;
    p        = __fmaf_rn(p, f, 1.5403530393381608e-04f);
    p        = __fmaf_rn(p, f, 1.3333558146428443e-03f);
    p        = __fmaf_rn(p, f, 9.6181291076284772e-03f);
    p        = __fmaf_rn(p, f, 5.5504108664821580e-02f);
    p        = __fmaf_rn(p, f, 2.4022650695910071e-01f);
    p        = __fmaf_rn(p, f, 6.9314718055994531e-01f);
    p        = __fmaf_rn(p, f, 1.0f);
    int ki   = (int)k;
    int k1   = ki / 2;
    int k2   = ki - k1;
    float s1 = __uint_as_float((uint32_t)(k1 + 127) << 23);
    float s2 = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// x^y with fast_safe_pow's special cases (y == 0 -> 1, x == 0 -> 0,
// y == 1 -> x, y == 2 -> min(x*x, FLT_MAX)); x < 0 -> 0; else
// 2^(log2(x) y).
OSLMI_DEV float fsafe_pow(float x, float y)
{
    if (y == 0.0f)
        return 1.0f;
    if (x == 0.0f)
        return 0.0f;
    if (y == 1.0f)
        return x;
    if (y == 2.0f) {
        float x2 = x * x;
        return x2 < 3.40282346638528859812e+38f ? x2 : 3.40282346638528859812e+38f;
    }
    if (x < 0.0f)
        return 0.0f;
    return fexp2(flog2(x) * y);
}

// Dual2 exp and cos (dual.h:830-836, :1122-1127): f(a) with f'(a) * da.
OSLMI_DEV Df dexp(const Df& a)
{
    float f = fexp(a.v);
    return Df(f, f * a.dx, f * a.dy);
}
OSLMI_DEV Df dcos(const Df& a)
{
    float s, c;
    fsincos(a.v, &s, &c);
    return Df(c, -s * a.dx, -s * a.dy);
}

}  // namespace oslmi
