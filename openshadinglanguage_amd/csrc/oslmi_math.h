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
