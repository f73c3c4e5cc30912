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

}  // namespace oref
