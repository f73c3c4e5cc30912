// CDNA4 device library: color system (blackbody, color-space transforms).
//
// Reference: src/liboslexec/opcolor_impl.h (BB_TABLE_UNMAP :163-170,
// lookup_blackbody_rgb :518-555, compute_blackbody_rgb :558-565) and
// opcolor.h:46-58 (XYZ_to_RGB, luminance). The 317-entry blackbody table,
// the XYZ<->RGB matrices and the luminance weights are built once on the
// host (ColorSystem::set_colorspace, opcolor.cpp:142-211) and handed to the
// kernels; kernels stage the table in LDS.
#pragma once
#include "oslmi_math.h"
#include "oslmi_types.h"

namespace oslmi {

#define OSLMI_BB_TABLE 317

// Device-side color system (plain data, copied to constant or HBM).
struct ColorSysDev {
    float xyz2rgb[9];   // row-major m[r][c]
    float rgb2xyz[9];
    float lum[3];
    float cie[243];     // CIE 1931 x/y/z bar, 380..780 nm step 5
    float bb[OSLMI_BB_TABLE * 3];
};

OSLMI_DEV V3 mul_row(V3 c, const float* m)
{
    return V3(c.x * m[0] + c.y * m[3] + c.z * m[6], c.x * m[1] + c.y * m[4] + c.z * m[7],
              c.x * m[2] + c.y * m[5] + c.z * m[8]);
}

OSLMI_DEV V3 xyz_to_rgb(const float* xyz2rgb, V3 c) { return mul_row(c, xyz2rgb); }
OSLMI_DEV float luminance(const float* lum, V3 c) { return c.x * lum[0] + c.y * lum[1] + c.z * lum[2]; }

// fast expm1 with the same structure as the host-side table build is not
// needed here: temperatures >= 12000 K integrate Planck's law with expm1
// formed as exp(x)-1 for |x| >= 0.5 and a Taylor series below.
OSLMI_DEV float fexpm1(float x)
{
    if (fabsf(x) < 0.5f) {
        float p = 1.0f / 362880.0f;
        p       = __fmaf_rn(p, x, 1.0f / 40320.0f);
        p       = __fmaf_rn(p, x, 1.0f / 5040.0f);
        p       = __fmaf_rn(p, x, 1.0f / 720.0f);
        p       = __fmaf_rn(p, x, 1.0f / 120.0f);
        p       = __fmaf_rn(p, x, 1.0f / 24.0f);
        p       = __fmaf_rn(p, x, 1.0f / 6.0f);
        p       = __fmaf_rn(p, x, 0.5f);
        p       = __fmaf_rn(p, x, 1.0f);
        return p * x;
    }
    return fexp(x) - 1.0f;
}

// table: 317*3 floats (LDS or global); xyz2rgb/cie only for T >= 12000 K.
OSLMI_DEV V3 blackbody_rgb(const float* table, const float* xyz2rgb, const float* cie, float T)
{
    if (T < 12000.0f) {
        if (T < 800.0f)
            return V3(1.0e-6f, 0.0f, 0.0f);
        float t  = fcbrt((T - 800.0f) / 2.0f);
        t        = t * t;
        int ti   = (int)t;
        float rm = t - (float)ti;
        const float* a = table + 3 * ti;
        V3 rgb(a[0] * (1.0f - rm) + a[3] * rm, a[1] * (1.0f - rm) + a[4] * rm,
               a[2] * (1.0f - rm) + a[5] * rm);
        V3 rgb2 = rgb * rgb;
        V3 rgb4 = rgb2 * rgb2;
        return rgb4 * rgb;
    }
    float X = 0, Y = 0, Z = 0;
    const float dlambda = 5.0f * 1e-9;
    for (int i = 0; i < 81; ++i) {
        float lambda   = 380.0f + 5.0f * (float)i;
        float wlm      = lambda * 1e-9f;
        float wlm2     = wlm * wlm;
        float wlm4     = wlm2 * wlm2;
        float wlm5     = wlm4 * wlm;
        float inv_wlm5 = 1.0f / wlm5;
        float Me       = ((3.74183e-16f * inv_wlm5) / fexpm1(1.4388e-2f / (wlm * T))) * dlambda;
        X += Me * cie[3 * i];
        Y += Me * cie[3 * i + 1];
        Z += Me * cie[3 * i + 2];
    }
    V3 rgb = mul_row(V3(X, Y, Z), xyz2rgb);
    return V3(rgb.x < 0.0f ? 0.0f : rgb.x, rgb.y < 0.0f ? 0.0f : rgb.y, rgb.z < 0.0f ? 0.0f : rgb.z);
}

// ---- color-space transforms (opcolor_impl.h:246-490) -------------------
// Each transform is written once over a scalar type F, float for values or
// Df for Dual2 colors, so both follow the reference's own evaluation order
// (with derivatives carried by Dual2 arithmetic, dual.h). Spaces: rgb (the
// working space, also "RGB", "linear" and the system's own name), hsv,
// hsl, YIQ, XYZ, xyY, sRGB.
enum ColorSpaceCode { kCS_rgb = 0, kCS_hsv, kCS_hsl, kCS_YIQ, kCS_XYZ, kCS_xyY, kCS_sRGB };

template<class F> struct CTrip {
    F x, y, z;
};
OSLMI_DEV float cs_val(float x) { return x; }
OSLMI_DEV float cs_val(const Df& x) { return x.v; }
OSLMI_DEV CTrip<float> cs_split(V3 c) { return { c.x, c.y, c.z }; }
OSLMI_DEV CTrip<Df> cs_split(const Dv& c) { return { comp(c, 0), comp(c, 1), comp(c, 2) }; }
OSLMI_DEV V3 cs_join(const CTrip<float>& c) { return V3(c.x, c.y, c.z); }
OSLMI_DEV Dv cs_join(const CTrip<Df>& c) { return make_dv(c.x, c.y, c.z); }
// sfm::min_val / max_val (sfmath.h:219-229): the right operand unless the
// comparison picks the left
template<class F> OSLMI_DEV F cs_min(const F& l, const F& r) { return cs_val(r) > cs_val(l) ? l : r; }
template<class F> OSLMI_DEV F cs_max(const F& l, const F& r) { return cs_val(r) > cs_val(l) ? r : l; }

// OpenImageIO's safe_pow: 1 for y == 0, 0 for x == 0 or a negative x with
// a non-integer y, otherwise x^y (as 2^(log2|x| y), negated for a negative
// x and an odd integer y) clamped to the finite range
OSLMI_DEV float cs_safe_pow(float x, float y)
{
    if (y == 0.0f)
        return 1.0f;
    if (x == 0.0f || (x < 0.0f && y != floorf(y)))
        return 0.0f;
    float r = fexp2(flog2(fabsf(x)) * y);
    if (x < 0.0f && fmodf(y, 2.0f) != 0.0f)
        r = -r;
    const float big = 3.40282346638528859812e+38f;
    return r < -big ? -big : (r > big ? big : r);
}
// Dual2 safe_pow with a constant exponent (dual.h:1059-1067): the value is
// formed as u^(v-1) * u
OSLMI_DEV float cs_pow(float u, float v) { return cs_safe_pow(u, v); }
OSLMI_DEV Df cs_pow(const Df& u, float v)
{
    float pm1  = cs_safe_pow(u.v, v - 1.0f);
    float puv  = pm1 * u.v;
    float logu = u.v > 0.0f ? logf(u.v < 1.17549435e-38f ? 1.17549435e-38f : u.v) : 0.0f;
    float dfdu = v * pm1, dfdv = logu * puv;
    return Df(puv, dfdu * u.dx + dfdv * 0.0f, dfdu * u.dy + dfdv * 0.0f);
}
OSLMI_DEV float cs_floor(float x) { return floorf(x); }
OSLMI_DEV float cs_floor(const Df& x) { return floorf(x.v); }  // loses derivatives

// row 3-vector times a 3x3 matrix given row-major (Imath's v * M)
template<class F>
OSLMI_DEV CTrip<F> cs_mul(const CTrip<F>& c, const float m[9])
{
    return { c.x * m[0] + c.y * m[3] + c.z * m[6], c.x * m[1] + c.y * m[4] + c.z * m[7],
             c.x * m[2] + c.y * m[5] + c.z * m[8] };
}

template<class F> OSLMI_DEV CTrip<F> hsv_to_rgb(const CTrip<F>& hsv)
{
    F h = hsv.x, s = hsv.y, v = hsv.z;
    if (cs_val(s) < 0.0001f)
        return { v, v, v };
    h      = 6.0f * (h - cs_floor(h));
    int hi = (int)floorf(cs_val(h));
    F f    = h - F((float)hi);
    F p    = v * (1.0f - s);
    F q    = v * (1.0f - s * f);
    F t    = v * (1.0f - s * (1.0f - f));
    if (hi == 0)
        return { v, t, p };
    if (hi == 1)
        return { q, v, p };
    if (hi == 2)
        return { p, v, t };
    if (hi == 3)
        return { p, q, v };
    if (hi == 4)
        return { t, p, v };
    return { v, p, q };
}

template<class F> OSLMI_DEV CTrip<F> rgb_to_hsv(const CTrip<F>& rgb)
{
    F r = rgb.x, g = rgb.y, b = rgb.z;
    F mincomp = cs_min(r, cs_min(g, b));
    F maxcomp = cs_max(r, cs_max(g, b));
    F delta   = maxcomp - mincomp;
    F v       = maxcomp;
    F s       = F(0.0f);
    if (cs_val(maxcomp) > 0.0f)
        s = delta / maxcomp;
    F h = F(0.0f);
    if (cs_val(s) > 0.0f) {
        float k;
        F x, y;
        if (cs_val(r) >= cs_val(maxcomp)) {
            k = 0.0f / 6.0f, x = g, y = b;
        } else if (cs_val(g) >= cs_val(maxcomp)) {
            k = 2.0f / 6.0f, x = b, y = r;
        } else {
            k = 4.0f / 6.0f, x = r, y = g;
        }
        h = k + (x - y) / (6.0f * delta);
        if (cs_val(h) < 0.0f)
            h = h + 1.0f;
    }
    return { h, s, v };
}

template<class F> OSLMI_DEV CTrip<F> hsl_to_rgb(const CTrip<F>& hsl)
{
    F h = hsl.x, s = hsl.y, l = hsl.z;
    F v = (cs_val(l) <= 0.5f) ? (l * (1.0f + s)) : (l * (1.0f - s) + s);
    if (cs_val(v) <= 0.0f)
        return { F(0.0f), F(0.0f), F(0.0f) };
    F mn = 2.0f * l - v;
    s    = (v - mn) / v;
    return hsv_to_rgb(CTrip<F>{ h, s, v });
}

template<class F> OSLMI_DEV CTrip<F> rgb_to_hsl(const CTrip<F>& rgb)
{
    F minval      = cs_min(rgb.x, cs_min(rgb.y, rgb.z));
    CTrip<F> hsv  = rgb_to_hsv(rgb);
    F maxval      = hsv.z;
    F h = hsv.x, s, l = (minval + maxval) / 2.0f;
    if (cs_val(minval) == cs_val(maxval)) {
        s = F(0.0f);
    } else {
        F min2max = maxval - minval;
        F divisor = cs_val(l) <= 0.5f ? maxval + minval : 2.0f - maxval - minval;
        s         = min2max / divisor;
    }
    return { h, s, l };
}

// YIQ matrices (opcolor_impl.h:408-424), row-major
template<class F> OSLMI_DEV CTrip<F> yiq_to_rgb(const CTrip<F>& c)
{
    const float m[9] = { 1.0000f, 1.0000f, 1.0000f, 0.9557f, -0.2716f, -1.1082f, 0.6199f, -0.6469f, 1.7051f };
    return cs_mul(c, m);
}
template<class F> OSLMI_DEV CTrip<F> rgb_to_yiq(const CTrip<F>& c)
{
    const float m[9] = { 0.299f, 0.596f, 0.212f, 0.587f, -0.275f, -0.523f, 0.114f, -0.321f, 0.311f };
    return cs_mul(c, m);
}

template<class F> OSLMI_DEV CTrip<F> xyz_to_xyY(const CTrip<F>& c)
{
    F n     = c.x + c.y + c.z;
    F n_inv = cs_val(n) >= 1.0e-6f ? 1.0f / n : F(0.0f);
    return { c.x * n_inv, c.y * n_inv, c.y };
}
template<class F> OSLMI_DEV CTrip<F> xyY_to_xyz(const CTrip<F>& c)
{
    F Y   = c.z;
    F Y_y = cs_val(c.y) > 1.0e-6f ? Y / c.y : F(0.0f);
    F X   = Y_y * c.x;
    F Z   = Y_y * (1.0f - c.x - c.y);
    return { X, Y, Z };
}
template<class F> OSLMI_DEV F srgb_to_linear1(const F& x)
{
    return cs_val(x) <= 0.04045f ? x * (1.0f / 12.92f)
                                 : cs_pow((x + 0.055f) * (1.0f / 1.055f), 2.4f);
}
template<class F> OSLMI_DEV F linear_to_srgb1(const F& x)
{
    return cs_val(x) <= 0.0031308f ? 12.92f * x : 1.055f * cs_pow(x, 1.f / 2.4f) - 0.055f;
}

// ColorSystem::transformc (opcolor.cpp:324-380): from -> rgb -> to
template<class F>
OSLMI_DEV CTrip<F> cs_transform(const ColorSysDev* CS, int from, int to, CTrip<F> c)
{
    if (from == kCS_hsv) c = hsv_to_rgb(c);
    else if (from == kCS_hsl) c = hsl_to_rgb(c);
    else if (from == kCS_YIQ) c = yiq_to_rgb(c);
    else if (from == kCS_XYZ) c = cs_mul(c, CS->xyz2rgb);
    else if (from == kCS_xyY) c = cs_mul(xyY_to_xyz(c), CS->xyz2rgb);
    else if (from == kCS_sRGB) c = { srgb_to_linear1(c.x), srgb_to_linear1(c.y), srgb_to_linear1(c.z) };
    if (to == kCS_hsv) c = rgb_to_hsv(c);
    else if (to == kCS_hsl) c = rgb_to_hsl(c);
    else if (to == kCS_YIQ) c = rgb_to_yiq(c);
    else if (to == kCS_XYZ) c = cs_mul(c, CS->rgb2xyz);
    else if (to == kCS_xyY) c = xyz_to_xyY(cs_mul(c, CS->rgb2xyz));
    else if (to == kCS_sRGB) c = { linear_to_srgb1(c.x), linear_to_srgb1(c.y), linear_to_srgb1(c.z) };
    return c;
}
template<int FROM, int TO> OSLMI_DEV V3 transformc(const ColorSysDev* CS, V3 c)
{
    return cs_join(cs_transform<float>(CS, FROM, TO, cs_split(c)));
}
template<int FROM, int TO> OSLMI_DEV Dv transformc(const ColorSysDev* CS, const Dv& c)
{
    return cs_join(cs_transform<Df>(CS, FROM, TO, cs_split(c)));
}

}  // namespace oslmi
