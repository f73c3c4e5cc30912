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

}  // namespace oslmi
