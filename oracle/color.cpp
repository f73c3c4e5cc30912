// CPU oracle: color system.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// Follows src/liboslexec/opcolor.cpp (chroma table :38-54, set_colorspace
// :142-211) and opcolor_impl.h (CIE table :55-90, BB_TABLE_MAP/UNMAP
// :146-170, bb_spectrum :183-222, spectrum_to_XYZ :226-245,
// lookup_blackbody_rgb :518-555, compute_blackbody_rgb :558-565).
// The table build uses libm expm1f/powf (the reference uses OIIO's
// fast_expm1 and powf); the per-point cube root uses the oracle's
// ofast_cbrtf where the reference calls OIIO::fast_cbrt.
#include "color.h"

#include <cstring>
#include <string>

#include "omath.h"

namespace oref {

namespace {

struct Chroma {
    float xR, yR, xG, yG, xB, yB, xW, yW;
};
#define ILL_C 0.3101, 0.3162
#define ILL_D65 0.3127, 0.3291
#define ILL_E 0.33333333, 0.33333333
#define ILL_ACES 0.32168, 0.33767
const Chroma kChroma[13] = {
    { 0.64, 0.33, 0.30, 0.60, 0.15, 0.06, ILL_D65 },      // Rec709
    { 0.64, 0.33, 0.30, 0.60, 0.15, 0.06, ILL_D65 },      // sRGB
    { 0.67, 0.33, 0.21, 0.71, 0.14, 0.08, ILL_C },        // NTSC
    { 0.64, 0.33, 0.29, 0.60, 0.15, 0.06, ILL_D65 },      // EBU
    { 0.64, 0.33, 0.29, 0.60, 0.15, 0.06, ILL_D65 },      // PAL
    { 0.64, 0.33, 0.29, 0.60, 0.15, 0.06, ILL_D65 },      // SECAM
    { 0.630, 0.340, 0.310, 0.595, 0.155, 0.070, ILL_D65 },  // SMPTE
    { 0.670, 0.330, 0.210, 0.710, 0.150, 0.060, ILL_D65 },  // HDTV
    { 0.7355, 0.2645, 0.2658, 0.7243, 0.1669, 0.0085, ILL_E },  // CIE
    { 0.64, 0.33, 0.21, 0.71, 0.15, 0.06, ILL_D65 },      // AdobeRGB
    { 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, ILL_E },              // XYZ
    { 0.7347, 0.2653, 0.0, 1.0, 0.0001, -0.077, ILL_ACES },  // ACES2065-1
    { 0.713, 0.293, 0.165, 0.83, 0.128, 0.044, ILL_ACES },   // ACEScg
};
const char* kNames[13] = { "Rec709", "sRGB",     "NTSC", "EBU",        "PAL",   "SECAM", "SMPTE",
                           "HDTV",   "CIE",      "AdobeRGB", "XYZ", "ACES2065-1", "ACEScg" };

const float kCIE[81 * 3] = {
    0.0014, 0.0000, 0.0065, 0.0022, 0.0001, 0.0105, 0.0042, 0.0001, 0.0201, 0.0076, 0.0002, 0.0362,
    0.0143, 0.0004, 0.0679, 0.0232, 0.0006, 0.1102, 0.0435, 0.0012, 0.2074, 0.0776, 0.0022, 0.3713,
    0.1344, 0.0040, 0.6456, 0.2148, 0.0073, 1.0391, 0.2839, 0.0116, 1.3856, 0.3285, 0.0168, 1.6230,
    0.3483, 0.0230, 1.7471, 0.3481, 0.0298, 1.7826, 0.3362, 0.0380, 1.7721, 0.3187, 0.0480, 1.7441,
    0.2908, 0.0600, 1.6692, 0.2511, 0.0739, 1.5281, 0.1954, 0.0910, 1.2876, 0.1421, 0.1126, 1.0419,
    0.0956, 0.1390, 0.8130, 0.0580, 0.1693, 0.6162, 0.0320, 0.2080, 0.4652, 0.0147, 0.2586, 0.3533,
    0.0049, 0.3230, 0.2720, 0.0024, 0.4073, 0.2123, 0.0093, 0.5030, 0.1582, 0.0291, 0.6082, 0.1117,
    0.0633, 0.7100, 0.0782, 0.1096, 0.7932, 0.0573, 0.1655, 0.8620, 0.0422, 0.2257, 0.9149, 0.0298,
    0.2904, 0.9540, 0.0203, 0.3597, 0.9803, 0.0134, 0.4334, 0.9950, 0.0087, 0.5121, 1.0000, 0.0057,
    0.5945, 0.9950, 0.0039, 0.6784, 0.9786, 0.0027, 0.7621, 0.9520, 0.0021, 0.8425, 0.9154, 0.0018,
    0.9163, 0.8700, 0.0017, 0.9786, 0.8163, 0.0014, 1.0263, 0.7570, 0.0011, 1.0567, 0.6949, 0.0010,
    1.0622, 0.6310, 0.0008, 1.0456, 0.5668, 0.0006, 1.0026, 0.5030, 0.0003, 0.9384, 0.4412, 0.0002,
    0.8544, 0.3810, 0.0002, 0.7514, 0.3210, 0.0001, 0.6424, 0.2650, 0.0000, 0.5419, 0.2170, 0.0000,
    0.4479, 0.1750, 0.0000, 0.3608, 0.1382, 0.0000, 0.2835, 0.1070, 0.0000, 0.2187, 0.0816, 0.0000,
    0.1649, 0.0610, 0.0000, 0.1212, 0.0446, 0.0000, 0.0874, 0.0320, 0.0000, 0.0636, 0.0232, 0.0000,
    0.0468, 0.0170, 0.0000, 0.0329, 0.0119, 0.0000, 0.0227, 0.0082, 0.0000, 0.0158, 0.0057, 0.0000,
    0.0114, 0.0041, 0.0000, 0.0081, 0.0029, 0.0000, 0.0058, 0.0021, 0.0000, 0.0041, 0.0015, 0.0000,
    0.0029, 0.0010, 0.0000, 0.0020, 0.0007, 0.0000, 0.0014, 0.0005, 0.0000, 0.0010, 0.0004, 0.0000,
    0.0007, 0.0002, 0.0000, 0.0005, 0.0002, 0.0000, 0.0003, 0.0001, 0.0000, 0.0002, 0.0001, 0.0000,
    0.0002, 0.0001, 0.0000, 0.0001, 0.0000, 0.0000, 0.0001, 0.0000, 0.0000, 0.0001, 0.0000, 0.0000,
    0.0000, 0.0000, 0.0000
};

float bb_table_map(float i)
{
    float is = sqrtf(i);
    float ip = is * is * is;
    return ip * 2.0f + 800.0f;
}

float bb_table_unmap(float T)
{
    float t  = (T - 800.0f) / 2.0f;
    float ic = ofast_cbrtf(t);
    return ic * ic;
}

// table build (host, once): libm expm1f; per point (T >= 12000 K): the
// oracle's fixed expm1.
float bb_spectrum(float T, float lambda_nm, bool per_point)
{
    float wlm       = lambda_nm * 1e-9f;
    const float c1  = 3.74183e-16f;
    const float c2  = 1.4388e-2f;
    float wlm2      = wlm * wlm;
    float wlm4      = wlm2 * wlm2;
    float wlm5      = wlm4 * wlm;
    float inv_wlm5  = 1.0f / wlm5;
    float x         = c2 / (wlm * T);
    return (c1 * inv_wlm5) / (per_point ? ofast_expm1f(x) : expm1f(x));
}

V3 spectrum_to_XYZ(float T, bool per_point = false)
{
    float X = 0, Y = 0, Z = 0;
    const float dlambda = 5.0f * 1e-9;
    for (int i = 0; i < 81; ++i) {
        float lambda = 380.0f + 5.0f * (float)i;
        float Me     = bb_spectrum(T, lambda, per_point) * dlambda;
        X += Me * kCIE[3 * i + 0];
        Y += Me * kCIE[3 * i + 1];
        Z += Me * kCIE[3 * i + 2];
    }
    return V3(X, Y, Z);
}

V3 mul_row(V3 c, const float m[3][3])
{
    return V3(c.x * m[0][0] + c.y * m[1][0] + c.z * m[2][0],
              c.x * m[0][1] + c.y * m[1][1] + c.z * m[2][1],
              c.x * m[0][2] + c.y * m[1][2] + c.z * m[2][2]);
}

void clamp_zero(V3& c)
{
    if (c.x < 0.0f)
        c.x = 0.0f;
    if (c.y < 0.0f)
        c.y = 0.0f;
    if (c.z < 0.0f)
        c.z = 0.0f;
}

// Imath Matrix33::inverse(singExc=false) for a non-affine matrix: the
// cofactor matrix divided elementwise by the determinant, with the same
// underflow guard as Matrix22::inverse.
void inverse33(const float x[3][3], float out[3][3])
{
    float s[3][3] = {
        { x[1][1] * x[2][2] - x[2][1] * x[1][2], x[2][1] * x[0][2] - x[0][1] * x[2][2],
          x[0][1] * x[1][2] - x[1][1] * x[0][2] },
        { x[2][0] * x[1][2] - x[1][0] * x[2][2], x[0][0] * x[2][2] - x[2][0] * x[0][2],
          x[1][0] * x[0][2] - x[0][0] * x[1][2] },
        { x[1][0] * x[2][1] - x[2][0] * x[1][1], x[2][0] * x[0][1] - x[0][0] * x[2][1],
          x[0][0] * x[1][1] - x[1][0] * x[0][1] }
    };
    float r = x[0][0] * s[0][0] + x[0][1] * s[1][0] + x[0][2] * s[2][0];
    if (fabsf(r) >= 1.0f) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                s[i][j] /= r;
    } else {
        float mr = fabsf(r) / 1.17549435e-38f;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                if (mr > fabsf(s[i][j])) {
                    s[i][j] /= r;
                } else {
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            out[a][b] = (a == b) ? 1.0f : 0.0f;
                    return;
                }
            }
    }
    std::memcpy(out, s, sizeof(s));
}

}  // namespace

bool colorsys_init(ColorSys& cs, const char* name)
{
    int idx = -1;
    for (int i = 0; i < 13; ++i)
        if (std::string(name) == kNames[i])
            idx = i;
    if (idx < 0)
        return false;
    const Chroma& c = kChroma[idx];
    V3 R(c.xR, c.yR, 0.0f), G(c.xG, c.yG, 0.0f), B(c.xB, c.yB, 0.0f), W(c.xW, c.yW, 0.0f);
    R.z = 1.0f - (R.x + R.y);
    G.z = 1.0f - (G.x + G.y);
    B.z = 1.0f - (B.x + B.y);
    W.z = 1.0f - (W.x + W.y);
    V3 r(G.y * B.z - B.y * G.z, B.x * G.z - G.x * B.z, G.x * B.y - B.x * G.y);
    V3 g(B.y * R.z - R.y * B.z, R.x * B.z - B.x * R.z, B.x * R.y - R.x * B.y);
    V3 b(R.y * G.z - G.y * R.z, G.x * R.z - R.x * G.z, R.x * G.y - G.x * R.y);
    V3 w(dot(r, W), dot(g, W), dot(b, W));
    if (W.y != 0.0f)
        w = w * (1.0f / W.y);
    r = r / w.x;
    g = g / w.y;
    b = b / w.z;
    float m[3][3] = { { r.x, g.x, b.x }, { r.y, g.y, b.y }, { r.z, g.z, b.z } };
    std::memcpy(cs.xyz2rgb, m, sizeof(m));
    inverse33(cs.xyz2rgb, cs.rgb2xyz);
    cs.lum_scale = V3(cs.rgb2xyz[0][1], cs.rgb2xyz[1][1], cs.rgb2xyz[2][1]);
    float lum2   = (1.0f - cs.lum_scale.x - cs.lum_scale.y);
    if (fabsf(lum2 - cs.lum_scale.z) < 0.001f)
        cs.lum_scale.z = lum2;
    float lastT = 0;
    int i       = 0;
    for (; lastT <= 12000.0f; ++i) {
        float T = bb_table_map((float)i);
        lastT   = T;
        V3 rgb  = mul_row(spectrum_to_XYZ(T), cs.xyz2rgb);
        clamp_zero(rgb);
        rgb            = V3(powf(rgb.x, 1.0f / 5.0f), powf(rgb.y, 1.0f / 5.0f), powf(rgb.z, 1.0f / 5.0f));
        cs.bb_table[i] = rgb;
    }
    cs.bb_entries = i;
    for (; i < 317; ++i)
        cs.bb_table[i] = V3(0.0f);
    return true;
}

V3 blackbody_rgb(const ColorSys& cs, float T)
{
    if (T < 12000.0f) {
        if (T < 800.0f)
            return V3(1.0e-6f, 0.0f, 0.0f);
        float t  = bb_table_unmap(T);
        int ti   = (int)t;
        float rm = t - (float)ti;
        V3 a = cs.bb_table[ti], b = cs.bb_table[ti + 1];
        V3 rgb(a.x * (1.0f - rm) + b.x * rm, a.y * (1.0f - rm) + b.y * rm,
               a.z * (1.0f - rm) + b.z * rm);
        V3 rgb2 = rgb * rgb;
        V3 rgb4 = rgb2 * rgb2;
        return rgb4 * rgb;
    }
    V3 rgb = mul_row(spectrum_to_XYZ(T, true), cs.xyz2rgb);
    clamp_zero(rgb);
    return rgb;
}

V3 xyz_to_rgb(const ColorSys& cs, V3 c) { return mul_row(c, cs.xyz2rgb); }

D2<V3> xyz_to_rgb(const ColorSys& cs, const D2<V3>& c)
{
    return D2<V3>(mul_row(c.v, cs.xyz2rgb), mul_row(c.dx, cs.xyz2rgb), mul_row(c.dy, cs.xyz2rgb));
}

float luminance(const ColorSys& cs, V3 c) { return dot(c, cs.lum_scale); }

// ---- color spaces (opcolor_impl.h:246-490, Foley & van Dam) ---------------
namespace {

V3 hsv2rgb(V3 hsv)
{
    float h = hsv.x, s = hsv.y, v = hsv.z;
    if (s < 0.0001f)
        return V3(v, v, v);
    h       = 6.0f * (h - std::floor(h));
    int hi  = (int)std::floor(h);
    float f = h - (float)hi;
    float p = v * (1.0f - s);
    float q = v * (1.0f - s * f);
    float t = v * (1.0f - s * (1.0f - f));
    switch (hi) {
    case 0: return V3(v, t, p);
    case 1: return V3(q, v, p);
    case 2: return V3(p, v, t);
    case 3: return V3(p, q, v);
    case 4: return V3(t, p, v);
    default: return V3(v, p, q);
    }
}

// min / max keeping sfm::min_val / max_val's choice on ties (sfmath.h:219-229)
float mn(float l, float r) { return r > l ? l : r; }
float mx(float l, float r) { return r > l ? r : l; }

V3 rgb2hsv(V3 c)
{
    float lo = mn(c.x, mn(c.y, c.z)), hi = mx(c.x, mx(c.y, c.z));
    float delta = hi - lo, s = 0.0f, h = 0.0f;
    if (hi > 0.0f)
        s = delta / hi;
    if (s > 0.0f) {
        // the reference's form with 1/6 folded into the constants
        float k, x, y;
        if (c.x >= hi) {
            k = 0.0f / 6.0f;
            x = c.y;
            y = c.z;
        } else if (c.y >= hi) {
            k = 2.0f / 6.0f;
            x = c.z;
            y = c.x;
        } else {
            k = 4.0f / 6.0f;
            x = c.x;
            y = c.y;
        }
        h = k + (x - y) / (6.0f * delta);
        if (h < 0.0f)
            h += 1.0f;
    }
    return V3(h, s, hi);
}

V3 hsl2rgb(V3 hsl)
{
    float h = hsl.x, s = hsl.y, l = hsl.z;
    float v = (l <= 0.5f) ? (l * (1.0f + s)) : (l * (1.0f - s) + s);
    if (v <= 0.0f)
        return V3(0.0f);
    float lo = 2.0f * l - v;
    return hsv2rgb(V3(h, (v - lo) / v, v));
}

V3 rgb2hsl(V3 c)
{
    float lo  = mn(c.x, mn(c.y, c.z));
    V3 hsv    = rgb2hsv(c);
    float hi  = hsv.z;
    float l   = (lo + hi) / 2.0f;
    float s   = 0.0f;
    if (lo != hi)
        s = (hi - lo) / (l <= 0.5f ? hi + lo : 2.0f - hi - lo);
    return V3(hsv.x, s, l);
}

// row vector times a row-major 3x3
V3 rowmul(V3 c, const float m[3][3])
{
    return V3(c.x * m[0][0] + c.y * m[1][0] + c.z * m[2][0],
              c.x * m[0][1] + c.y * m[1][1] + c.z * m[2][1],
              c.x * m[0][2] + c.y * m[1][2] + c.z * m[2][2]);
}
const float kYIQ2RGB[3][3] = { { 1.0000f, 1.0000f, 1.0000f },
                               { 0.9557f, -0.2716f, -1.1082f },
                               { 0.6199f, -0.6469f, 1.7051f } };
const float kRGB2YIQ[3][3] = { { 0.299f, 0.596f, 0.212f },
                               { 0.587f, -0.275f, -0.523f },
                               { 0.114f, -0.321f, 0.311f } };

V3 xyz2xyY(V3 c)
{
    float n   = c.x + c.y + c.z;
    float inv = n >= 1.0e-6f ? float(1.0f / n) : float(0.0f);
    return V3(c.x * inv, c.y * inv, c.y);
}
V3 xyY2xyz(V3 c)
{
    float Y_y = c.y > 1.0e-6f ? float(c.z / c.y) : float(0.0f);
    return V3(Y_y * c.x, c.z, Y_y * (1.0f - c.x - c.y));
}

// OpenImageIO safe_pow, with the oracle's own pow (omath.h)
float spow(float x, float y) { return osafe_powf(x, y); }
float s2lin(float x)
{
    return x <= 0.04045f ? x * (1.0f / 12.92f) : spow((x + 0.055f) * (1.0f / 1.055f), 2.4f);
}
float lin2s(float x) { return x <= 0.0031308f ? 12.92f * x : 1.055f * spow(x, 1.f / 2.4f) - 0.055f; }

}  // namespace

int color_space_code(const char* name, bool tc)
{
    const std::string n(name);
    if (n == "rgb" || n == "RGB" || n == "Rec709" || (tc && n == "linear"))
        return 0;
    const char* names[] = { "hsv", "hsl", "YIQ", "XYZ", "xyY" };
    for (int k = 0; k < 5; ++k)
        if (n == names[k])
            return k + 1;
    if (tc && n == "sRGB")
        return 6;
    return -1;
}

V3 transformc(const ColorSys& cs, int from, int to, V3 c)
{
    V3 rgb = c;
    switch (from) {
    case 1: rgb = hsv2rgb(c); break;
    case 2: rgb = hsl2rgb(c); break;
    case 3: rgb = rowmul(c, kYIQ2RGB); break;
    case 4: rgb = rowmul(c, cs.xyz2rgb); break;
    case 5: rgb = rowmul(xyY2xyz(c), cs.xyz2rgb); break;
    case 6: rgb = V3(s2lin(c.x), s2lin(c.y), s2lin(c.z)); break;
    default: break;
    }
    switch (to) {
    case 1: return rgb2hsv(rgb);
    case 2: return rgb2hsl(rgb);
    case 3: return rowmul(rgb, kRGB2YIQ);
    case 4: return rowmul(rgb, cs.rgb2xyz);
    case 5: return xyz2xyY(rowmul(rgb, cs.rgb2xyz));
    case 6: return V3(lin2s(rgb.x), lin2s(rgb.y), lin2s(rgb.z));
    default: return rgb;
    }
}

}  // namespace oref
