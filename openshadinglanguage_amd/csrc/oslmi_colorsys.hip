// Host-side color system and Gabor parameter construction.
#include "oslmi_colorsys.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace oslmi {

namespace {

struct Chroma {
    float xR, yR, xG, yG, xB, yB, xW, yW;
};
// White points and primaries: src/liboslexec/opcolor.cpp:31-54
const struct {
    const char* name;
    Chroma c;
} kSpaces[] = {
    { "Rec709", { 0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.3127, 0.3291 } },
    { "sRGB", { 0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.3127, 0.3291 } },
    { "NTSC", { 0.67, 0.33, 0.21, 0.71, 0.14, 0.08, 0.3101, 0.3162 } },
    { "EBU", { 0.64, 0.33, 0.29, 0.60, 0.15, 0.06, 0.3127, 0.3291 } },
    { "PAL", { 0.64, 0.33, 0.29, 0.60, 0.15, 0.06, 0.3127, 0.3291 } },
    { "SECAM", { 0.64, 0.33, 0.29, 0.60, 0.15, 0.06, 0.3127, 0.3291 } },
    { "SMPTE", { 0.630, 0.340, 0.310, 0.595, 0.155, 0.070, 0.3127, 0.3291 } },
    { "HDTV", { 0.670, 0.330, 0.210, 0.710, 0.150, 0.060, 0.3127, 0.3291 } },
    { "CIE", { 0.7355, 0.2645, 0.2658, 0.7243, 0.1669, 0.0085, 0.33333333, 0.33333333 } },
    { "AdobeRGB", { 0.64, 0.33, 0.21, 0.71, 0.15, 0.06, 0.3127, 0.3291 } },
    { "XYZ", { 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.33333333, 0.33333333 } },
    { "ACES2065-1", { 0.7347, 0.2653, 0.0, 1.0, 0.0001, -0.077, 0.32168, 0.33767 } },
    { "ACEScg", { 0.713, 0.293, 0.165, 0.83, 0.128, 0.044, 0.32168, 0.33767 } },
};

// CIE 1931 colour matching functions, 380..780 nm every 5 nm
// (src/liboslexec/opcolor_impl.h:55-90).
const float kCIE[243] = {
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

struct H3 {
    float x, y, z;
};
H3 h3(float a, float b, float c) { return H3 { a, b, c }; }
float hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

H3 row_mul(H3 c, const float* m)
{
    return h3(c.x * m[0] + c.y * m[3] + c.z * m[6], c.x * m[1] + c.y * m[4] + c.z * m[7],
              c.x * m[2] + c.y * m[5] + c.z * m[8]);
}

// Planck spectrum integrated against the CIE curves (table build only).
H3 bb_xyz(float T)
{
    float X = 0, Y = 0, Z = 0;
    const float dlambda = 5.0f * 1e-9;
    for (int i = 0; i < 81; ++i) {
        float lambda   = 380.0f + 5.0f * (float)i;
        float wlm      = lambda * 1e-9f;
        float wlm2     = wlm * wlm;
        float wlm4     = wlm2 * wlm2;
        float wlm5     = wlm4 * wlm;
        float inv_wlm5 = 1.0f / wlm5;
        float Me       = ((3.74183e-16f * inv_wlm5) / expm1f(1.4388e-2f / (wlm * T))) * dlambda;
        X += Me * kCIE[3 * i];
        Y += Me * kCIE[3 * i + 1];
        Z += Me * kCIE[3 * i + 2];
    }
    return h3(X, Y, Z);
}

// Imath Matrix33::inverse (cofactors / determinant)
void inverse33(const float* x, float* out)
{
    float s[9] = { x[4] * x[8] - x[7] * x[5], x[7] * x[2] - x[1] * x[8], x[1] * x[5] - x[4] * x[2],
                   x[6] * x[5] - x[3] * x[8], x[0] * x[8] - x[6] * x[2], x[3] * x[2] - x[0] * x[5],
                   x[3] * x[7] - x[6] * x[4], x[6] * x[1] - x[0] * x[7], x[0] * x[4] - x[3] * x[1] };
    float r = x[0] * s[0] + x[1] * s[3] + x[2] * s[6];
    if (fabsf(r) >= 1.0f) {
        for (int i = 0; i < 9; ++i)
            s[i] /= r;
    } else {
        float mr = fabsf(r) / 1.17549435e-38f;
        for (int i = 0; i < 9; ++i) {
            if (mr > fabsf(s[i])) {
                s[i] /= r;
            } else {
                for (int k = 0; k < 9; ++k)
                    out[k] = (k % 4 == 0) ? 1.0f : 0.0f;
                return;
            }
        }
    }
    std::memcpy(out, s, sizeof(s));
}

}  // namespace

bool build_colorsys(const std::string& name, ColorSysDev& cs)
{
    const Chroma* c = nullptr;
    for (auto& s : kSpaces)
        if (name == s.name)
            c = &s.c;
    if (!c)
        return false;
    H3 R = h3(c->xR, c->yR, 0.0f), G = h3(c->xG, c->yG, 0.0f), B = h3(c->xB, c->yB, 0.0f),
       W = h3(c->xW, c->yW, 0.0f);
    R.z  = 1.0f - (R.x + R.y);
    G.z  = 1.0f - (G.x + G.y);
    B.z  = 1.0f - (B.x + B.y);
    W.z  = 1.0f - (W.x + W.y);
    H3 r = h3(G.y * B.z - B.y * G.z, B.x * G.z - G.x * B.z, G.x * B.y - B.x * G.y);
    H3 g = h3(B.y * R.z - R.y * B.z, R.x * B.z - B.x * R.z, B.x * R.y - R.x * B.y);
    H3 b = h3(R.y * G.z - G.y * R.z, G.x * R.z - R.x * G.z, R.x * G.y - G.x * R.y);
    H3 w = h3(hdot(r, W), hdot(g, W), hdot(b, W));
    if (W.y != 0.0f) {
        float k = 1.0f / W.y;
        w       = h3(w.x * k, w.y * k, w.z * k);
    }
    r = h3(r.x / w.x, r.y / w.x, r.z / w.x);
    g = h3(g.x / w.y, g.y / w.y, g.z / w.y);
    b = h3(b.x / w.z, b.y / w.z, b.z / w.z);
    float m[9] = { r.x, g.x, b.x, r.y, g.y, b.y, r.z, g.z, b.z };
    std::memcpy(cs.xyz2rgb, m, sizeof(m));
    inverse33(cs.xyz2rgb, cs.rgb2xyz);
    cs.lum[0]  = cs.rgb2xyz[1];
    cs.lum[1]  = cs.rgb2xyz[4];
    cs.lum[2]  = cs.rgb2xyz[7];
    float lum2 = (1.0f - cs.lum[0] - cs.lum[1]);
    if (fabsf(lum2 - cs.lum[2]) < 0.001f)
        cs.lum[2] = lum2;
    std::memcpy(cs.cie, kCIE, sizeof(kCIE));
    std::memset(cs.bb, 0, sizeof(cs.bb));
    float lastT = 0;
    for (int i = 0; lastT <= 12000.0f; ++i) {
        float is  = sqrtf((float)i);
        float ip  = is * is * is;
        float T   = ip * 2.0f + 800.0f;
        lastT     = T;
        H3 rgb    = row_mul(bb_xyz(T), cs.xyz2rgb);
        rgb.x     = rgb.x < 0.0f ? 0.0f : rgb.x;
        rgb.y     = rgb.y < 0.0f ? 0.0f : rgb.y;
        rgb.z     = rgb.z < 0.0f ? 0.0f : rgb.z;
        cs.bb[3 * i]     = powf(rgb.x, 1.0f / 5.0f);
        cs.bb[3 * i + 1] = powf(rgb.y, 1.0f / 5.0f);
        cs.bb[3 * i + 2] = powf(rgb.z, 1.0f / 5.0f);
    }
    return true;
}

GaborUniform build_gabor_uniform(const GaborOptions& o)
{
    GaborUniform g;
    float bandwidth = std::min(std::max(o.bandwidth, 0.01f), 100.0f);
    float two_bw    = exp2f(bandwidth);
    const float SQRT_PI_OVER_LN2 = 2.128934e+00f;
    g.a          = 2.0f * ((two_bw - 1.0f) / (two_bw + 1.0f)) * SQRT_PI_OVER_LN2;
    g.radius     = sqrtf(-logf(0.02f) / float(M_PI)) / g.a;
    g.radius2    = g.radius * g.radius;
    g.radius3    = g.radius2 * g.radius;
    g.radius_inv = 1.0f / g.radius;
    float imp    = std::min(std::max(o.impulses, 1.0f), 32.0f);
    g.lambda     = imp / (float(1.33333 * M_PI) * g.radius3);
    g.sqrt_lambda_inv = 1.0f / sqrtf(g.lambda);
    g.poisson_g       = expf(-(g.lambda * g.radius3));
    float a3          = g.a * g.a * g.a;
    float gabor_var   = 1.0f / (4.0f * sqrtf(2.0f) * a3);
    float scale       = 1.0f / (3.0f * sqrtf(gabor_var));
    scale *= 0.5f;
    g.scale       = scale;
    g.dir_x       = o.direction[0];
    g.dir_y       = o.direction[1];
    g.dir_z       = o.direction[2];
    g.anisotropic = o.anisotropic;
    g.do_filter   = o.do_filter;
    return g;
}

}  // namespace oslmi
