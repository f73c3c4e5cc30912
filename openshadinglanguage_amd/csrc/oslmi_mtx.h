// MaterialX layered closure on the device: libbsdl's MTX dielectric,
// generalized-Schlick and sheen lobes with their energy-compensation
// tables, flattened and MIS-combined the way testrender's CompositeBSDF
// does for layer(sheen_bsdf, layer(dielectric_bsdf, generalized_schlick_bsdf)).
//
//   GGXDist D / G1 / G2_G1 / sample / sample_for_refl / D_refl_D
//                                  libbsdl/include/BSDL/microfacet_tools_impl.h:19-118
//   TabulatedEnergyCurve           microfacet_tools_impl.h:120-218
//   Turquin multiple-scattering reflection   microfacet_tools_impl.h:345-394
//   Frame, square_to_unit_disc, sample_uniform_hemisphere, reflect
//                                  libbsdl/include/BSDL/tools.h:209-283, :307-318, :468-541
//   visible_normal, regularize_roughness     libbsdl/include/BSDL/bsdf_impl.h:48-83, :155-160
//   DielectricFresnel / DielectricLobe       MTX/bsdf_dielectric_impl.h:21-310
//   SchlickFresnel / SchlickLobe             MTX/bsdf_schlick_impl.h:14-141
//   SheenLobe (Conty-Kulla, non-legacy shadowing)   MTX/bsdf_sheen_impl.h:19-210
//   world/local wrapping           src/testrender/shading.cpp:73-106
//   layer opacity and weights      src/testrender/shading.cpp:1199-1272, :1645-1662
//   CompositeBSDF prepare/eval/sample        src/testrender/shading.h:319-393
//   MIS::update_eval               src/testrender/sampling.h:173-206
//
// Every function is a sequence of IEEE binary32 operations in the order
// of the reference expressions (the library is built with
// -ffp-contract=off), so the CPU restatement gives identical bits. The
// run-time trigonometry is fsincos and pow is fsafe_pow (oslmi_math.h),
// standing in for OpenImageIO's fast_cos/fast_sin/fast_safe_pow of
// testrender's BSDLConfig. Functions marked OSLMI_HD are also compiled for
// the host, where the library bakes the energy tables with libm
// trigonometry as libbsdl's genluts does.
//
// Only reflection is modelled (transmission tints must be zero); the
// library rejects other materials.
#pragma once
#include "oslmi_math.h"
#include "oslmi_types.h"

#if !defined(__HIPCC_RTC__)
#include <cmath>
#endif

namespace oslmi {
namespace mtx {

constexpr float kPI        = 3.14159265358979323846f;
constexpr float kONEOVERPI = 1.0f / kPI;
constexpr float kEPSILON   = 1e-4f;
constexpr float kIOR_MIN   = 1.001f;
constexpr float kIOR_MAX   = 5.0f;
constexpr float kFLT_MIN   = 1.17549435e-38f;

// table geometry (16 cosines x 16 roughnesses x 1 or 32 Fresnel indices)
constexpr int kNc = 16, kNr = 16, kNfRefl = 32;
constexpr int kLutMiniGGX   = 0;
constexpr int kLutReflFront = kNc * kNr;
constexpr int kLutFloats    = kNc * kNr * (1 + kNfRefl);

OSLMI_HD float SQR(float x) { return x * x; }
// std::max / std::min, and libbsdl's MIN template; libbsdl's MAX on floats
// is std::max. They differ on ties of signed zeros and on NaN.
OSLMI_HD float smax(float a, float b) { return a < b ? b : a; }
OSLMI_HD float smin(float a, float b) { return b < a ? b : a; }
OSLMI_HD float bmin(float a, float b) { return a < b ? a : b; }
OSLMI_HD float CLAMP(float x, float a, float b) { return bmin(smax(x, a), b); }
OSLMI_HD float LERP(float f, float a, float b)
{
    f = CLAMP(f, 0.0f, 1.0f);
    return (1 - f) * a + f * b;
}
OSLMI_HD V3 LERP3(float f, V3 a, V3 b)
{
    f = CLAMP(f, 0.0f, 1.0f);
    return a * (1 - f) + b * f;
}
OSLMI_HD float MAX_RGB(V3 c) { return smax(c.x, smax(c.y, c.z)); }
OSLMI_HD float MAX_ABS_XYZ(V3 v) { return smax(fabsf(v.x), smax(fabsf(v.y), fabsf(v.z))); }
OSLMI_HD float pmax(V3 c) { return smax(smax(c.x, c.y), c.z); }
OSLMI_HD float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Imath Vec3 length (with the lengthTiny branch) and normalized()
OSLMI_HD float vlength(V3 v)
{
    float l2 = vdot(v, v);
    if (l2 < 2.0f * kFLT_MIN) {
        float ax = fabsf(v.x), ay = fabsf(v.y), az = fabsf(v.z);
        float m = ax;
        if (m < ay)
            m = ay;
        if (m < az)
            m = az;
        if (m == 0.0f)
            return 0.0f;
        ax = ax / m;
        ay = ay / m;
        az = az / m;
        return m * sqrtf(ax * ax + ay * ay + az * az);
    }
    return sqrtf(l2);
}
OSLMI_HD V3 normalized(V3 v)
{
    float l = vlength(v);
    if (l == 0.0f)
        return V3(0.0f);
    return V3(v.x / l, v.y / l, v.z / l);
}
OSLMI_HD V3 reflect(V3 E, V3 N) { return N * (2 * vdot(N, E)) - E; }

struct Frame {
    V3 X, Y, Z;
    OSLMI_HD V3 local(V3 a) const { return V3(vdot(a, X), vdot(a, Y), vdot(a, Z)); }
    OSLMI_HD V3 world(V3 a) const
    {
        return V3(X.x * a.x + Y.x * a.y + Z.x * a.z, X.y * a.x + Y.y * a.y + Z.y * a.z,
                  X.z * a.x + Y.z * a.y + Z.z * a.z);
    }
};
// Duff et al. orthonormal basis (tools.h:468-481)
OSLMI_HD void ortho_build(V3 Z, V3* X, V3* Y)
{
    const float s = copysignf(1.0f, Z.z);
    const float a = -1.0f / (s + Z.z);
    const float b = Z.x * Z.y * a;
    *X            = V3(1.0f + s * Z.x * Z.x * a, s * b, -s * Z.x);
    *Y            = V3(b, s + Z.y * Z.y * a, -Z.y);
}
OSLMI_HD Frame frame_zx(V3 Z, V3 X)
{
    Frame f;
    f.Z = Z;
    f.X = X;
    if (MAX_ABS_XYZ(X) < 1e-4f || fabsf(vdot(Z, normalized(X))) > 0.999f) {
        ortho_build(Z, &f.X, &f.Y);
    } else {
        f.Y = normalized(cross(Z, X));
        f.X = cross(f.Y, Z);
    }
    return f;
}
OSLMI_HD V3 visible_normal(V3 wo, V3 Ngf, V3 N)
{
    if (vdot(wo, N) > 0.0f)
        return N;
    V3 V = cross(wo, N);
    if (MAX_ABS_XYZ(V) < 1e-4f) {
        V = cross(wo, Ngf);
        if (MAX_ABS_XYZ(V) < 1e-4f) {
            const float s = copysignf(1.0f, wo.z);
            const float a = -1.0f / (s + wo.z);
            V             = V3(wo.x * wo.y * a, s + SQR(wo.y) * a, -wo.y);
        } else
            V = normalized(V);
    } else
        V = normalized(V);
    return normalized(cross(V, wo) + 1e-4f * wo);
}
// camera rays: path_roughness = 0
OSLMI_HD float regularize_roughness(float r)
{
    const float roughness_product = 1.0f - 0.0f;
    return 1.0f - (1.0f - r) * roughness_product;
}

// Clarberg's quadrant polynomials for cos(x pi/4), sin(x pi/4)
OSLMI_HD float fast_cos_quadrant(float x)
{
    const float c1 = 0.99998736f;
    const float c2 = -0.30837047f;
    const float c3 = 0.01578646f;
    const float c4 = -0.00029826362f;
    float x2       = x * x;
    float cp       = c3 + c4 * x2;
    cp             = c2 + cp * x2;
    cp             = c1 + cp * x2;
    return cp;
}
OSLMI_HD float fast_sin_quadrant(float x)
{
    const float s1 = 0.7853975892066955566406250000000000f;
    const float s2 = -0.0807407423853874206542968750000000f;
    const float s3 = 0.0024843954015523195266723632812500f;
    const float s4 = -0.0000341485538228880614042282104492f;
    float x2       = x * x;
    float sp       = s3 + s4 * x2;
    sp             = s2 + sp * x2;
    sp             = s1 + sp * x2;
    return sp * x;
}
// [0,1)^2 -> unit disc (concentric, quadrant polynomials): (x, y) in *dx, *dy
OSLMI_HD void square_to_disc(float u, float v, float* dx, float* dy, float* rad_out)
{
    const float a = 2 * u - 1, qa = fabsf(a);
    const float b = 2 * v - 1, qb = fabsf(b);
    const float rad = qa > qb ? qa : qb;
    const float phi = qa > qb ? qb / qa : ((qa == qb) ? 1.0f : 2 - qa / qb);
    *dx             = copysignf(rad * fast_cos_quadrant(phi), a);
    *dy             = copysignf(rad * fast_sin_quadrant(phi), b);
    *rad_out        = rad;
}
OSLMI_HD V3 sample_uniform_hemisphere(float u, float v)
{
    float x, y, rad;
    square_to_disc(u, v, &x, &y, &rad);
    const float cos_theta = 1 - rad * rad;
    const float sin_theta = sqrtf(2 - rad * rad);
    return V3(sin_theta * x, sin_theta * y, cos_theta);
}

// ---- GGX --------------------------------------------------------------
struct GGX {
    float ax, ay;
};
OSLMI_HD GGX ggx_make(float rough, float aniso, bool flip)
{
    GGX d;
    d.ax = SQR(rough);
    d.ay = d.ax;
    if (flip)
        aniso = -aniso;
    d.ax = smax(d.ax * (1 + aniso), 1e-5f);
    d.ay = smax(d.ay * (1 - aniso), 1e-5f);
    return d;
}
OSLMI_HD float ggx_D(const GGX& d, V3 Hr)
{
    const float cosPhi2st2 = SQR(Hr.x / d.ax);
    const float sinPhi2st2 = SQR(Hr.y / d.ay);
    const float cosThetaM2 = SQR(Hr.z);
    const float sinThetaM2 = cosPhi2st2 + sinPhi2st2;
    return 1.0f / (kPI * d.ax * d.ay * SQR(cosThetaM2 + sinThetaM2));
}
OSLMI_HD float ggx_G1(const GGX& d, V3 w)
{
    w = V3(w.x * d.ax, w.y * d.ay, w.z);
    return 2.0f * w.z / (w.z + vlength(w));
}
OSLMI_HD float ggx_G2_G1(const GGX& d, V3 wi, V3 wo)
{
    wi             = V3(wi.x * d.ax, wi.y * d.ay, wi.z);
    wo             = V3(wo.x * d.ax, wo.y * d.ay, wo.z);
    const float nl = vlength(wi);
    const float nv = vlength(wo);
    return wi.z * (wo.z + nv) / (wo.z * nl + wi.z * nv);
}
OSLMI_HD float ggx_D_refl_D(const GGX& d, V3 wo)
{
    const float aix  = wo.x * d.ax, aiy = wo.y * d.ay;
    const float len2 = aix * aix + aiy * aiy;
    const float t    = sqrtf(len2 + SQR(wo.z));
    const float a    = CLAMP(smin(d.ax, d.ay), 0.0f, 1.0f);
    const float s    = 1 + sqrtf(SQR(wo.x) + SQR(wo.y));
    const float a2   = SQR(a);
    const float s2   = SQR(s);
    const float k    = (1 - a2) * s2 / (s2 + a2 * SQR(wo.z));
    return 2 * wo.z / (k * wo.z + t);
}
// visible-normal sampling (Walter), used by the MiniMicrofacetGGX bake
OSLMI_HD V3 ggx_sample_vndf(const GGX& d, V3 wo, float randu, float randv)
{
    const V3 V  = normalized(V3(d.ax * wo.x, d.ay * wo.y, wo.z));
    const V3 T1 = V.z < 0.9999f ? normalized(V3(V.y, -V.x, 0)) : V3(1, 0, 0);
    const V3 T2 = cross(T1, V);
    float px, py, rad;
    square_to_disc(randu, randv, &px, &py, &rad);
    const float s   = 0.5f * (1 + V.z);
    const float p2o = s * py + (1 - s) * sqrtf(1 - px * px);
    const float p3  = sqrtf(smax(1.0f - SQR(px) - SQR(p2o), 0.0f));
    const V3 N      = px * T1 + p2o * T2 + p3 * V;
    return normalized(V3(d.ax * N.x, d.ay * N.y, smax(N.z, 0.0f)));
}
// bounded visible-normal sampling for reflection (Eto-Tokuyoshi); Trig::sc
// gives sin and cos of the azimuth
template<class Trig> OSLMI_HD V3 ggx_sample_bvndf(const GGX& d, V3 wo, float randu, float randv)
{
    V3 i_std             = normalized(V3(wo.x * d.ax, wo.y * d.ay, wo.z));
    const float phi      = 2.0f * kPI * randu;
    const float a        = CLAMP(smin(d.ax, d.ay), 0.0f, 1.0f);
    const float s        = 1 + sqrtf(SQR(wo.x) + SQR(wo.y));
    const float a2       = SQR(a);
    const float s2       = SQR(s);
    const float k        = (1 - a2) * s2 / (s2 + a2 * SQR(wo.z));
    const float b        = k * i_std.z;
    const float z        = (1 - randv) * (1 + b) - b;
    const float sinTheta = sqrtf(CLAMP(1 - SQR(z), 0.0f, 1.0f));
    float sn, c;
    Trig::sc(phi, &sn, &c);
    V3 o_std = V3(sinTheta * c, sinTheta * sn, z);
    V3 m_std = i_std + o_std;
    return normalized(V3(m_std.x * d.ax, m_std.y * d.ay, m_std.z));
}
struct TrigFast {
    OSLMI_DEV static void sc(float x, float* s, float* c) { fsincos(x, s, c); }
};

// ---- energy tables ----------------------------------------------------
OSLMI_HD float get_cosine(int table, int i)
{
    const float x = float(i) * (1.0f / 15.0f);
    return table == kLutMiniGGX ? smax(SQR(x), 1e-6f) : smax(x, 1e-6f);
}
// Row indices are clamped into the table (the reference asserts
// 0 <= roughness, fresnel_index <= 1, where the clamps change nothing), so
// no input reads outside it.
OSLMI_HD int clamp_index(int k, int hi) { return k < 0 ? 0 : (k > hi ? hi : k); }
OSLMI_HD float interpolate_emiss(const float* E, int Nf, float roughness, float fresnel_index,
                                 int i)
{
    float rf = roughness * (kNr - 1);
    int ra   = clamp_index(static_cast<int>(rf), kNr - 1);
    int rb   = ra + 1 < kNr - 1 ? ra + 1 : kNr - 1;
    rf -= ra;
    if (Nf == 1)
        return LERP(rf, E[ra * kNc + i], E[rb * kNc + i]);
    float ff = fresnel_index * (Nf - 1);
    int fa   = clamp_index(static_cast<int>(ff), Nf - 1);
    int fb   = fa + 1 < Nf - 1 ? fa + 1 : Nf - 1;
    ff -= fa;
    return LERP(ff, LERP(rf, E[(fa * kNr + ra) * kNc + i], E[(fa * kNr + rb) * kNc + i]),
                LERP(rf, E[(fb * kNr + ra) * kNc + i], E[(fb * kNr + rb) * kNc + i]));
}
// TabulatedEnergyCurve::Emiss_eval on table `table` of the table set
OSLMI_HD float emiss_eval(const float* tables, int table, float roughness, float fresnel_index,
                          float c)
{
    const float* E = tables + table;
    const int Nf   = table == kLutMiniGGX ? 1 : kNfRefl;
    float cos0     = get_cosine(table, 0);
    if (c <= cos0)
        return interpolate_emiss(E, Nf, roughness, fresnel_index, 0);
    for (int i = 1; i < kNc; i++) {
        const float cos1 = get_cosine(table, i);
        if (c < cos1) {
            float q = (c - cos0) / (cos1 - cos0);
            return LERP(q, interpolate_emiss(E, Nf, roughness, fresnel_index, i - 1),
                        interpolate_emiss(E, Nf, roughness, fresnel_index, i));
        }
        cos0 = cos1;
    }
    return interpolate_emiss(E, Nf, roughness, fresnel_index, kNc - 1);
}

// ---- Fresnel ------------------------------------------------------------
OSLMI_HD float dielectric_fresnel(float eta, float c)
{
    float g = (eta - 1.0f) * (eta + 1.0f) + c * c;
    if (g > 0) {
        g       = sqrtf(g);
        float A = (g - c) / (g + c);
        float B = (c * (g + c) - 1) / (c * (g - c) + 1);
        return 0.5f * A * A * (1 + B * B);
    }
    return 1.0f;
}
OSLMI_HD float dfresnel_eta(float eta, bool backside)
{
    if (backside)
        eta = 1 / eta;
    return eta >= 1 ? CLAMP(eta, kIOR_MIN, kIOR_MAX) : CLAMP(eta, 1 / kIOR_MAX, 1 / kIOR_MIN);
}
OSLMI_HD float dfresnel_table_index(float eta)
{
    const float feta = eta;
    const float seta = CLAMP(feta < 1 ? 1 / feta : feta, kIOR_MIN, kIOR_MAX);
    const float x    = (seta - kIOR_MIN) * (1 / (kIOR_MAX - kIOR_MIN));
    return sqrtf(x);
}

// microfacet reflection state of the two MTX microfacet lobes: GGX,
// Fresnel (kind 0 dielectric, 1 generalized Schlick) and the
// MiniMicrofacetGGX multiple-scattering energy E_ms of the lobe's
// DielectricBSDF
struct Spec {
    GGX d;
    int kind;
    float eta;
    V3 F0, F90;
    float exponent, tir_cos;
    float E_ms;
    OSLMI_DEV V3 fresnel(float c) const
    {
        if (kind == 0)
            return V3(dielectric_fresnel(eta, c));
        c = CLAMP(c, 0.0f, 1.0f);
        if (c < tir_cos)
            return V3(1.0f);
        return LERP3(fsafe_pow(1 - c, exponent), F0, F90);
    }
};
// the dielectric-only variant the table bake runs on the host
struct SpecDielectric {
    GGX d;
    float eta;
    float E_ms;
    OSLMI_HD V3 fresnel(float c) const { return V3(dielectric_fresnel(eta, c)); }
};

struct Sample {
    V3 wi;
    V3 weight;
    float pdf;
    float roughness;
};
OSLMI_HD Sample zero_sample()
{
    Sample s;
    s.wi        = V3(0.0f);
    s.weight    = V3(0.0f);
    s.pdf       = 0.0f;
    s.roughness = 0.0f;
    return s;
}

// eval_turquin_microms_reflection (microfacet_tools_impl.h:345-378)
template<class S> OSLMI_HD Sample eval_turquin(const S& sp, V3 wo, V3 wi)
{
    const float cosNO = wo.z;
    const float cosNI = wi.z;
    if (cosNI <= 0 || cosNO <= 0)
        return zero_sample();
    V3 m            = normalized(wo + wi);
    float cosMO     = vdot(m, wo);
    const float D   = ggx_D(sp.d, m);
    float D_refl_D  = ggx_D_refl_D(sp.d, wo);
    float D_refl    = D_refl_D * D;
    const float G1  = ggx_G1(sp.d, wo);
    float pdf       = D_refl / (4 * cosNO);
    const float out = ggx_G2_G1(sp.d, wi, wo) * G1 / D_refl_D;
    const V3 F      = sp.fresnel(cosMO);
    const float msf = sp.E_ms / smax(0.01f, 1 - sp.E_ms);
    Sample s;
    s.wi        = wi;
    s.weight    = (F * out) * (V3(1.0f) + F * msf);
    s.pdf       = pdf;
    s.roughness = -1.0f;
    return s;
}

// ---- sheen (Conty-Kulla) --------------------------------------------------
OSLMI_DEV float conty_kulla_albedo(float a, float cosNO)
{
    float rx = 13.67300f, ry = 1.0f;
    rx = rx + -68.78018f * cosNO;
    ry = ry + 61.57746f * cosNO;
    rx = rx + 799.08825f * a;
    ry = ry + 442.78211f * a;
    rx = rx + -905.00061f * cosNO * a;
    ry = ry + 2597.49308f * cosNO * a;
    rx = rx + 60.28956f * cosNO * cosNO;
    ry = ry + 121.81241f * cosNO * cosNO;
    rx = rx + 1086.96473f * a * a;
    ry = ry + 3045.55075f * a * a;
    return CLAMP(rx / ry, 0.0f, 1.0f);
}
OSLMI_DEV float conty_kulla_D(float a, V3 Hr)
{
    float cos_theta = CLAMP(Hr.z, 0.0f, 1.0f);
    float sin_theta = sqrtf(1.0f - SQR(cos_theta));
    return fsafe_pow(sin_theta, 1 / a) * (2 + 1 / a) * 0.5f * kONEOVERPI;
}
// SheenMicrofacet<ContyKullaDist<false>>::eval, weight times the tint
OSLMI_DEV Sample sheen_mf_eval(float a, V3 tint, V3 wo, V3 wi)
{
    const float cosNO = wo.z;
    const float cosNI = wi.z;
    if (cosNI <= 1e-5f || cosNO <= 1e-5f)
        return zero_sample();
    const float D = conty_kulla_D(a, normalized(wo + wi));
    if ((double)D < 1e-6)
        return zero_sample();
    const float ni = smin(1.0f, wi.z);
    const float no = smin(1.0f, wo.z);
    const float G2 = (ni * no) / (ni + no - ni * no);
    Sample s;
    s.wi        = wi;
    s.weight    = V3(D * G2 * 0.5f * kPI / cosNO) * tint;
    s.pdf       = 0.5f * kONEOVERPI;
    s.roughness = 0.0f;
    return s;
}

// ---- one MaterialX lobe of a closure list --------------------------------
// A libbsdl MTX lobe as testrender's BSDL_WRAP builds it at wo (the
// renderer's -I; src/testrender/shading.cpp:73-106): the lobe frame
// (visible normal, tangent U; the sheen's X is wo), wo in that frame and
// the state its eval, sample, albedo and layer opacity read. Front facing,
// path roughness 0, reflection only.
enum MxKind { kMxDielectric = 0, kMxSchlick = 1, kMxSheen = 2 };
struct MxLobe {
    Frame f;
    V3 wol;     // wo in f
    V3 tint;    // reflection tint; sheen albedo
    Spec sp;    // dielectric / generalized Schlick specular
    float E;    // dielectric / Schlick E_ms (0 without reflection); sheen Emiss
    float a;    // sheen ContyKulla alpha
    int dorefl;
};

// DielectricLobe (kind 0) / SchlickLobe (kind 1) constructors
// (MTX/bsdf_dielectric_impl.h:210-262, MTX/bsdf_schlick_impl.h:14-60):
// anisotropic GGX from roughness_x / roughness_y, Fresnel, the
// MiniMicrofacetGGX and DielectricReflFront energies at cos(N, wo).
OSLMI_DEV void mx_micro(MxLobe* M, int kind, V3 N, V3 U, V3 refl_tint, float roughness_x,
                        float roughness_y, float ior, V3 F0, V3 F90, float exponent, V3 wo,
                        const float* T)
{
    M->f      = frame_zx(visible_normal(wo, N, N), U);
    M->wol    = M->f.local(wo);
    M->tint   = refl_tint;
    M->dorefl = MAX_RGB(refl_tint) > 0;
    M->a      = 0.0f;
    const float rx    = CLAMP(roughness_x, kEPSILON, 2.0f);
    const float ry    = CLAMP(roughness_y, kEPSILON, 2.0f);
    const float ax    = smax(rx, ry);
    const float ay    = smin(rx, ry);
    const float b     = ay / ax;
    const float aniso = (1 - b) / (1 + b);
    const float rough = regularize_roughness(sqrtf(ax / (1 + aniso)));
    const float cosNO = vdot(M->f.Z, wo);
    if (kind == kMxDielectric) {
        M->sp.kind     = 0;
        M->sp.eta      = dfresnel_eta(CLAMP(ior, kIOR_MIN, kIOR_MAX) / 1.0f, false);
        M->sp.F0       = V3(0.0f);
        M->sp.F90      = V3(0.0f);
        M->sp.exponent = 0.0f;
        M->sp.tir_cos  = 0.0f;
    } else {
        const float avg_F0 = CLAMP((F0.x + F0.y + F0.z) * (1.0f / 3), 0.0f, 0.99f);
        const float sqF0   = sqrtf(avg_F0);
        M->sp.kind         = 1;
        M->sp.eta          = dfresnel_eta((1 + sqF0) / (1 - sqF0), false);
        M->sp.F0 = V3(smax(0.0f, smin(1.0f, F0.x)), smax(0.0f, smin(1.0f, F0.y)), smax(0.0f, smin(1.0f, F0.z)));
        M->sp.F90 = V3(smax(0.0f, smin(1.0f, F90.x)), smax(0.0f, smin(1.0f, F90.y)),
                       smax(0.0f, smin(1.0f, F90.z)));
        M->sp.exponent = exponent;
        M->sp.tir_cos  = M->sp.eta >= 1 ? 0 : sqrtf(1 - SQR(M->sp.eta));
    }
    M->sp.d    = ggx_make(rough, aniso, rx < ry);
    M->sp.E_ms = emiss_eval(T, kLutMiniGGX, rough, 0.0f, cosNO);
    M->E = M->dorefl ? emiss_eval(T, kLutReflFront, rough, dfresnel_table_index(M->sp.eta), cosNO)
                     : 0.0f;
}

// SheenLobe constructor, Conty-Kulla mode (MTX/bsdf_sheen_impl.h:19-48)
OSLMI_DEV void mx_sheen(MxLobe* M, V3 N, V3 albedo, float roughness, V3 wo)
{
    M->f      = frame_zx(visible_normal(wo, N, N), wo);
    M->wol    = M->f.local(wo);
    M->tint   = albedo;
    M->dorefl = 1;
    const float s_rough = regularize_roughness(CLAMP(roughness, 0.0f, 1.0f));
    M->a                = CLAMP(smax(0.06f, s_rough), 0.06f, 1.0f);
    const float cosNO   = CLAMP(vdot(M->f.Z, wo), 0.0f, 1.0f);
    M->E                = 1 - smin(conty_kulla_albedo(M->a, cosNO) * pmax(M->tint), 1.0f);
}

// eval_impl at wi in the lobe frame (wo: the built one)
OSLMI_DEV Sample mx_eval_local(const MxLobe& M, int kind, V3 wi)
{
    if (kind == kMxSheen) {
        if (!(wi.z > 0 && M.wol.z >= 0))
            return zero_sample();
        return sheen_mf_eval(M.a, M.tint, M.wol, wi);
    }
    if (!M.dorefl)
        return zero_sample();
    Sample s = eval_turquin(M.sp, M.wol, wi);
    s.weight = s.weight * (s.wi.z > 0 ? M.tint : V3(0.0f));
    return s;
}

// sample_impl's direction in the lobe frame (false: no sample); the
// sample's weight and pdf are mx_eval_local there
OSLMI_DEV bool mx_sample_dir(const MxLobe& M, int kind, float rx, float ry, V3* wl)
{
    if (kind == kMxSheen) {
        *wl = sample_uniform_hemisphere(rx, ry);
        return true;
    }
    if (!M.dorefl || M.wol.z <= 0)
        return false;
    const V3 m = ggx_sample_bvndf<TrigFast>(M.sp.d, M.wol, rx, ry);
    *wl        = reflect(M.wol, m);
    return !(wl->z <= 0 || MAX_ABS_XYZ(*wl) < kEPSILON);
}

// albedo_impl (CompositeBSDF::prepare) and filter_o (layer opacity)
OSLMI_HD V3 mx_albedo(const MxLobe& M) { return M.tint * (1 - M.E); }
OSLMI_HD float mx_filter_o(const MxLobe& M) { return M.E; }

}  // namespace mtx
}  // namespace oslmi
