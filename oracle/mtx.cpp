// CPU oracle: libbsdl MaterialX lobes and testrender's layered closure.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
//   Imath Vec3 length / normalized (lengthTiny for |v|^2 < 2 FLT_MIN)
//   GGXDist D, G1, G2_G1, sample, sample_for_refl, D_refl_D
//                                     libbsdl/include/BSDL/microfacet_tools_impl.h:19-118
//   TabulatedEnergyCurve               microfacet_tools_impl.h:120-218
//   eval/sample_turquin_microms_reflection   microfacet_tools_impl.h:345-394
//   square_to_unit_disc, sample_uniform_hemisphere, fast_{cos,sin}_quadrant,
//   reflect, Frame                     libbsdl/include/BSDL/tools.h:209-283, :307-318, :468-541
//   Sample::update                     libbsdl/include/BSDL/bsdf_impl.h:16-38
//   BsdfGlobals::visible_normal / regularize_roughness   bsdf_impl.h:48-83, :155-160
//   DielectricFresnel, DielectricBSDF, DielectricLobe
//                                     MTX/bsdf_dielectric_impl.h:21-310
//   SchlickFresnel, SchlickLobe        MTX/bsdf_schlick_impl.h:14-141
//   SheenLobe, ContyKullaDist (non-legacy shadowing), SheenMicrofacet,
//   ContyKullaSheenMTX::albedo         MTX/bsdf_sheen_impl.h:19-210
//   BSDL_WRAP world<->local            src/testrender/shading.cpp:73-106
//   evaluate_layer_opacity, MX_LAYER   src/testrender/shading.cpp:1199-1272, :1645-1662
//   CompositeBSDF prepare/eval/sample  src/testrender/shading.h:319-393
//   MIS::update_eval                   src/testrender/sampling.h:173-206
//
// Run-time fast math (testrender/bsdl_config.h:17-73): OIIO fast_cos/
// fast_sin are replaced by ofast_sincosf and fast_safe_pow by
// ofast_safe_powf (omath.h). The table bake uses libm cosf/sinf like
// genluts.cpp's BSDLDefaultConfig (libbsdl/include/BSDL/config.h:52-90).
#include "mtx.h"

#include <dlfcn.h>

#include <atomic>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "omath.h"

namespace oref {
namespace mtx {

namespace {

const float kPI        = float(M_PI);
const float kONEOVERPI = 1.0f / kPI;
const float kEPSILON   = 1e-4f;
const float kIOR_MIN   = 1.001f;
const float kIOR_MAX   = 5.0f;

inline float SQR(float x) { return x * x; }
// std::max / std::min, and BSDL's MIN template (tools.h:92-97); BSDL's
// MAX on floats is std::max (tools.h:16-20). They differ on ties of
// signed zeros and on NaN.
inline float smax(float a, float b) { return a < b ? b : a; }
inline float smin(float a, float b) { return b < a ? b : a; }
inline float bmin(float a, float b) { return a < b ? a : b; }
inline float CLAMP(float x, float a, float b) { return bmin(smax(x, a), b); }
inline float LERP(float f, float a, float b)
{
    f = CLAMP(f, 0.0f, 1.0f);
    return (1 - f) * a + f * b;
}
inline V3 LERP3(float f, V3 a, V3 b)
{
    f = CLAMP(f, 0.0f, 1.0f);
    return a * (1 - f) + b * f;
}
inline float MAX_RGB(V3 c) { return smax(c.x, smax(c.y, c.z)); }
inline float MAX_ABS_XYZ(V3 v) { return smax(fabsf(v.x), smax(fabsf(v.y), fabsf(v.z))); }
inline float pmax(V3 c)  // Power::max
{
    float m = c.x;
    m       = smax(m, c.y);
    m       = smax(m, c.z);
    return m;
}

float vlength(V3 v)
{
    float l2 = dot(v, v);
    if (l2 < 2.0f * FLT_MIN) {
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
V3 normalized(V3 v)
{
    float l = vlength(v);
    if (l == 0.0f)
        return V3(0.0f);
    return V3(v.x / l, v.y / l, v.z / l);
}
V3 reflect(V3 E, V3 N) { return N * (2 * dot(N, E)) - E; }

struct Frame {
    V3 X, Y, Z;
    V3 local(V3 a) const { return V3(dot(a, X), dot(a, Y), dot(a, Z)); }
    V3 world(V3 a) const
    {
        return V3(X.x * a.x + Y.x * a.y + Z.x * a.z, X.y * a.x + Y.y * a.y + Z.y * a.z,
                  X.z * a.x + Y.z * a.y + Z.z * a.z);
    }
};
void ortho_build(V3 Z, V3* X, V3* Y)
{
    const float s = copysignf(1.0f, Z.z);
    const float a = -1.0f / (s + Z.z);
    const float b = Z.x * Z.y * a;
    *X            = V3(1.0f + s * Z.x * Z.x * a, s * b, -s * Z.x);
    *Y            = V3(b, s + Z.y * Z.y * a, -Z.y);
}
Frame frame_zx(V3 Z, V3 X)
{
    Frame f;
    f.Z = Z;
    f.X = X;
    if (MAX_ABS_XYZ(X) < 1e-4f || fabsf(dot(Z, normalized(X))) > 0.999f) {
        ortho_build(Z, &f.X, &f.Y);
    } else {
        f.Y = normalized(cross(Z, X));
        f.X = cross(f.Y, Z);
    }
    return f;
}

V3 visible_normal(V3 wo, V3 Ngf, V3 N)
{
    if (dot(wo, N) > 0.0f)
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

// path_roughness = 0 for the camera-ray shading point
float regularize_roughness(float r)
{
    const float roughness_product = 1.0f - 0.0f;
    return 1.0f - (1.0f - r) * roughness_product;
}

float fast_cos_quadrant(float x)
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
float fast_sin_quadrant(float x)
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

// ---- GGX ------------------------------------------------------------
struct GGX {
    float ax, ay;
};
GGX ggx_make(float rough, float aniso, bool flip)
{
    GGX d;
    d.ax = SQR(rough);
    d.ay = d.ax;
    if (flip)
        aniso = -aniso;
    const float ALPHA_MIN = 1e-5f;
    d.ax                  = smax(d.ax * (1 + aniso), ALPHA_MIN);
    d.ay                  = smax(d.ay * (1 - aniso), ALPHA_MIN);
    return d;
}

// ---- energy tables ----------------------------------------------------
float get_cosine(int table, int i)
{
    const float x = float(i) * (1.0f / 15.0f);
    return table == kLutMiniGGX ? smax(SQR(x), 1e-6f) : smax(x, 1e-6f);
}
// Row indices are clamped into the table: the reference asserts
// 0 <= roughness, fresnel_index <= 1, where the clamps change nothing.
inline int clamp_index(int k, int hi) { return k < 0 ? 0 : (k > hi ? hi : k); }
float interpolate_emiss(const float* E, int Nf, float roughness, float fresnel_index, int i)
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
    return LERP(ff,
                LERP(rf, E[(fa * kNr + ra) * kNc + i], E[(fa * kNr + rb) * kNc + i]),
                LERP(rf, E[(fb * kNr + ra) * kNc + i], E[(fb * kNr + rb) * kNc + i]));
}
float emiss_eval_t(const float* tables, int table, float roughness, float fresnel_index, float c)
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
float dfresnel_eta(float eta, bool backside)
{
    if (backside)
        eta = 1 / eta;
    return eta >= 1 ? CLAMP(eta, kIOR_MIN, kIOR_MAX)
                    : CLAMP(eta, 1 / kIOR_MAX, 1 / kIOR_MIN);
}
float dfresnel_table_index(float eta)
{
    const float feta = eta;
    const float seta = CLAMP(feta < 1 ? 1 / feta : feta, kIOR_MIN, kIOR_MAX);
    const float x    = (seta - kIOR_MIN) * (1 / (kIOR_MAX - kIOR_MIN));
    return sqrtf(x);
}

// DielectricFresnel (kind 0) or SchlickFresnel (kind 1)
struct Fresnel {
    int kind;
    float eta;
    V3 F0, F90;
    float exponent, tir_cos;
    V3 eval(float c) const
    {
        if (kind == 0)
            return V3(dielectric_fresnel(eta, c));
        c = CLAMP(c, 0.0f, 1.0f);
        if (c < tir_cos)
            return V3(1.0f);
        return LERP3(ofast_safe_powf(1 - c, exponent), F0, F90);
    }
};

// ---- microfacet reflection with Turquin multiple scattering ------------
struct Spec {
    GGX d;
    Fresnel f;
    float E_ms;
};
Sample eval_turquin(const Spec& sp, V3 wo, V3 wi)
{
    const float cosNO = wo.z;
    const float cosNI = wi.z;
    if (cosNI <= 0 || cosNO <= 0)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    V3 m            = normalized(wo + wi);
    float cosMO     = dot(m, wo);
    const float D   = ggx_D(sp.d.ax, sp.d.ay, m);
    float D_refl_D  = ggx_D_refl_D(sp.d.ax, sp.d.ay, wo, m);
    float D_refl    = D_refl_D * D;
    const float G1  = ggx_G1(sp.d.ax, sp.d.ay, wo);
    float pdf       = D_refl / (4 * cosNO);
    const float out = ggx_G2_G1(sp.d.ax, sp.d.ay, wi, wo) * G1 / D_refl_D;
    const V3 F      = sp.f.eval(cosMO);
    const V3 F_ms   = F;
    const float msf = sp.E_ms / smax(0.01f, 1 - sp.E_ms);
    const V3 O      = (F * out) * (V3(1.0f) + F_ms * msf);
    return Sample { wi, O, pdf, -1.0f };
}
Sample sample_turquin(const Spec& sp, V3 wo, float rx, float ry, bool libm_trig)
{
    const float cosNO = wo.z;
    if (cosNO <= 0)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    V3 m  = ggx_sample_bvndf(sp.d.ax, sp.d.ay, wo, rx, ry, libm_trig);
    V3 wi = reflect(wo, m);
    if (wi.z <= 0)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    return eval_turquin(sp, wo, wi);
}

// ---- MTX lobes ------------------------------------------------------------
// DielectricLobe / SchlickLobe (reflection only: refr_tint == 0)
struct MicroLobe {
    Frame frame;
    Spec spec;
    V3 refl_tint, refr_tint;
    float E_ms;
    bool dorefl, dorefr;
    float roughness;
};
struct SheenLobeS {
    Frame frame;
    V3 tint;
    float a;  // ContyKullaDist alpha
    float Emiss;
    bool backfacing;
    float roughness;
};

float microlobe_roughness(float roughness_x, float roughness_y, float* aniso, bool* flip)
{
    const float rx = CLAMP(roughness_x, kEPSILON, 2.0f);
    const float ry = CLAMP(roughness_y, kEPSILON, 2.0f);
    const float ax = smax(rx, ry);
    const float ay = smin(rx, ry);
    const float b  = ay / ax;
    *aniso         = (1 - b) / (1 + b);
    *flip          = rx < ry;
    return regularize_roughness(sqrtf(ax / (1 + *aniso)));
}

// DielectricLobe (kind 0) / SchlickLobe (kind 1) constructors at wo
// (MTX/bsdf_dielectric_impl.h:210-262, MTX/bsdf_schlick_impl.h:14-60):
// frame (visible normal, tangent U), anisotropic GGX, Fresnel, energies
void micro_build(MicroLobe* L, int kind, V3 N, V3 U, V3 refl_tint, V3 refr_tint, float roughness_x,
                 float roughness_y, float ior, V3 F0, V3 F90, float exponent, V3 wo)
{
    L->frame     = frame_zx(visible_normal(wo, N, N), U);
    L->refl_tint = refl_tint;
    L->refr_tint = refr_tint;
    L->dorefl    = MAX_RGB(refl_tint) > 0;
    L->dorefr    = MAX_RGB(refr_tint) > 0;
    float aniso;
    bool flip;
    const float roughness = microlobe_roughness(roughness_x, roughness_y, &aniso, &flip);
    const float cosNO     = dot(L->frame.Z, wo);
    Fresnel f;
    if (kind == 0) {
        const float IOR = CLAMP(ior, kIOR_MIN, kIOR_MAX);
        f.kind          = 0;
        f.eta           = dfresnel_eta(IOR / 1.0f, false);
        f.exponent      = 0;
        f.tir_cos       = 0;
    } else {
        const float avg_F0         = CLAMP((F0.x + F0.y + F0.z) * (1.0f / 3), 0.0f, 0.99f);
        const float sqrt_F0        = sqrtf(avg_F0);
        const float refraction_ior = (1 + sqrt_F0) / (1 - sqrt_F0);
        f.kind     = 1;
        f.eta      = dfresnel_eta(refraction_ior, false);
        f.F0       = V3(smax(0.0f, smin(1.0f, F0.x)), smax(0.0f, smin(1.0f, F0.y)),
                        smax(0.0f, smin(1.0f, F0.z)));
        f.F90      = V3(smax(0.0f, smin(1.0f, F90.x)), smax(0.0f, smin(1.0f, F90.y)),
                        smax(0.0f, smin(1.0f, F90.z)));
        f.exponent = exponent;
        const float zero = 0.0f;
        f.tir_cos  = f.eta >= 1 ? zero : sqrtf(1 - SQR(f.eta));
    }
    L->E_ms      = 0;
    L->spec.d    = ggx_make(roughness, aniso, flip);
    L->spec.f    = f;
    L->spec.E_ms = 0;
    if (!L->dorefr)
        L->spec.E_ms = emiss_eval_t(luts(), kLutMiniGGX, roughness, 0.0f, cosNO);
    if (L->dorefl && !L->dorefr)
        L->E_ms = emiss_eval_t(luts(), kLutReflFront, roughness, dfresnel_table_index(f.eta), cosNO);
    L->roughness = roughness;
}

void dielectric_lobe(MicroLobe* L, const LayeredParams& p, V3 wo)
{
    micro_build(L, 0, p.N, V3(0.0f), p.coat_refl_tint, p.coat_refr_tint, p.coat_roughness,
                p.coat_roughness, p.coat_ior, V3(0.0f), V3(0.0f), 0.0f, wo);
}

void schlick_lobe(MicroLobe* L, const LayeredParams& p, V3 wo)
{
    micro_build(L, 1, p.N, V3(0.0f), p.base_refl_tint, p.base_refr_tint, p.base_roughness,
                p.base_roughness, 0.0f, p.base_f0, p.base_f90, p.base_exponent, wo);
}

Sample microlobe_eval_local(const MicroLobe& L, V3 wo, V3 wi)
{
    if (!L.dorefl && !L.dorefr)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    Sample s  = eval_turquin(L.spec, wo, wi);
    s.weight  = s.weight * (s.wi.z > 0 ? L.refl_tint : L.refr_tint);
    s.roughness = L.roughness;
    return s;
}
Sample microlobe_sample_local(const MicroLobe& L, V3 wo, float rx, float ry)
{
    if (!L.dorefl && !L.dorefr)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    Sample s = sample_turquin(L.spec, wo, rx, ry, false);
    if (MAX_ABS_XYZ(s.wi) < kEPSILON)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    s.weight    = s.weight * (s.wi.z > 0 ? L.refl_tint : L.refr_tint);
    s.roughness = L.roughness;
    return s;
}

float ck_G2(V3 wo, V3 wi)
{
    float cosNI = smin(1.0f, wi.z);
    float cosNO = smin(1.0f, wo.z);
    return (cosNI * cosNO) / (cosNI + cosNO - cosNI * cosNO);
}
Sample sheen_mf_eval(float a, V3 wo, V3 wi)
{
    const float cosNO = wo.z;
    const float cosNI = wi.z;
    if (cosNI <= 1e-5f || cosNO <= 1e-5f)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    const float D = conty_kulla_D(a, normalized(wo + wi));
    if ((double)D < 1e-6)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    const float G2 = ck_G2(wo, wi);
    return Sample { wi, V3(D * G2 * 0.5f * kPI / cosNO), 0.5f * kONEOVERPI, 0.0f };
}

// SheenLobe constructor, Conty-Kulla mode (MTX/bsdf_sheen_impl.h:19-48)
void sheen_build(SheenLobeS* L, V3 N, V3 albedo, float roughness, V3 wo)
{
    L->frame      = frame_zx(visible_normal(wo, N, N), wo);
    L->roughness  = regularize_roughness(CLAMP(roughness, 0.0f, 1.0f));
    L->tint       = albedo;
    L->Emiss      = 1.0f;
    L->backfacing = false;
    const float sheen_alpha = smax(0.06f, L->roughness);
    L->a                    = CLAMP(sheen_alpha, 0.06f, 1.0f);
    const float cosNO       = CLAMP(dot(L->frame.Z, wo), 0.0f, 1.0f);
    if (L->backfacing)
        L->Emiss = 1;
    else
        L->Emiss = 1 - smin(conty_kulla_albedo(L->a, cosNO) * pmax(L->tint), 1.0f);
}
void sheen_lobe(SheenLobeS* L, const LayeredParams& p, V3 wo)
{
    sheen_build(L, p.N, p.sheen_albedo, p.sheen_roughness, wo);
}
Sample sheen_eval_local(const SheenLobeS& L, V3 wo, V3 wi)
{
    const float cosNO        = wo.z;
    const float cosNI        = wi.z;
    const bool is_reflection = cosNI > 0 && cosNO >= 0;
    const bool do_self       = is_reflection && !L.backfacing;
    if (!do_self)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    Sample s    = sheen_mf_eval(L.a, wo, wi);
    s.weight    = s.weight * L.tint;
    s.roughness = L.roughness;
    return s;
}
Sample sheen_sample_local(const SheenLobeS& L, V3 wo, float rx, float ry)
{
    if (L.backfacing)
        return Sample { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    V3 wi       = sample_uniform_hemisphere(rx, ry);
    Sample s    = sheen_mf_eval(L.a, wo, wi);
    s.weight    = s.weight * L.tint;
    s.roughness = L.roughness;
    return s;
}

void mis_update_eval(V3* w, float* pdf, V3 ow, float opdf, float b)
{
    if (b > FLT_MIN) {
        opdf *= b;
        ow = ow * (1 / b);
        float mis;
        if (*pdf < opdf)
            mis = 1 / (1 + *pdf / opdf);
        else if (opdf < *pdf)
            mis = 1 - 1 / (1 + opdf / *pdf);
        else
            mis = 0.5f;
        *w = *w * (1 - mis) + ow * mis;
        *pdf += opdf;
    }
}

// OIIO clamp on one channel
inline float oclamp(float a, float lo, float hi)
{
    float v = a;
    if (!(lo <= v))
        v = lo;
    if (!(v <= hi))
        v = hi;
    return v;
}
inline V3 oclamp3(V3 c) { return V3(oclamp(c.x, 0.f, 1.f), oclamp(c.y, 0.f, 1.f), oclamp(c.z, 0.f, 1.f)); }
inline bool is_black(V3 c) { return c.x == 0 && c.y == 0 && c.z == 0; }

}  // namespace

// ---- exposed building blocks ---------------------------------------------
float ggx_D(float ax, float ay, V3 Hr)
{
    const float cosPhi2st2 = SQR(Hr.x / ax);
    const float sinPhi2st2 = SQR(Hr.y / ay);
    const float cosThetaM2 = SQR(Hr.z);
    const float sinThetaM2 = cosPhi2st2 + sinPhi2st2;
    return 1.0f / (kPI * ax * ay * SQR(cosThetaM2 + sinThetaM2));
}
float ggx_G1(float ax, float ay, V3 w)
{
    w = V3(w.x * ax, w.y * ay, w.z);
    return 2.0f * w.z / (w.z + vlength(w));
}
float ggx_G2_G1(float ax, float ay, V3 wi, V3 wo)
{
    wi             = V3(wi.x * ax, wi.y * ay, wi.z);
    wo             = V3(wo.x * ax, wo.y * ay, wo.z);
    const float nl = vlength(wi);
    const float nv = vlength(wo);
    return wi.z * (wo.z + nv) / (wo.z * nl + wi.z * nv);
}
float ggx_D_refl_D(float ax, float ay, V3 wo, V3 m)
{
    (void)m;
    const float aix  = wo.x * ax, aiy = wo.y * ay;
    const float len2 = aix * aix + aiy * aiy;
    const float t    = sqrtf(len2 + SQR(wo.z));
    const float a    = CLAMP(smin(ax, ay), 0.0f, 1.0f);
    const float s    = 1 + sqrtf(SQR(wo.x) + SQR(wo.y));
    const float a2   = SQR(a);
    const float s2   = SQR(s);
    const float k    = (1 - a2) * s2 / (s2 + a2 * SQR(wo.z));
    return 2 * wo.z / (k * wo.z + t);
}
V3 ggx_sample_vndf(float ax, float ay, V3 wo, float randu, float randv)
{
    const V3 V  = normalized(V3(ax * wo.x, ay * wo.y, wo.z));
    const V3 T1 = V.z < 0.9999f ? normalized(V3(V.y, -V.x, 0)) : V3(1, 0, 0);
    const V3 T2 = cross(T1, V);
    V2 p        = square_to_unit_disc(randu, randv);
    const float s   = 0.5f * (1 + V.z);
    const float p2o = s * p.y + (1 - s) * sqrtf(1 - p.x * p.x);
    const float p3  = sqrtf(smax(1.0f - SQR(p.x) - SQR(p2o), 0.0f));
    const V3 N      = p.x * T1 + p2o * T2 + p3 * V;
    return normalized(V3(ax * N.x, ay * N.y, smax(N.z, 0.0f)));
}
V3 ggx_sample_bvndf(float ax, float ay, V3 wo, float randu, float randv, bool libm_trig)
{
    V3 i_std             = normalized(V3(wo.x * ax, wo.y * ay, wo.z));
    const float phi      = 2.0f * kPI * randu;
    const float a        = CLAMP(smin(ax, ay), 0.0f, 1.0f);
    const float s        = 1 + sqrtf(SQR(wo.x) + SQR(wo.y));
    const float a2       = SQR(a);
    const float s2       = SQR(s);
    const float k        = (1 - a2) * s2 / (s2 + a2 * SQR(wo.z));
    const float b        = k * i_std.z;
    const float z        = (1 - randv) * (1 + b) - b;
    const float sinTheta = sqrtf(CLAMP(1 - SQR(z), 0.0f, 1.0f));
    float c, sn;
    if (libm_trig) {
        c  = cosf(phi);
        sn = sinf(phi);
    } else {
        ofast_sincosf(phi, &sn, &c);
    }
    V3 o_std = V3(sinTheta * c, sinTheta * sn, z);
    V3 m_std = i_std + o_std;
    return normalized(V3(m_std.x * ax, m_std.y * ay, m_std.z));
}
V2 square_to_unit_disc(float u, float v)
{
    const float a = 2 * u - 1, qa = fabsf(a);
    const float b = 2 * v - 1, qb = fabsf(b);
    const float rad = qa > qb ? qa : qb;
    const float one = 1.0f;
    const float phi = qa > qb ? qb / qa : ((qa == qb) ? one : 2 - qa / qb);
    const float x   = copysignf(rad * fast_cos_quadrant(phi), a);
    const float y   = copysignf(rad * fast_sin_quadrant(phi), b);
    return V2 { x, y };
}
V3 sample_uniform_hemisphere(float u, float v)
{
    const float a = 2 * u - 1, qa = fabsf(a);
    const float b = 2 * v - 1, qb = fabsf(b);
    const float rad = qa > qb ? qa : qb;
    const float one = 1.0f;
    const float phi = qa > qb ? qb / qa : ((qa == qb) ? one : 2 - qa / qb);
    const float x   = copysignf(rad * fast_cos_quadrant(phi), a);
    const float y   = copysignf(rad * fast_sin_quadrant(phi), b);
    const float cos_theta = 1 - rad * rad;
    const float sin_theta = sqrtf(2 - rad * rad);
    return V3(sin_theta * x, sin_theta * y, cos_theta);
}
float dielectric_fresnel(float eta, float c)
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
float emiss_eval(int table, float roughness, float fresnel_index, float c)
{
    return emiss_eval_t(luts(), table, roughness, fresnel_index, c);
}
float conty_kulla_albedo(float a, float cosNO)
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
float conty_kulla_D(float a, V3 Hr)
{
    float cos_theta = CLAMP(Hr.z, 0.0f, 1.0f);
    float sin_theta = sqrtf(1.0f - SQR(cos_theta));
    return ofast_safe_powf(sin_theta, 1 / a) * (2 + 1 / a) * 0.5f * kONEOVERPI;
}
void sample_update(V3* w, float* pdf, V3 ow, float opdf, float cpdf)
{
    const float PDF_MIN = 1e-6f;
    if (cpdf > PDF_MIN) {
        opdf *= cpdf;
        ow = ow * (1 / cpdf);
        if (opdf == *pdf)
            *w = LERP3(0.5f, *w, ow);
        else if (opdf < *pdf)
            *w = LERP3(1 / (1 + opdf / *pdf), ow, *w);
        else
            *w = LERP3(1 / (1 + *pdf / opdf), *w, ow);
        *pdf += opdf;
    }
}

// ---- the layered material ------------------------------------------------
struct LayeredBsdf {
    int n;             // lobes present (1..3), in closure order
    int kind[3];       // 0 sheen, 1 dielectric, 2 schlick
    V3 weight[3];
    float pdf[3];
    SheenLobeS sheen;
    MicroLobe coat, base;
};

// ---- one MaterialX lobe of a closure list ---------------------------------
namespace {
struct MxL {
    int kind;  // 0 dielectric, 1 generalized Schlick, 2 sheen
    MicroLobe m;
    SheenLobeS s;
};
static_assert(sizeof(MxL) <= sizeof(MxLobe::state), "MxLobe storage");
const MxL* mxl(const MxLobe& L) { return reinterpret_cast<const MxL*>(L.state); }
}  // namespace

void mx_build(MxLobe* L, const MxParams& p, V3 wo)
{
    MxL* x = reinterpret_cast<MxL*>(L->state);
    *x     = MxL();
    x->kind = p.kind;
    if (p.kind == 2)
        sheen_build(&x->s, p.N, p.tint, p.rx, wo);
    else
        micro_build(&x->m, p.kind, p.N, p.U, p.tint, V3(0.0f), p.rx, p.ry, p.ior, p.F0, p.F90,
                    p.exponent, wo);
}
// BSDL_WRAP::eval / sample (src/testrender/shading.cpp:88-101)
Sample mx_eval(const MxLobe& L, V3 wo, V3 wi)
{
    const MxL* x = mxl(L);
    Sample s;
    if (x->kind == 2)
        s = sheen_eval_local(x->s, x->s.frame.local(wo), x->s.frame.local(wi));
    else
        s = microlobe_eval_local(x->m, x->m.frame.local(wo), x->m.frame.local(wi));
    return Sample { wi, s.weight, s.pdf, s.roughness };
}
Sample mx_sample(const MxLobe& L, V3 wo, float rx, float ry)
{
    const MxL* x   = mxl(L);
    const Frame& f = x->kind == 2 ? x->s.frame : x->m.frame;
    Sample s       = x->kind == 2 ? sheen_sample_local(x->s, f.local(wo), rx, ry)
                                  : microlobe_sample_local(x->m, f.local(wo), rx, ry);
    return Sample { f.world(s.wi), s.weight, s.pdf, s.roughness };
}
// albedo_impl, filter_o (MTX/bsdf_*_decl.h)
V3 mx_albedo(const MxLobe& L)
{
    const MxL* x = mxl(L);
    return x->kind == 2 ? x->s.tint * (1 - x->s.Emiss) : x->m.refl_tint * (1 - x->m.E_ms);
}
float mx_filter_o(const MxLobe& L)
{
    const MxL* x = mxl(L);
    return x->kind == 2 ? x->s.Emiss : x->m.E_ms;
}

LayeredBsdf* layered_new() { return new LayeredBsdf(); }
void layered_free(LayeredBsdf* b) { delete b; }

namespace {
Sample lobe_eval_world(const LayeredBsdf* B, int k, V3 wo, V3 wi)
{
    Sample s;
    if (k == 0)
        s = sheen_eval_local(B->sheen, B->sheen.frame.local(wo), B->sheen.frame.local(wi));
    else {
        const MicroLobe& L = k == 1 ? B->coat : B->base;
        s                  = microlobe_eval_local(L, L.frame.local(wo), L.frame.local(wi));
    }
    return Sample { wi, s.weight, s.pdf, s.roughness };
}
Sample lobe_sample_world(const LayeredBsdf* B, int k, V3 wo, float rx, float ry)
{
    Sample s;
    const Frame* f;
    if (k == 0) {
        f = &B->sheen.frame;
        s = sheen_sample_local(B->sheen, f->local(wo), rx, ry);
    } else {
        const MicroLobe& L = k == 1 ? B->coat : B->base;
        f                  = &L.frame;
        s                  = microlobe_sample_local(L, f->local(wo), rx, ry);
    }
    return Sample { f->world(s.wi), s.weight, s.pdf, s.roughness };
}
V3 lobe_albedo(const LayeredBsdf* B, int k)
{
    if (k == 0)
        return B->sheen.tint * (1 - B->sheen.Emiss);
    const MicroLobe& L = k == 1 ? B->coat : B->base;
    return L.refl_tint * (1 - L.E_ms);
}
}  // namespace

void layered_build(LayeredBsdf* B, const LayeredParams& p, V3 wo)
{
    sheen_lobe(&B->sheen, p, wo);
    dielectric_lobe(&B->coat, p, wo);
    schlick_lobe(&B->base, p, wo);
    // process_bsdf_closure on layer(sheen, layer(coat, base)), all closure
    // component weights 1; opacity of a top layer = w * (1 - filter_o)
    const V3 one(1.0f);
    B->n           = 0;
    V3 weight      = one;
    // outer layer
    V3 op_sheen    = one * (one * (one - V3(B->sheen.Emiss)));
    V3 base_w      = weight * (one - oclamp3(op_sheen));
    B->kind[B->n]  = 0;
    B->weight[B->n++] = weight * one;
    if (!is_black(base_w)) {
        weight        = base_w;
        V3 cw         = weight * one;
        V3 op_coat    = one * (one * (one - V3(B->coat.E_ms)));
        V3 base_w2    = weight * (one - oclamp3(op_coat));
        B->kind[B->n] = 1;
        B->weight[B->n++] = cw * one;
        if (!is_black(base_w2)) {
            B->kind[B->n]     = 2;
            B->weight[B->n++] = base_w2 * one;
        }
    }
    // CompositeBSDF::prepare(wo, path_weight = 1, absorb = false)
    const V3 pw(1.0f);
    float total = 0;
    for (int i = 0; i < B->n; i++) {
        B->pdf[i] = dot(B->weight[i], pw * lobe_albedo(B, B->kind[i])) / (pw.x + pw.y + pw.z);
        total += B->pdf[i];
    }
    if (total > 0 || total > 1)
        for (int i = 0; i < B->n; i++)
            B->pdf[i] /= total;
}

Sample layered_eval(const LayeredBsdf* B, V3 wo, V3 wi)
{
    Sample s { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    for (int i = 0; i < B->n; i++) {
        Sample b = lobe_eval_world(B, B->kind[i], wo, wi);
        b.weight = b.weight * B->weight[i];
        mis_update_eval(&s.weight, &s.pdf, b.weight, b.pdf, B->pdf[i]);
        s.roughness += b.roughness * B->pdf[i];
    }
    return s;
}

Sample layered_sample(const LayeredBsdf* B, V3 wo, float rx, float ry, float rz)
{
    (void)rz;  // reflection-only lobes consume two random numbers
    const Sample zero { V3(0.0f), V3(0.0f), 0.0f, 0.0f };
    float accum = 0;
    for (int i = 0; i < B->n; i++) {
        if (rx < (B->pdf[i] + accum)) {
            rx       = (rx - accum) / B->pdf[i];
            rx       = smin(rx, 0.99999994f);
            Sample s = lobe_sample_world(B, B->kind[i], wo, rx, ry);
            s.weight = s.weight * (B->weight[i] * (1 / B->pdf[i]));
            s.pdf *= B->pdf[i];
            if (s.pdf == 0.0f)
                return zero;
            for (int j = 0; j < B->n; j++) {
                if (i != j) {
                    Sample b = lobe_eval_world(B, B->kind[j], wo, s.wi);
                    b.weight = b.weight * B->weight[j];
                    mis_update_eval(&s.weight, &s.pdf, b.weight, b.pdf, B->pdf[j]);
                }
            }
            return s;
        }
        accum += B->pdf[i];
    }
    return zero;
}

void layered_weights(const LayeredBsdf* B, float* w9, float* pdf3)
{
    for (int k = 0; k < 3; k++) {
        w9[3 * k] = w9[3 * k + 1] = w9[3 * k + 2] = 0;
        pdf3[k]                                   = 0;
    }
    for (int i = 0; i < B->n; i++) {
        int k         = B->kind[i];
        w9[3 * k]     = B->weight[i].x;
        w9[3 * k + 1] = B->weight[i].y;
        w9[3 * k + 2] = B->weight[i].z;
        pdf3[k]       = B->pdf[i];
    }
}

// ---- energy tables: bake and cache ---------------------------------------
// Restates libbsdl/src/genluts.cpp:50-192. Every entry is 1 - E, E the
// running mean of bsdf.sample(wo, rnd).weight.max() over a 128 x 128
// scrambled Larcher-Pillichshammer point set with wo at the entry's cosine
// and scrambles from fasthash64 of the (fresnel, roughness) entry indices.
// MiniMicrofacetGGX (microfacet_tools_impl.h:220-237) is baked first; the
// DielectricReflFront entries sample the Turquin reflection, whose
// multiple-scattering term reads it. libm cosf/sinf, as genluts.cpp uses
// through BSDLDefaultConfig. Cached in <dir of liboslref.so>/bsdl_luts.bin.
namespace {

std::once_flag g_lut_once;
std::vector<float> g_lut;
const char kLutMagic[8] = { 'B', 'S', 'D', 'L', 'L', 'U', 'T', '1' };

std::string lut_cache_path()
{
    Dl_info info;
    if (dladdr((void*)&layered_new, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.find_last_of('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/bsdl_luts.bin";
    }
    return "bsdl_luts.bin";
}
bool lut_load(const std::string& path, float* out)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f)
        return false;
    char magic[8];
#ifdef OREF_F64
    // the tables are binary32 data; the FP64 build widens them on load
    std::vector<oref_f32> t(kLutFloats);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kLutMagic, 8) == 0
              && fread(t.data(), 4, kLutFloats, f) == (size_t)kLutFloats;
    for (int i = 0; ok && i < kLutFloats; i++)
        out[i] = t[i];
#else
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kLutMagic, 8) == 0
              && fread(out, 4, kLutFloats, f) == (size_t)kLutFloats;
#endif
    fclose(f);
    return ok;
}

#if !defined(OREF_COUNTING) && !defined(OREF_F64)
uint32_t ri_LP(uint32_t i)
{
    uint32_t r = 0;
    for (uint32_t v = 1U << 31; i; i >>= 1, v |= v >> 1)
        if (i & 1)
            r ^= v;
    return r;
}
uint32_t ri_LP_inv(uint32_t i)
{
    uint32_t r = 0;
    for (uint32_t v = 3U << 30; i; i >>= 1, v >>= 1)
        if (i & 1)
            r ^= v;
    return r;
}
V3 lp_point(int si, int AA, uint32_t sx, uint32_t sy, uint32_t sz)
{
    const uint32_t ex      = si % AA;
    const uint32_t ey      = si / AA;
    const uint32_t upper   = (ex ^ (sx >> 16)) << 16;
    const uint32_t lpUpper = ri_LP(upper) ^ sy;
    const uint32_t delta   = (ey << 16) ^ (lpUpper & 0xFFFF0000u);
    const uint32_t lower   = ri_LP_inv(delta);
    const uint32_t index   = upper | lower;
    const uint32_t x       = index ^ sx;
    const uint32_t y       = lpUpper ^ delta;
    const float jx         = (x & 65535) * (1 / 65536.0f);
    const float jy         = (y & 65535) * (1 / 65536.0f);
    uint32_t rz = sz, ii = index;
    for (uint64_t v2 = uint64_t(3) << 62; ii; ii >>= 1, v2 ^= v2 >> 1)
        if (ii & 1)
            rz ^= uint32_t(v2 >> 31);
    return V3((ex + jx) / AA, (ey + jy) / AA, rz * 2.3283063e-10f);
}
uint64_t fasthash_mix(uint64_t h)
{
    h ^= h >> 23;
    h *= 0x2127599bf4325c37ULL;
    h ^= h >> 47;
    return h;
}
uint32_t randhash3(uint32_t x, uint32_t y, uint32_t z)
{
    const uint64_t m    = 0x880355f21e6d1965ULL;
    const uint64_t v[2] = { (uint64_t(x) << 32) + y, uint64_t(z) };
    uint64_t h          = (2 * sizeof(uint64_t)) * m;
    for (int k = 0; k < 2; k++) {
        h ^= fasthash_mix(v[k]);
        h *= m;
    }
    return (uint32_t)fasthash_mix(h);
}

// MiniMicrofacetGGX::sample(...).weight.max()
float mini_ggx_weight(const GGX& d, V3 wo, V3 rnd)
{
    const V3 m = ggx_sample_vndf(d.ax, d.ay, wo, rnd.x, rnd.y);
    if (dot(m, wo) > 0) {
        const V3 wi = reflect(wo, m);
        if (wi.z > 0)
            return ggx_G2_G1(d.ax, d.ay, wi, wo);
    }
    return 0.0f;
}
// DielectricReflFront::sample(...).weight.max(): Turquin reflection with
// libm trigonometry
float reflfront_weight(const Spec& sp, V3 wo, V3 rnd)
{
    if (wo.z <= 0)
        return 0.0f;
    const V3 m  = ggx_sample_bvndf(sp.d.ax, sp.d.ay, wo, rnd.x, rnd.y, true);
    const V3 wi = reflect(wo, m);
    if (wi.z <= 0)
        return 0.0f;
    return pmax(eval_turquin(sp, wo, wi).weight);
}

template<class W> float compute_E(float cos_theta, W weight, uint32_t f, uint32_t r)
{
    const int AA          = 128;
    const int NUM_SAMPLES = AA * AA;
    const V3 wo(sqrtf(1 - SQR(cos_theta)), 0, cos_theta);
    float E           = 0;
    const uint32_t sx = randhash3(f, r, 0);
    const uint32_t sy = randhash3(f, r, 1);
    const uint32_t sz = randhash3(f, r, 2);
    for (int i = 0; i < NUM_SAMPLES; i++) {
        const V3 rnd    = lp_point(i, AA, sx, sy, sz);
        const float out = weight(wo, rnd);
        E               = LERP(1.0f / (1.0f + i), E, out);
    }
    return smin(E, 1.0f);
}

template<class F> void parallel_for_n(int n, int nthreads, F fn)
{
    std::atomic<int> next { 0 };
    auto work = [&]() {
        for (int i = next++; i < n; i = next++)
            fn(i);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; t++)
        th.emplace_back(work);
    work();
    for (auto& t : th)
        t.join();
}
#endif  // OREF_COUNTING

}  // namespace

#if !defined(OREF_COUNTING) && !defined(OREF_F64)
void bake_luts(float* out, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
    float* mini = out + kLutMiniGGX;
    parallel_for_n(kNr, nthreads, [&](int r) {
        const float rough = float(r) * (1.0f / (kNr - 1));
        const GGX d       = ggx_make(rough, 0.0f, false);
        for (int c = 0; c < kNc; c++) {
            const float cth = get_cosine(kLutMiniGGX, c);
            mini[r * kNc + c] =
                1 - compute_E(cth, [&](V3 wo, V3 rnd) { return mini_ggx_weight(d, wo, rnd); },
                              0u, (uint32_t)r);
        }
    });
    float* refl = out + kLutReflFront;
    parallel_for_n(kNfRefl * kNr, nthreads, [&](int fr) {
        const int f               = fr / kNr;
        const int r               = fr % kNr;
        const float fresnel_index = float(f) * (1.0f / (kNfRefl - 1));
        const float rough         = float(r) * (1.0f / (kNr - 1));
        Spec sp;
        sp.d          = ggx_make(rough, 0.0f, false);
        sp.f.kind     = 0;
        sp.f.eta      = dfresnel_eta(LERP(SQR(fresnel_index), kIOR_MIN, kIOR_MAX), false);
        sp.f.exponent = 0;
        sp.f.tir_cos  = 0;
        for (int c = 0; c < kNc; c++) {
            const float cth = get_cosine(kLutReflFront, c);
            sp.E_ms         = emiss_eval_t(out, kLutMiniGGX, rough, 0.0f, cth);
            refl[(f * kNr + r) * kNc + c] =
                1 - compute_E(cth, [&](V3 wo, V3 rnd) { return reflfront_weight(sp, wo, rnd); },
                              (uint32_t)f, (uint32_t)r);
        }
    });
}

bool save_luts(const char* path, const float* tables)
{
    FILE* f = fopen(path, "wb");
    if (!f)
        return false;
    bool ok = fwrite(kLutMagic, 1, 8, f) == 8
              && fwrite(tables, 4, kLutFloats, f) == (size_t)kLutFloats;
    fclose(f);
    return ok;
}
#endif

const float* luts()
{
    std::call_once(g_lut_once, []() {
        g_lut.assign(kLutFloats, 0.0f);
        const std::string path = lut_cache_path();
        if (!lut_load(path, g_lut.data())) {
#if !defined(OREF_COUNTING) && !defined(OREF_F64)
            unsigned hc = std::thread::hardware_concurrency();
            bake_luts(g_lut.data(), hc ? (int)(hc < 16 ? hc : 16) : 4);
            save_luts(path.c_str(), g_lut.data());
#else
            fprintf(stderr, "oracle: %s missing (run make in oracle/)\n", path.c_str());
            abort();
#endif
        }
    });
    return g_lut.data();
}

}  // namespace mtx
}  // namespace oref
