// CPU oracle: closure lobes and their MIS combination.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
//   Diffuse<0>::eval                 src/testrender/shading.cpp:301-312
//   Microfacet<GGXDist,0> eval/albedo src/testrender/shading.cpp:572-615,
//                                     evalLambda/G/D :713-757, GGXDist :497-506
//   TangentFrame                     src/testrender/sampling.h:15-65
//   fresnel_dielectric               src/testrender/optics.h:13-31
//   CompositeBSDF prepare/eval       src/testrender/shading.h:323-372
//   MIS::update_eval                 src/testrender/sampling.h:173-206
//   CompositeBSDF::sample            src/testrender/shading.h:376-402
//   Diffuse<0>::sample               src/testrender/shading.cpp:313-320,
//     sample_cosine_hemisphere / to_unit_disk  sampling.h:70-100
//   Microfacet<GGXDist,0>::sample    src/testrender/shading.cpp:655-685,
//     sampleMicronormal :759-795, GGXDist::sampleSlope :483-507
// to_unit_disk calls OIIO::fast_sincos, which is not in the reference tree;
// the oracle's own sincos (omath.h) stands in (parity unpinned at ulp level).
#include "closure.h"

#include <limits>

#include <algorithm>
#include <cfloat>

#include "omath.h"

namespace oref {

namespace {

struct Frame {
    V3 u, v, w;
};

// from_normal_and_tangent with a zero tangent (microfacet() passes
// U = vector(0)) falls back to from_normal (Duff et al. ONB).
Frame frame_from_normal(V3 n)
{
    const float sign = copysignf(1.0f, n.z);
    const float a    = -1 / (sign + n.z);
    const float b    = n.x * n.y * a;
    Frame f;
    f.u = V3(1 + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.v = V3(b, sign + n.y * n.y * a, -n.y);
    f.w = n;
    return f;
}
Frame frame_from_normal_tangent(V3 n, V3 t)
{
    V3 x        = t - n * dot(n, t);
    float xlen2 = dot(x, x);
    if (xlen2 > 0) {
        x = x * (1.0f / sqrtf(xlen2));
        return Frame { x, cross(n, x), n };
    }
    return frame_from_normal(n);
}
V3 tolocal(const Frame& f, V3 a) { return V3(dot(a, f.u), dot(a, f.v), dot(a, f.w)); }

V3 normalize(V3 v)
{
    float l = sqrtf(dot(v, v));
    if (l != 0.0f)
        return V3(v.x / l, v.y / l, v.z / l);
    return v;
}

float SQR(float x) { return x * x; }

float ggx_lambda(V3 w, float xa, float ya)
{
    float cosTheta2  = SQR(w.z);
    float cosPhi2st2 = SQR(w.x * xa);
    float sinPhi2st2 = SQR(w.y * ya);
    float a2         = cosTheta2 / (cosPhi2st2 + sinPhi2st2);
    return 0.5f * (-1.0f + sqrtf(1.0f + 1.0f / a2));
}

float ggx_D(V3 Hr, float xa, float ya)
{
    float cosThetaM = Hr.z;
    if (cosThetaM > 0) {
        float cosPhi2st2 = SQR(Hr.x / xa);
        float sinPhi2st2 = SQR(Hr.y / ya);
        float cosThetaM2 = SQR(cosThetaM);
        float cosThetaM4 = SQR(cosThetaM2);
        float tanThetaM2 = (cosPhi2st2 + sinPhi2st2) / cosThetaM2;
        float F          = 1 / (float(M_PI) * (1 + tanThetaM2) * (1 + tanThetaM2));
        return F / (xa * ya * cosThetaM4);
    }
    return 0;
}

// BeckmannDist (src/testrender/shading.cpp:510-522): F = exp(-tan^2) / pi
// with the oracle's exp for OIIO::fast_exp, and Walter et al.'s rational
// Lambda
float beckmann_lambda(V3 w, float xa, float ya)
{
    const float a2 = SQR(w.z) / (SQR(w.x * xa) + SQR(w.y * ya));
    const float a  = sqrtf(a2);
    return a < 1.6f ? (1.0f - 1.259f * a + 0.396f * a2) / (3.535f * a + 2.181f * a2)
                    : float(0.0f);
}
float beckmann_D(V3 Hr, float xa, float ya)
{
    const float c = Hr.z;
    if (!(c > 0))
        return 0;
    const float c2   = SQR(c);
    const float tan2 = (SQR(Hr.x / xa) + SQR(Hr.y / ya)) / c2;
    return (float(1 / M_PI) * ofast_expf(-tan2)) / (xa * ya * SQR(c2));
}

BsdfSample lobe_eval(const Lobe& l, V3 wo, V3 wi)
{
    if (l.kind == kLobeMx) {
        const mtx::Sample s = mtx::mx_eval(l.mx, wo, wi);
        return { s.weight, s.pdf };
    }
    if (l.kind == kLobeDiffuse || l.kind == kLobeTranslucent) {
        const float pdf = std::max<float>(dot(l.N, wi), 0.0f) * float(M_1_PI);
        return { V3(1.0f), pdf };
    }
    if (l.kind >= kLobeReflection && l.kind <= kLobeEmission)
        return { V3(0.0f), 0.0f };
    if (l.kind == kLobePhong) {
        // Phong::eval (src/testrender/shading.cpp:323-340)
        const float cosNI = dot(l.N, wi);
        const float cosNO = dot(l.N, wo);
        if (cosNI > 0 && cosNO > 0) {
            const V3 R        = l.N * (2 * cosNO) - wo;
            const float cosRI = dot(R, wi);
            if (cosRI > 0) {
                const float e   = l.alpha;
                const float pdf = (e + 1) * float(M_1_PI / 2) * ofast_safe_powf(cosRI, e);
                return { V3(cosNI * (e + 2) / (e + 1)), pdf };
            }
        }
        return { V3(0.0f), 0.0f };
    }
    Frame tf        = frame_from_normal_tangent(l.N, l.aniso ? l.U : V3(0.0f));
    const V3 wo_l   = tolocal(tf, wo);
    const V3 wi_l   = tolocal(tf, wi);
    const float xa  = l.alpha, ya = l.aniso ? l.ay : l.alpha;
    const bool bk   = l.kind == kLobeBeckmannRefl;
    // Microfacet<Dist, Refract>::eval (src/testrender/shading.cpp:589-646)
    if ((l.refract == 0 || l.refract == 2) && wo_l.z > 0 && wi_l.z > 0) {
        const V3 m           = normalize(wi_l + wo_l);
        const float D        = bk ? beckmann_D(m, xa, ya) : ggx_D(m, xa, ya);
        const float Lambda_o = bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya);
        const float Lambda_i = bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya);
        const float G2       = 1 / (Lambda_o + Lambda_i + 1);
        const float G1       = 1 / (Lambda_o + 1);
        const float Fr       = fresnel_dielectric(dot(m, wo_l), l.eta);
        float pdf            = (G1 * D * 0.25f) / wo_l.z;
        float out            = G2 / G1;
        if (l.refract == 2) {
            pdf *= Fr;
            return { V3(out), pdf };
        }
        return { V3(out * Fr), pdf };
    }
    if ((l.refract == 1 || l.refract == 2) && wi_l.z < 0 && wo_l.z > 0.0f) {
        // half-vector of the refraction (Walter et al. 2007, eq. 16)
        V3 ht = -(wi_l * l.eta + wo_l);
        if (l.eta < 1.0f)
            ht = -ht;
        const V3 Ht       = normalize(ht);
        const float cosHO = dot(Ht, wo_l);
        const float Ft    = 1.0f - fresnel_dielectric(cosHO, l.eta);
        if (Ft > 0) {
            const float cosHI = dot(Ht, wi_l);
            if (Ht.z <= 0.0f)
                return { V3(0.0f), 0.0f };
            const float Dt       = bk ? beckmann_D(Ht, xa, ya) : ggx_D(Ht, xa, ya);
            const float Lambda_o = bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya);
            const float Lambda_i = bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya);
            const float G2       = 1 / (Lambda_o + Lambda_i + 1);
            const float G1       = 1 / (Lambda_o + 1);
            const float invHt2   = 1 / dot(ht, ht);
            float pdf = (std::fabs(cosHI * cosHO) * (l.eta * l.eta) * (G1 * Dt) * invHt2) / wo_l.z;
            float out = G2 / G1;
            if (l.refract == 2) {
                pdf *= Ft;
                return { V3(out), pdf };
            }
            return { V3(out * Ft), pdf };
        }
    }
    return { V3(0.0f), 0.0f };
}

// fresnel_refraction (src/testrender/optics.h:33-62)
float fresnel_refraction(V3 I, V3 N, float eta, V3& T)
{
    float cosi = -dot(I, N);
    V3 Nn;
    float neta;
    if (cosi > 0) {
        neta = 1 / eta;
        Nn   = N;
    } else {
        cosi = -cosi;
        neta = eta;
        Nn   = -N;
    }
    float arg = 1.0f - (neta * neta * (1.0f - cosi * cosi));
    if (arg >= 0) {
        float dnp = sqrtf(arg);
        float nK  = (neta * cosi) - dnp;
        T         = I * neta + Nn * nK;
        return 1 - fresnel_dielectric(cosi, eta);
    }
    T = V3(0.0f);
    return 0;
}

// TangentFrame::get: x u + y v + z w
V3 frame_get(const Frame& f, float x, float y, float z) { return f.u * x + f.v * y + f.w * z; }

// concentric square-to-disk map (Shirley and Chiu)
void to_unit_disk(float& x, float& y)
{
    const float quarter_pi = float(M_PI_4);
    const float half_pi    = float(M_PI_2);
    float phi, r;
    const float a = 2 * x - 1;
    const float b = 2 * y - 1;
    if (a * a > b * b) {
        r   = a;
        phi = quarter_pi * (b / a);
    } else if (b != 0) {
        r   = b;
        phi = half_pi - quarter_pi * (a / b);
    } else {
        r   = 0;
        phi = 0;
    }
    // OIIO::fast_sincos(phi, &x, &y): x takes the sine, y the cosine
    ofast_sincosf(phi, &x, &y);
    x = x * r;
    y = y * r;
}

// GGX visible-slope sampling for a stretched view direction at cos_theta
void ggx_sample_slope(float cos_theta, float u, float v, float* sx, float* sy)
{
    const float c = cos_theta < 1e-6f ? float(1e-6f) : cos_theta;
    const float Q = (1 + c) * u - c;
    const float num = c * sqrtf((1 - c) * (1 + c)) - Q * sqrtf((1 - Q) * (1 + Q));
    float den       = (Q - c) * (Q + c);
    const float eps = 1.0f / 4294967296.0f;
    if (fabsf(den) < eps)
        den = copysignf(eps, den);
    *sx            = num / den;
    const float Ru = 1 - 2 * v;
    const float u2 = fabsf(Ru);
    const float z  = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f))
                    / (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.0f) + 0.597999f);
    *sy = copysignf(1.0f, Ru) * z * sqrtf(1.0f + *sx * *sx);
}

// Beckmann visible-slope sampling (BeckmannDist::sampleSlope,
// src/testrender/shading.cpp:516-566: Jakob's improved routine, a coarse
// root, one Newton step and a second when the first leaves a residual),
// with the oracle's erf / erf^-1 / exp standing in for OpenImageIO's fast_
// versions
void beckmann_sample_slope(float cos_theta, float u, float v, float* sx, float* sy)
{
    const float sqrt_pi_inv = 1 / sqrtf(float(M_PI));
    const float ct          = cos_theta < 1e-6f ? float(1e-6f) : cos_theta;
    const float tan_i       = sqrtf(1 - ct * ct) / ct;
    const float cot_i       = 1 / tan_i;
    const float c           = ofast_erff(cot_i);
    const float K           = tan_i * sqrt_pi_inv;
    const float y_approx    = u * (1.0f + c + K * (1 - c * c));
    const float y_exact     = u * (1.0f + c + K * ofast_expf(-cot_i * cot_i));
    float b = K > 0 ? (0.5f - sqrtf(K * (K - y_approx + 1.0f) + 0.25f)) / K : y_approx - 1.0f;
    float inv_erf = ofast_ierff(b);
    float value   = 1.0f + b + K * ofast_expf(-inv_erf * inv_erf) - y_exact;
    if (fabsf(value) > 1e-6f) {
        b -= value / (1 - inv_erf * tan_i);
        inv_erf = ofast_ierff(b);
        value   = 1.0f + b + K * ofast_expf(-inv_erf * inv_erf) - y_exact;
        b -= value / (1 - inv_erf * tan_i);
        *sx = ofast_ierff(b);
    } else {
        *sx = inv_erf;
    }
    *sy = ofast_ierff(2.0f * v - 1.0f);
}

V3 ggx_sample_micronormal(V3 wo, float xa, float ya, float u, float v, bool beckmann)
{
    V3 swo(wo.x * xa, wo.y * ya, wo.z);
    swo                   = normalize(swo);
    const float cos_theta = std::max<float>(swo.z, 0.0f);
    float cos_phi = 1, sin_phi = 0;
    if (cos_theta < 0.99999f) {
        const float inv = 1 / sqrtf(SQR(swo.x) + SQR(swo.y));
        cos_phi         = swo.x * inv;
        sin_phi         = swo.y * inv;
    }
    float slx, sly;
    if (beckmann)
        beckmann_sample_slope(cos_theta, u, v, &slx, &sly);
    else
        ggx_sample_slope(cos_theta, u, v, &slx, &sly);
    float tx         = cos_phi * slx - sin_phi * sly;
    float ty         = sin_phi * slx + cos_phi * sly;
    tx               = tx * xa;
    ty               = ty * ya;
    const float mlen = sqrtf(tx * tx + ty * ty + 1);
    return V3(fabsf(tx) < mlen ? -tx / mlen : float(1.0f),
              fabsf(ty) < mlen ? -ty / mlen : float(1.0f),
              1.0f / mlen);
}

BsdfDirSample lobe_sample(const Lobe& l, V3 wo, float rx, float ry, float rz)
{
    BsdfDirSample r = { V3(0.0f), V3(0.0f), 0.0f };
    if (l.kind == kLobeMx) {
        const mtx::Sample s = mtx::mx_sample(l.mx, wo, rx, ry);
        return { s.wi, s.weight, s.pdf };
    }
    if (l.kind == kLobeEmission)
        return r;
    if (l.kind >= kLobeReflection && l.kind <= kLobeTransparent) {
        // Reflection / Refraction / Transparent::sample (shading.cpp:1098-1149)
        const float inf = std::numeric_limits<float>::infinity();
        if (l.kind == kLobeTransparent)
            return { -wo, V3(1.0f), inf };
        if (l.kind == kLobeRefraction) {
            V3 wi;
            const float Ft = fresnel_refraction(-wo, l.N, l.eta, wi);
            return { wi, V3(Ft), inf };
        }
        const float cosNO = dot(l.N, wo);
        if (cosNO > 0)
            return { l.N * (2 * cosNO) - wo, V3(fresnel_dielectric(cosNO, l.eta)), inf };
        return r;
    }
    if (l.kind == kLobePhong) {
        // Phong::sample (shading.cpp:342-362)
        const float cosNO = dot(l.N, wo);
        if (cosNO > 0) {
            const V3 R      = l.N * (2 * cosNO) - wo;
            const float phi = 2 * float(M_PI) * rx;
            float sp, cp;
            ofast_sincosf(phi, &sp, &cp);
            const float cosTheta  = ofast_safe_powf(ry, 1 / (l.alpha + 1));
            const float sinTheta2 = 1 - cosTheta * cosTheta;
            const float sinTheta  = sinTheta2 > 0 ? sqrtf(sinTheta2) : float(0.0f);
            const V3 wi = frame_get(frame_from_normal(R), cp * sinTheta, sp * sinTheta, cosTheta);
            const BsdfSample e = lobe_eval(l, wo, wi);
            r.wi     = wi;
            r.weight = e.weight;
            r.pdf    = e.pdf;
        }
        return r;
    }
    if (l.kind == kLobeDiffuse || l.kind == kLobeTranslucent) {
        to_unit_disk(rx, ry);
        const float cos_theta = sqrtf(std::max<float>(1 - rx * rx - ry * ry, 0.0f));
        r.wi     = frame_get(frame_from_normal(l.N), rx, ry, cos_theta);
        r.weight = V3(1.0f);
        r.pdf    = cos_theta * float(M_1_PI);
        return r;
    }
    const Frame tf    = frame_from_normal_tangent(l.N, l.aniso ? l.U : V3(0.0f));
    const V3 wo_l     = tolocal(tf, wo);
    const float cosNO = wo_l.z;
    if (!(cosNO > 0))
        return r;
    const float xa = l.alpha, ya = l.aniso ? l.ay : l.alpha;
    const bool bk        = l.kind == kLobeBeckmannRefl;
    const V3 m           = ggx_sample_micronormal(wo_l, xa, ya, rx, ry, bk);
    const float cosMO    = dot(m, wo_l);
    const float F        = fresnel_dielectric(cosMO, l.eta);
    // Microfacet<GGX, Refract>::sample (src/testrender/shading.cpp:650-712)
    if (l.refract == 0 || (l.refract == 2 && rz < F)) {
        const V3 wi_l        = m * (2.0f * cosMO) - wo_l;
        const float D        = (bk ? beckmann_D(m, xa, ya) : ggx_D(m, xa, ya));
        const float Lambda_o = (bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya));
        const float Lambda_i = (bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya));
        const float G2       = 1 / (Lambda_o + Lambda_i + 1);
        const float G1       = 1 / (Lambda_o + 1);
        r.wi                 = frame_get(tf, wi_l.x, wi_l.y, wi_l.z);
        r.pdf                = (G1 * D * 0.25f) / cosNO;
        if (l.refract == 2) {
            r.pdf *= F;
            r.weight = V3(G2 / G1);
        } else {
            r.weight = V3(F * (G2 / G1));
        }
        return r;
    }
    const V3 M = frame_get(tf, m.x, m.y, m.z);
    V3 wi;
    const float Ft       = fresnel_refraction(-wo, M, l.eta, wi);
    const V3 wi_l        = tolocal(tf, wi);
    const float cosHO    = dot(m, wo_l);
    const float cosHI    = dot(m, wi_l);
    const float D        = (bk ? beckmann_D(m, xa, ya) : ggx_D(m, xa, ya));
    const float Lambda_o = (bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya));
    const float Lambda_i = (bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya));
    const float G2       = 1 / (Lambda_o + Lambda_i + 1);
    const float G1       = 1 / (Lambda_o + 1);
    const V3 ht          = -(wi_l * l.eta + wo_l);
    const float invHt2   = 1.0f / dot(ht, ht);
    r.wi  = wi;
    r.pdf = (std::fabs(cosHI * cosHO) * (l.eta * l.eta) * (G1 * D) * invHt2) / std::fabs(wo_l.z);
    if (l.refract == 2) {
        r.pdf *= Ft;
        r.weight = V3(G2 / G1);
    } else {
        r.weight = V3(Ft * (G2 / G1));
    }
    return r;
}

V3 lobe_albedo(const Lobe& l, V3 wo)
{
    if (l.kind == kLobeMx)
        return mtx::mx_albedo(l.mx);
    if (l.kind == kLobeDiffuse || l.kind == kLobeTranslucent || l.kind == kLobePhong
        || l.kind == kLobeTransparent)
        return V3(1.0f);  // BSDF::get_albedo (src/testrender/shading.h:287-290)
    if (l.kind == kLobeEmission)
        return V3(0.0f);  // never picked by the sampler
    if (l.kind == kLobeReflection || l.kind == kLobeFresnelReflection) {
        const float cosNO = dot(l.N, wo);  // Reflection::get_albedo (shading.cpp:1086-1092)
        return V3(cosNO > 0 ? fresnel_dielectric(cosNO, l.eta) : float(1.0f));
    }
    if (l.kind == kLobeRefraction)
        return V3(1 - fresnel_dielectric(dot(l.N, wo), l.eta));  // shading.cpp:1117-1121
    // Microfacet::get_albedo (shading.cpp:580-588)
    if (l.refract == 2)
        return V3(1.0f);
    float fr = fresnel_dielectric(dot(l.N, wo), l.eta);
    return V3(l.refract ? 1 - fr : fr);
}

void update_eval(V3* w, float* pdf, V3 ow, float opdf, float b)
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

}  // namespace

float fresnel_dielectric(float cosi, float eta)
{
    if (eta == 0)
        return 1;
    if (cosi < 0.0f)
        eta = 1.0f / eta;
    float c = fabsf(cosi);
    float g = eta * eta - 1 + c * c;
    if (g > 0) {
        g       = sqrtf(g);
        float A = (g - c) / (g + c);
        float B = (c * (g + c) - 1) / (c * (g - c) + 1);
        return 0.5f * A * A * (1 + B * B);
    }
    return 1.0f;
}

void Composite::prepare(V3 wo)
{
    // path_weight = (1,1,1), absorb = false
    const V3 pw(1.0f);
    float total = 0;
    for (int i = 0; i < n; i++) {
        pdfs[i] = dot(lobes[i].weight, pw * lobe_albedo(lobes[i], wo)) / (pw.x + pw.y + pw.z);
        total += pdfs[i];
    }
    if (total > 0 || total > 1)
        for (int i = 0; i < n; i++)
            pdfs[i] /= total;
}

BsdfDirSample Composite::sample(V3 wo, float rx, float ry, float rz) const
{
    (void)rz;  // reflection-only lobes draw no third number
    float accum = 0;
    for (int i = 0; i < n; i++) {
        if (rx < (pdfs[i] + accum)) {
            rx = (rx - accum) / pdfs[i];
            rx = std::min<float>(rx, 0.99999994f);
            BsdfDirSample s = lobe_sample(lobes[i], wo, rx, ry, rz);
            const V3 wsc    = lobes[i].weight * (1 / pdfs[i]);
            s.weight        = s.weight * wsc;
            s.pdf           = s.pdf * pdfs[i];
            if (s.pdf == 0.0f)
                return { V3(0.0f), V3(0.0f), 0.0f };
            for (int j = 0; j < n; j++) {
                if (i != j) {
                    BsdfSample b = lobe_eval(lobes[j], wo, s.wi);
                    b.weight     = b.weight * lobes[j].weight;
                    update_eval(&s.weight, &s.pdf, b.weight, b.pdf, pdfs[j]);
                }
            }
            return s;
        }
        accum += pdfs[i];
    }
    return { V3(0.0f), V3(0.0f), 0.0f };
}

BsdfSample Composite::eval(V3 wo, V3 wi) const
{
    BsdfSample s = { V3(0.0f), 0.0f };
    for (int i = 0; i < n; i++) {
        BsdfSample b = lobe_eval(lobes[i], wo, wi);
        b.weight     = b.weight * lobes[i].weight;
        update_eval(&s.weight, &s.pdf, b.weight, b.pdf, pdfs[i]);
    }
    return s;
}

}  // namespace oref
