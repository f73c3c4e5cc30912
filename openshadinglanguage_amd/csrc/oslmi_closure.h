// CDNA4 device library: closure lobes and their MIS combination.
//
// OSL builds a per-point closure tree (ADD/MUL nodes over ClosureComponents,
// src/include/OSL/oslclosure.h:41-145); the renderer flattens it to at most
// eight weighted lobes and evaluates them with multiple-importance weights
// (testrender: CompositeBSDF, src/testrender/shading.h:319-437). On the GPU
// the flattened form is the only form: a fixed-capacity lobe list held in
// registers, built directly by the emitted shader code.
//
//   Diffuse<0>                 src/testrender/shading.cpp:301-312
//   Microfacet<GGXDist,0>      src/testrender/shading.cpp:572-615, :713-757
//   GGXDist                    src/testrender/shading.cpp:497-506
//   BeckmannDist F, Lambda     src/testrender/shading.cpp:510-522 (eval only;
//     OIIO::fast_exp there, the library's own exp here)
//   TangentFrame::from_normal  src/testrender/sampling.h:17-42
//   fresnel_dielectric         src/testrender/optics.h:13-31
//   CompositeBSDF prepare/eval src/testrender/shading.h:323-372
//   MIS::update_eval           src/testrender/sampling.h:173-206
//   CompositeBSDF::sample      src/testrender/shading.h:376-402
//   Diffuse<0>::sample         src/testrender/shading.cpp:313-320 with
//     sample_cosine_hemisphere, to_unit_disk  sampling.h:70-100
//   Microfacet<GGXDist,0>::sample  src/testrender/shading.cpp:655-685,
//     sampleMicronormal :759-795, GGXDist::sampleSlope :483-507
//
// Kernels whose closures include MaterialX lobes (dielectric_bsdf,
// generalized_schlick_bsdf, sheen_bsdf, layer) are compiled with
// OSLMI_MX_CLOSURES defined: their lobes then also carry a libbsdl lobe
// state (oslmi_mtx.h), and closure values the layer opacity
// (testrender's evaluate_layer_opacity) of their last component.
#pragma once
#include "oslmi_math.h"
#include "oslmi_types.h"
// after oslmi_types.h, which gives the run-time compiler its integer types
#include "oslmi_groupargs.h"
#ifdef OSLMI_MX_CLOSURES
#include "oslmi_mtx.h"
#endif

namespace oslmi {

// Lobe kinds (the library's own numbering).
enum ClosureId {
    kClosureDiffuse = 1, kClosureMicrofacetGGX = 2, kClosureMicrofacetBeckmann = 3,
    kClosureTranslucent = 4, kClosurePhong = 5,
    // delta lobes: one direction with an infinite pdf, nothing to evaluate
    kClosureReflection = 6, kClosureFresnelReflection = 7, kClosureRefraction = 8,
    kClosureTransparent = 9,
    // emission(): light, not a BSDF: kept in the list for the closure export
    // (its weight is the emitted radiance) and ignored by prepare / eval /
    // sample, as testrender adds it to Le instead of the CompositeBSDF
    kClosureEmission = 10,
    kClosureMxDielectric = 11, kClosureMxSchlick = 12, kClosureMxSheen = 13
};

struct Lobe {
    int id;
    V3 weight;
    V3 N;         // translucent: the closure's N negated (testrender's Diffuse<1>)
    float alpha;  // microfacet xalpha (and yalpha unless anisotropic); phong exponent
    float eta;
#ifdef OSLMI_FULL_MICROFACET
    // general microfacet lobes: the tangent U of the lobe's frame
    // (TangentFrame::from_normal_and_tangent), yalpha and the refract mode
    // (0 reflection, 1 refraction, 2 both)
    V3 U;
    float ay;
    int refract;
#endif
#ifdef OSLMI_MX_CLOSURES
    mtx::MxLobe mx;
#endif
#ifdef OSLMI_CLOSURE_EXPORT
    // the closure's parameters as testrender registers them
    // (src/testrender/shading.cpp:185-296), for the closure export
    float xp[OSLMI_CLOSURE_PARAMS];
#endif
};

// Every access below uses a compile-time slot index (loops are unrolled
// and guarded by the run-time count), so a list lives in registers; a
// run-time index would put it in scratch memory.
template<int MAXL> struct LobeList {
    int n;
    Lobe l[MAXL];
    float pdf[MAXL];
#ifdef OSLMI_MX_CLOSURES
    // nn: the value is not the null closure (a layer of two null closures
    // is not null yet holds no lobe); op: evaluate_layer_opacity of it.
    // testrender's closure walks stop at the first null pointer they pop
    // (a layer pushes its base even when that is null): tr / opend, the
    // walk of process_bsdf_closure / evaluate_layer_opacity ends inside
    // this value, so nothing queued before it is visited.
    int nn, tr, opend;
    V3 op;
    OSLMI_DEV LobeList() : n(0), nn(0), tr(0), opend(0), op(0.0f) {}
#else
    OSLMI_DEV LobeList() : n(0) {}
#endif
    // copy of a list of another capacity (lobes beyond MAXL are dropped,
    // as add() drops them)
    template<int K2> OSLMI_DEV LobeList(const LobeList<K2>& o) : n(0)
    {
#pragma unroll
        for (int k = 0; k < (K2 < MAXL ? K2 : MAXL); ++k)
            if (k < o.n)
                l[k] = o.l[k];
        n = o.n < MAXL ? o.n : MAXL;
#ifdef OSLMI_MX_CLOSURES
        nn    = o.nn;
        tr    = o.tr;
        opend = o.opend;
        op    = o.op;
#endif
    }
    OSLMI_DEV void add(const Lobe& x)
    {
#pragma unroll
        for (int k = 0; k < MAXL; ++k)
            if (k == n)
                l[k] = x;
        n = n < MAXL ? n + 1 : n;
    }
};

struct BsdfEval {
    V3 weight;
    float pdf;
};

OSLMI_DEV float fresnel_dielectric(float cosi, float eta)
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

// fresnel_refraction (src/testrender/optics.h:33-62): the refracted
// direction T of the incoming direction I (pointing at the surface) about
// N, and the Fresnel transmittance; 0 and T = 0 on total internal
// reflection
OSLMI_DEV float fresnel_refraction(V3 I, V3 N, float eta, V3* T)
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
    const float arg = 1.0f - (neta * neta * (1.0f - cosi * cosi));
    if (arg >= 0) {
        const float dnp = sqrtf(arg);
        const float nK  = (neta * cosi) - dnp;
        *T              = I * neta + Nn * nK;
        return 1 - fresnel_dielectric(cosi, eta);
    }
    *T = V3(0.0f);
    return 0;
}

struct Frame {
    V3 u, v, w;
};
// from_normal_and_tangent(n, 0) falls back to from_normal (Duff et al.)
OSLMI_DEV Frame frame_from_normal(V3 n)
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
OSLMI_DEV V3 tolocal(const Frame& f, V3 a) { return V3(dot(a, f.u), dot(a, f.v), dot(a, f.w)); }
// TangentFrame::from_normal_and_tangent (src/testrender/sampling.h:30-42):
// x = t - n (n.t) normalised, y = n x x; a tangent parallel to n (or
// zero) falls back to from_normal
OSLMI_DEV Frame frame_from_normal_tangent(V3 n, V3 t)
{
    V3 x              = t - n * dot(n, t);
    const float xlen2 = dot(x, x);
    if (xlen2 > 0) {
        x = x * (1.0f / sqrtf(xlen2));
        Frame f;
        f.u = x;
        f.v = cross(n, x);
        f.w = n;
        return f;
    }
    return frame_from_normal(n);
}
// the frame and yalpha of a microfacet lobe
OSLMI_DEV Frame lobe_frame(const Lobe& l)
{
#ifdef OSLMI_FULL_MICROFACET
    return frame_from_normal_tangent(l.N, l.U);
#else
    return frame_from_normal(l.N);
#endif
}
OSLMI_DEV float lobe_yalpha(const Lobe& l)
{
#ifdef OSLMI_FULL_MICROFACET
    return l.ay;
#else
    return l.alpha;
#endif
}
OSLMI_DEV int lobe_refract(const Lobe& l)
{
#ifdef OSLMI_FULL_MICROFACET
    return l.refract;
#else
    (void)l;
    return 0;
#endif
}
// Vec3::normalize as the reference's microfacet code uses it: unchanged
// when the length is zero
OSLMI_DEV V3 normalized_or_self(V3 h)
{
    const float hl = sqrtf(dot(h, h));
    return hl != 0.0f ? V3(h.x / hl, h.y / hl, h.z / hl) : h;
}

OSLMI_DEV float ggx_lambda(V3 w, float xa, float ya)
{
    float cosTheta2  = w.z * w.z;
    float cosPhi2st2 = (w.x * xa) * (w.x * xa);
    float sinPhi2st2 = (w.y * ya) * (w.y * ya);
    float a2         = cosTheta2 / (cosPhi2st2 + sinPhi2st2);
    return 0.5f * (-1.0f + sqrtf(1.0f + 1.0f / a2));
}

OSLMI_DEV float ggx_D(V3 Hr, float xa, float ya)
{
    float cosThetaM = Hr.z;
    if (cosThetaM > 0) {
        float cosPhi2st2 = (Hr.x / xa) * (Hr.x / xa);
        float sinPhi2st2 = (Hr.y / ya) * (Hr.y / ya);
        float cosThetaM2 = cosThetaM * cosThetaM;
        float cosThetaM4 = cosThetaM2 * cosThetaM2;
        float tanThetaM2 = (cosPhi2st2 + sinPhi2st2) / cosThetaM2;
        float F          = 1 / (3.14159274101257324219f * (1 + tanThetaM2) * (1 + tanThetaM2));
        return F / (xa * ya * cosThetaM4);
    }
    return 0;
}

// Microfacet<BeckmannDist,0>: the same evalLambda / evalD with Beckmann's
// slope distribution and Lambda approximation
OSLMI_DEV float beckmann_lambda(V3 w, float xa, float ya)
{
    float cosTheta2  = w.z * w.z;
    float cosPhi2st2 = (w.x * xa) * (w.x * xa);
    float sinPhi2st2 = (w.y * ya) * (w.y * ya);
    float a2         = cosTheta2 / (cosPhi2st2 + sinPhi2st2);
    float a          = sqrtf(a2);
    return a < 1.6f ? (1.0f - 1.259f * a + 0.396f * a2) / (3.535f * a + 2.181f * a2) : 0.0f;
}
OSLMI_DEV float beckmann_D(V3 Hr, float xa, float ya)
{
    float cosThetaM = Hr.z;
    if (cosThetaM > 0) {
        float cosPhi2st2 = (Hr.x / xa) * (Hr.x / xa);
        float sinPhi2st2 = (Hr.y / ya) * (Hr.y / ya);
        float cosThetaM2 = cosThetaM * cosThetaM;
        float cosThetaM4 = cosThetaM2 * cosThetaM2;
        float tanThetaM2 = (cosPhi2st2 + sinPhi2st2) / cosThetaM2;
        float F          = 0.31830987334251403809f * fexp(-tanThetaM2);  // float(1 / M_PI)
        return F / (xa * ya * cosThetaM4);
    }
    return 0;
}

// Phong::eval (src/testrender/shading.cpp:323-340): a cosine-power lobe
// about the mirror direction of wo; exponent in alpha
OSLMI_DEV BsdfEval phong_eval(const Lobe& l, V3 wo, V3 wi)
{
    BsdfEval r;
    r.weight          = V3(0.0f);
    r.pdf             = 0.0f;
    const float cosNI = dot(l.N, wi);
    const float cosNO = dot(l.N, wo);
    if (cosNI > 0 && cosNO > 0) {
        const V3 R        = l.N * (2 * cosNO) - wo;
        const float cosRI = dot(R, wi);
        if (cosRI > 0) {
            const float e = l.alpha;
            r.pdf         = (e + 1) * 0.15915493667125701904f * fsafe_pow(cosRI, e);  // float(M_1_PI / 2)
            r.weight      = V3(cosNI * (e + 2) / (e + 1));
        }
    }
    return r;
}

OSLMI_DEV BsdfEval lobe_eval(const Lobe& l, V3 wo, V3 wi)
{
    BsdfEval r;
#ifdef OSLMI_MX_CLOSURES
    // MaterialX lobes: eval_impl in the lobe frame at the wo they were
    // built at (BSDL_WRAP::eval, src/testrender/shading.cpp:88-93)
    if (l.id >= kClosureMxDielectric) {
        const mtx::Sample s = mtx::mx_eval_local(l.mx, l.id - kClosureMxDielectric, l.mx.f.local(wi));
        r.weight            = s.weight;
        r.pdf               = s.pdf;
        return r;
    }
#endif
    if (l.id == kClosureDiffuse || l.id == kClosureTranslucent) {
        r.weight = V3(1.0f);
        r.pdf    = fmaxf(dot(l.N, wi), 0.0f) * 0.31830987334251403809f;  // float(M_1_PI)
        return r;
    }
    if (l.id == kClosurePhong)
        return phong_eval(l, wo, wi);
    if (l.id >= kClosureReflection && l.id <= kClosureEmission) {
        r.weight = V3(0.0f);
        r.pdf    = 0.0f;
        return r;
    }
    Frame tf       = lobe_frame(l);
    V3 wo_l        = tolocal(tf, wo);
    V3 wi_l        = tolocal(tf, wi);
    const float xa = l.alpha, ya = lobe_yalpha(l);
    const int rf   = lobe_refract(l);
    const bool bk  = l.id == kClosureMicrofacetBeckmann;
    if (rf != 1 && wo_l.z > 0 && wi_l.z > 0) {
        V3 m     = normalized_or_self(wi_l + wo_l);
        float D  = bk ? beckmann_D(m, xa, ya) : ggx_D(m, xa, ya);
        float Lo = bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya);
        float Li = bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya);
        float G2 = 1 / (Lo + Li + 1);
        float G1 = 1 / (Lo + 1);
        float Fr = fresnel_dielectric(dot(m, wo_l), l.eta);
        r.pdf    = (G1 * D * 0.25f) / wo_l.z;
        if (rf == 2) {
            r.pdf *= Fr;
            r.weight = V3(G2 / G1);
        } else {
            r.weight = V3((G2 / G1) * Fr);
        }
        return r;
    }
#ifdef OSLMI_FULL_MICROFACET
    // transmission (shading.cpp:612-646): the half-vector of the
    // refraction, Fresnel transmittance, the Jacobian eta^2 |.|/|ht|^2
    if (rf != 0 && wi_l.z < 0 && wo_l.z > 0.0f) {
        V3 ht = -(wi_l * l.eta + wo_l);
        if (l.eta < 1.0f)
            ht = -ht;
        const V3 Ht       = normalized_or_self(ht);
        const float cosHO = dot(Ht, wo_l);
        const float Ft    = 1.0f - fresnel_dielectric(cosHO, l.eta);
        if (Ft > 0) {
            const float cosHI = dot(Ht, wi_l);
            if (Ht.z <= 0.0f) {
                r.weight = V3(0.0f);
                r.pdf    = 0.0f;
                return r;
            }
            const float Dt     = bk ? beckmann_D(Ht, xa, ya) : ggx_D(Ht, xa, ya);
            const float Lo     = bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya);
            const float Li     = bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya);
            const float G2     = 1 / (Lo + Li + 1);
            const float G1     = 1 / (Lo + 1);
            const float invHt2 = 1 / dot(ht, ht);
            r.pdf = (fabsf(cosHI * cosHO) * (l.eta * l.eta) * (G1 * Dt) * invHt2) / wo_l.z;
            if (rf == 2) {
                r.pdf *= Ft;
                r.weight = V3(G2 / G1);
            } else {
                r.weight = V3((G2 / G1) * Ft);
            }
            return r;
        }
    }
#endif
    r.weight = V3(0.0f);
    r.pdf    = 0.0f;
    return r;
}

OSLMI_DEV V3 lobe_albedo(const Lobe& l, V3 wo)
{
#ifdef OSLMI_MX_CLOSURES
    if (l.id >= kClosureMxDielectric)
        return mtx::mx_albedo(l.mx);
#endif
    if (l.id == kClosureDiffuse || l.id == kClosureTranslucent || l.id == kClosurePhong
        || l.id == kClosureTransparent)
        return V3(1.0f);  // BSDF::get_albedo (src/testrender/shading.h:287-290)
    if (l.id == kClosureEmission)
        return V3(0.0f);  // never picked by the sampler
    if (l.id == kClosureReflection || l.id == kClosureFresnelReflection) {
        // Reflection::get_albedo (shading.cpp:1086-1092)
        const float cosNO = dot(l.N, wo);
        return V3(cosNO > 0 ? fresnel_dielectric(cosNO, l.eta) : 1.0f);
    }
    if (l.id == kClosureRefraction)
        return V3(1 - fresnel_dielectric(dot(l.N, wo), l.eta));  // shading.cpp:1117-1121
    // Microfacet::get_albedo (shading.cpp:580-588)
    const int rf = lobe_refract(l);
    if (rf == 2)
        return V3(1.0f);
    const float fr = fresnel_dielectric(dot(l.N, wo), l.eta);
    return V3(rf ? 1 - fr : fr);
}

template<int MAXL> OSLMI_DEV void composite_prepare(LobeList<MAXL>& L, V3 wo)
{
    const V3 pw(1.0f);
    float total = 0;
#pragma unroll
    for (int i = 0; i < MAXL; i++)
        if (i < L.n) {
            L.pdf[i] = dot(L.l[i].weight, pw * lobe_albedo(L.l[i], wo)) / (pw.x + pw.y + pw.z);
            total += L.pdf[i];
        }
    if (total > 0 || total > 1)
#pragma unroll
        for (int i = 0; i < MAXL; i++)
            if (i < L.n)
                L.pdf[i] /= total;
}

OSLMI_DEV void mis_update(V3* w, float* pdf, V3 ow, float opdf, float b)
{
    if (b > 1.17549435e-38f) {
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

// ---- sampling -------------------------------------------------------------
struct BsdfDirSample {
    V3 wi;
    V3 weight;
    float pdf;
};

// TangentFrame::get: x u + y v + z w
OSLMI_DEV V3 frame_get(const Frame& f, float x, float y, float z)
{
    return f.u * x + f.v * y + f.w * z;
}

// concentric square-to-disk map; the sine lands in x, the cosine in y, as
// testrender's OIIO::fast_sincos(phi, &x, &y) call has it
OSLMI_DEV void to_unit_disk(float& x, float& y)
{
    float phi, r;
    const float a = 2 * x - 1;
    const float b = 2 * y - 1;
    if (a * a > b * b) {
        r   = a;
        phi = 0.785398185253143310547f * (b / a);  // float(M_PI_4)
    } else if (b != 0) {
        r   = b;
        phi = 1.57079637050628662109f - 0.785398185253143310547f * (a / b);
    } else {
        r   = 0;
        phi = 0;
    }
    fsincos(phi, &x, &y);
    x = x * r;
    y = y * r;
}

OSLMI_DEV void ggx_sample_slope(float cos_theta, float u, float v, float* sx, float* sy)
{
    const float c   = cos_theta < 1e-6f ? 1e-6f : cos_theta;
    const float Q   = (1 + c) * u - c;
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
// with the library's erf / erf^-1 / exp standing in for OpenImageIO's fast_
// versions
OSLMI_DEV void beckmann_sample_slope(float cos_theta, float u, float v, float* sx, float* sy)
{
    const float sqrt_pi_inv = 1 / sqrtf(3.14159274101257324219f);
    const float ct          = cos_theta < 1e-6f ? 1e-6f : cos_theta;
    const float tan_i       = sqrtf(1 - ct * ct) / ct;
    const float cot_i       = 1 / tan_i;
    const float c           = ferf(cot_i);
    const float K           = tan_i * sqrt_pi_inv;
    const float y_approx    = u * (1.0f + c + K * (1 - c * c));
    const float y_exact     = u * (1.0f + c + K * fexp(-cot_i * cot_i));
    float b = K > 0 ? (0.5f - sqrtf(K * (K - y_approx + 1.0f) + 0.25f)) / K : y_approx - 1.0f;
    float inv_erf = fierf(b);
    float value   = 1.0f + b + K * fexp(-inv_erf * inv_erf) - y_exact;
    if (fabsf(value) > 1e-6f) {
        b -= value / (1 - inv_erf * tan_i);
        inv_erf = fierf(b);
        value   = 1.0f + b + K * fexp(-inv_erf * inv_erf) - y_exact;
        b -= value / (1 - inv_erf * tan_i);
        *sx = fierf(b);
    } else {
        *sx = inv_erf;
    }
    *sy = fierf(2.0f * v - 1.0f);
}

OSLMI_DEV V3 ggx_sample_micronormal(V3 wo, float xa, float ya, float u, float v, bool beckmann)
{
    V3 swo(wo.x * xa, wo.y * ya, wo.z);
    const float sl = sqrtf(dot(swo, swo));
    if (sl != 0.0f)
        swo = V3(swo.x / sl, swo.y / sl, swo.z / sl);
    const float cos_theta = fmaxf(swo.z, 0.0f);
    float cos_phi = 1, sin_phi = 0;
    if (cos_theta < 0.99999f) {
        const float inv = 1 / sqrtf(swo.x * swo.x + swo.y * swo.y);
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
    return V3(fabsf(tx) < mlen ? -tx / mlen : 1.0f, fabsf(ty) < mlen ? -ty / mlen : 1.0f,
              1.0f / mlen);
}

OSLMI_DEV BsdfDirSample lobe_sample(const Lobe& l, V3 wo, float rx, float ry, float rz)
{
    BsdfDirSample r;
    r.wi     = V3(0.0f);
    r.weight = V3(0.0f);
    r.pdf    = 0.0f;
#ifdef OSLMI_MX_CLOSURES
    // MaterialX lobes: sample_impl's direction, weighted by eval_impl
    // there, back to world space (BSDL_WRAP::sample, shading.cpp:94-101)
    if (l.id >= kClosureMxDielectric) {
        const int kind = l.id - kClosureMxDielectric;
        V3 wl;
        if (!mtx::mx_sample_dir(l.mx, kind, rx, ry, &wl))
            return r;
        const mtx::Sample s = mtx::mx_eval_local(l.mx, kind, wl);
        r.wi                = l.mx.f.world(wl);
        r.weight            = s.weight;
        r.pdf               = s.pdf;
        return r;
    }
#endif
    if (l.id == kClosureEmission)
        return r;
    if (l.id >= kClosureReflection && l.id <= kClosureTransparent) {
        // Reflection / Refraction / Transparent::sample (shading.cpp:1098-1149):
        // the one direction, an infinite pdf
        const float inf = __uint_as_float(0x7f800000u);
        if (l.id == kClosureTransparent) {
            r.wi     = -wo;
            r.weight = V3(1.0f);
            r.pdf    = inf;
        } else if (l.id == kClosureRefraction) {
            V3 wi;
            const float Ft = fresnel_refraction(-wo, l.N, l.eta, &wi);
            r.wi     = wi;
            r.weight = V3(Ft);
            r.pdf    = inf;
        } else {
            const float cosNO = dot(l.N, wo);
            if (cosNO > 0) {
                r.wi     = l.N * (2 * cosNO) - wo;
                r.weight = V3(fresnel_dielectric(cosNO, l.eta));
                r.pdf    = inf;
            }
        }
        return r;
    }
    if (l.id == kClosurePhong) {
        // Phong::sample (shading.cpp:342-362): a cosine-power lobe about the
        // mirror direction, then the eval there
        const float cosNO = dot(l.N, wo);
        if (cosNO > 0) {
            const V3 R       = l.N * (2 * cosNO) - wo;
            const float phi  = 2 * 3.14159274101257324219f * rx;
            float sp, cp;
            fsincos(phi, &sp, &cp);
            const float cosTheta  = fsafe_pow(ry, 1 / (l.alpha + 1));
            const float sinTheta2 = 1 - cosTheta * cosTheta;
            const float sinTheta  = sinTheta2 > 0 ? sqrtf(sinTheta2) : 0;
            const V3 wi = frame_get(frame_from_normal(R), cp * sinTheta, sp * sinTheta, cosTheta);
            const BsdfEval e = phong_eval(l, wo, wi);
            r.wi     = wi;
            r.weight = e.weight;
            r.pdf    = e.pdf;
        }
        return r;
    }
    if (l.id == kClosureDiffuse || l.id == kClosureTranslucent) {
        to_unit_disk(rx, ry);
        const float cos_theta = sqrtf(fmaxf(1 - rx * rx - ry * ry, 0.0f));
        r.wi     = frame_get(frame_from_normal(l.N), rx, ry, cos_theta);
        r.weight = V3(1.0f);
        r.pdf    = cos_theta * 0.31830987334251403809f;
        return r;
    }
    const Frame tf    = lobe_frame(l);
    const V3 wo_l     = tolocal(tf, wo);
    const float cosNO = wo_l.z;
    if (!(cosNO > 0))
        return r;
    const float xa    = l.alpha, ya = lobe_yalpha(l);
    const int rf      = lobe_refract(l);
    const bool bk     = l.id == kClosureMicrofacetBeckmann;
    const V3 m        = ggx_sample_micronormal(wo_l, xa, ya, rx, ry, bk);
    const float cosMO = dot(m, wo_l);
    const float F     = fresnel_dielectric(cosMO, l.eta);
    if (rf == 0 || (rf == 2 && rz < F)) {
        // reflection; rz chooses it by the microfacet's Fresnel term when
        // the lobe both reflects and refracts
        const V3 wi_l  = m * (2.0f * cosMO) - wo_l;
        const float D  = (bk ? beckmann_D(m, xa, ya) : ggx_D(m, xa, ya));
        const float Lo = (bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya));
        const float Li = (bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya));
        const float G2 = 1 / (Lo + Li + 1);
        const float G1 = 1 / (Lo + 1);
        r.wi           = frame_get(tf, wi_l.x, wi_l.y, wi_l.z);
        r.pdf          = (G1 * D * 0.25f) / cosNO;
        if (rf == 2) {
            r.pdf *= F;
            r.weight = V3(G2 / G1);
        } else {
            r.weight = V3(F * (G2 / G1));
        }
        return r;
    }
#ifdef OSLMI_FULL_MICROFACET
    // transmission through the sampled microfacet (shading.cpp:685-710)
    const V3 M        = frame_get(tf, m.x, m.y, m.z);
    V3 wi;
    const float Ft    = fresnel_refraction(-wo, M, l.eta, &wi);
    const V3 wi_l     = tolocal(tf, wi);
    const float cosHO = dot(m, wo_l);
    const float cosHI = dot(m, wi_l);
    const float D     = (bk ? beckmann_D(m, xa, ya) : ggx_D(m, xa, ya));
    const float Lo    = (bk ? beckmann_lambda(wo_l, xa, ya) : ggx_lambda(wo_l, xa, ya));
    const float Li    = (bk ? beckmann_lambda(wi_l, xa, ya) : ggx_lambda(wi_l, xa, ya));
    const float G2    = 1 / (Lo + Li + 1);
    const float G1    = 1 / (Lo + 1);
    const V3 ht       = -(wi_l * l.eta + wo_l);
    const float invHt2 = 1.0f / dot(ht, ht);
    r.wi  = wi;
    r.pdf = (fabsf(cosHI * cosHO) * (l.eta * l.eta) * (G1 * D) * invHt2) / fabsf(wo_l.z);
    if (rf == 2) {
        r.pdf *= Ft;
        r.weight = V3(G2 / G1);
    } else {
        r.weight = V3(Ft * (G2 / G1));
    }
#endif
    return r;
}

template<int MAXL> OSLMI_DEV BsdfEval composite_eval(const LobeList<MAXL>& L, V3 wo, V3 wi);

// CompositeBSDF::sample: rx picks a lobe by the prepared pdfs and is
// rescaled for it; the other lobes are evaluated at the drawn direction and
// MIS-combined in closure order. Compile-time lobe slots throughout.
template<int MAXL>
OSLMI_DEV BsdfDirSample composite_sample(const LobeList<MAXL>& L, V3 wo, float rx, float ry,
                                         float rz)
{
    BsdfDirSample s;
    s.wi        = V3(0.0f);
    s.weight    = V3(0.0f);
    s.pdf       = 0.0f;
    float accum = 0;
    bool done   = false;
#pragma unroll
    for (int i = 0; i < MAXL; i++) {
        if (!done && i < L.n) {
            if (rx < (L.pdf[i] + accum)) {
                done    = true;
                float u = (rx - accum) / L.pdf[i];
                u       = fminf(u, 0.99999994f);
                s       = lobe_sample(L.l[i], wo, u, ry, rz);
                const V3 wsc = L.l[i].weight * (1 / L.pdf[i]);
                s.weight     = s.weight * wsc;
                s.pdf        = s.pdf * L.pdf[i];
                if (s.pdf == 0.0f) {
                    s.wi     = V3(0.0f);
                    s.weight = V3(0.0f);
                } else {
#pragma unroll
                    for (int j = 0; j < MAXL; j++)
                        if (j != i && j < L.n) {
                            BsdfEval b = lobe_eval(L.l[j], wo, s.wi);
                            mis_update(&s.weight, &s.pdf, b.weight * L.l[j].weight, b.pdf,
                                       L.pdf[j]);
                        }
                }
            } else {
                accum += L.pdf[i];
            }
        }
    }
    return s;
}

template<int MAXL> OSLMI_DEV BsdfEval composite_eval(const LobeList<MAXL>& L, V3 wo, V3 wi)
{
    BsdfEval s;
    s.weight = V3(0.0f);
    s.pdf    = 0.0f;
#pragma unroll
    for (int i = 0; i < MAXL; i++)
        if (i < L.n) {
            BsdfEval b = lobe_eval(L.l[i], wo, wi);
            mis_update(&s.weight, &s.pdf, b.weight * L.l[i].weight, b.pdf, L.pdf[i]);
        }
    return s;
}

#ifdef OSLMI_CLOSURE_EXPORT
// testrender's closure ids (src/testrender/shading.h:25-61) of the lobe kinds
OSLMI_DEV float export_id(int id)
{
    return id == kClosureDiffuse ? 3.0f                        // DIFFUSE_ID
           : id == kClosureTranslucent ? 5.0f                 // TRANSLUCENT_ID
           : id == kClosurePhong ? 6.0f                       // PHONG_ID
           : id == kClosureReflection ? 9.0f                  // REFLECTION_ID
           : id == kClosureFresnelReflection ? 10.0f          // FRESNEL_REFLECTION_ID
           : id == kClosureRefraction ? 11.0f                 // REFRACTION_ID
           : id == kClosureTransparent ? 12.0f                // TRANSPARENT_ID
           : id == kClosureEmission ? 1.0f                    // EMISSION_ID
           : id == kClosureMicrofacetGGX || id == kClosureMicrofacetBeckmann ? 8.0f  // MICROFACET_ID
           : id == kClosureMxDielectric ? 17.0f               // MX_DIELECTRIC_ID
           : id == kClosureMxSchlick ? 19.0f                  // MX_GENERALIZED_SCHLICK_ID
                                       : 23.0f;               // MX_SHEEN_ID
}
// Ci's lobes as process_bsdf_closure hands them to CompositeBSDF::add_bsdf:
// out plane 0 the count (at most maxl), then per lobe k planes 1 + k R + j
// (R = OSLMI_CLOSURE_RECORD): id, weight rgb, the parameters
template<int MAXL>
OSLMI_DEV void export_closure(const LobeList<MAXL>& L, float* out, int maxl, int64_t n, int64_t i)
{
    const int cnt = L.n < maxl ? L.n : maxl;
    out[i]        = (float)cnt;
#pragma unroll
    for (int k = 0; k < MAXL; k++)
        if (k < cnt) {
            float* r            = out + (int64_t)(1 + k * OSLMI_CLOSURE_RECORD) * n + i;
            r[0]                = export_id(L.l[k].id);
            r[n]                = L.l[k].weight.x;
            r[2 * n]            = L.l[k].weight.y;
            r[3 * n]            = L.l[k].weight.z;
#pragma unroll
            for (int j = 0; j < OSLMI_CLOSURE_PARAMS; j++)
                r[(4 + j) * n] = L.l[k].xp[j];
        }
}
#endif

}  // namespace oslmi
