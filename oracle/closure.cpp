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
#include "closure.h"

#include <algorithm>
#include <cfloat>

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

BsdfSample lobe_eval(const Lobe& l, V3 wo, V3 wi)
{
    if (l.kind == kLobeDiffuse) {
        const float pdf = std::max<float>(dot(l.N, wi), 0.0f) * float(M_1_PI);
        return { V3(1.0f), pdf };
    }
    Frame tf        = frame_from_normal_tangent(l.N, V3(0.0f));
    const V3 wo_l   = tolocal(tf, wo);
    const V3 wi_l   = tolocal(tf, wi);
    const float xa  = l.alpha, ya = l.alpha;
    if (wo_l.z > 0 && wi_l.z > 0) {
        const V3 m           = normalize(wi_l + wo_l);
        const float D        = ggx_D(m, xa, ya);
        const float Lambda_o = ggx_lambda(wo_l, xa, ya);
        const float Lambda_i = ggx_lambda(wi_l, xa, ya);
        const float G2       = 1 / (Lambda_o + Lambda_i + 1);
        const float G1       = 1 / (Lambda_o + 1);
        const float Fr       = fresnel_dielectric(dot(m, wo_l), l.eta);
        float pdf            = (G1 * D * 0.25f) / wo_l.z;
        float out            = G2 / G1;
        return { V3(out * Fr), pdf };
    }
    return { V3(0.0f), 0.0f };
}

V3 lobe_albedo(const Lobe& l, V3 wo)
{
    if (l.kind == kLobeDiffuse)
        return V3(1.0f);
    float fr = fresnel_dielectric(dot(l.N, wo), l.eta);
    return V3(fr);
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
