// CPU oracle: renderer-side closure evaluation (testrender).
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
#pragma once
#include "mtx.h"
#include "oracle_core.h"

namespace oref {

enum LobeKind {
    kLobeDiffuse = 0, kLobeGGXRefl = 1, kLobeBeckmannRefl = 2, kLobeMx = 3,
    kLobeTranslucent = 4,  // Diffuse<1>: N holds the closure's N negated
    kLobePhong = 5,        // alpha holds the exponent
    // delta lobes (one direction, infinite pdf, nothing to evaluate)
    kLobeReflection = 6, kLobeFresnelReflection = 7, kLobeRefraction = 8, kLobeTransparent = 9,
    // emission(): light, not a BSDF; kept for the closure export (weight =
    // emitted radiance), ignored by prepare / eval / sample
    kLobeEmission = 10
};

struct Lobe {
    int kind;
    V3 weight;
    V3 N;
    float alpha;  // microfacet xalpha
    float eta;
    mtx::MxLobe mx;    // kLobeMx: a MaterialX lobe built at the shading point's wo
    mtx::MxParams mxp;  // kLobeMx: its closure parameters
    // microfacet(dist, N, U, xalpha, yalpha, ...): with aniso set, the lobe's
    // frame is from_normal_and_tangent(N, U) and yalpha is ay; otherwise
    // yalpha = alpha about N
    bool aniso  = false;
    V3 U        = V3(0.0f);
    float ay    = 0.0f;
    int refract = 0;  // Microfacet<Dist, refract>: 0 reflect, 1 refract, 2 both
};

struct BsdfSample {
    V3 weight;
    float pdf;
};

// a sampled direction with its MIS-combined weight and pdf
struct BsdfDirSample {
    V3 wi;
    V3 weight;
    float pdf;
};

// Flattened closure list + CompositeBSDF::prepare / eval
// (src/testrender/shading.h:319-372, MIS::update_eval sampling.h:173-206)
struct Composite {
    int n = 0;
    Lobe lobes[8];
    float pdfs[8];
    void add(const Lobe& l)
    {
        if (n < 8)
            lobes[n++] = l;
    }
    void prepare(V3 wo);
    BsdfSample eval(V3 wo, V3 wi) const;
    BsdfDirSample sample(V3 wo, float rx, float ry, float rz) const;
};

float fresnel_dielectric(float cosi, float eta);  // src/testrender/optics.h:13-31

}  // namespace oref
