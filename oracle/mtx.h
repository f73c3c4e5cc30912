// CPU oracle: libbsdl MaterialX lobes (dielectric, generalized Schlick,
// sheen), their energy tables, and testrender's layering and CompositeBSDF.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
#pragma once
#include "oracle_core.h"

namespace oref {
namespace mtx {

// Energy table sizes (MTX/bsdf_dielectric_decl.h:34-37,
// microfacet_tools_decl.h:64-67).
constexpr int kNc = 16, kNr = 16, kNfRefl = 32;
constexpr int kLutMiniGGX    = 0;                       // 16 x 16 floats
constexpr int kLutReflFront  = kNc * kNr;               // 32 x 16 x 16 floats
constexpr int kLutFloats     = kNc * kNr * (1 + kNfRefl);

// The baked tables (genluts.cpp:128-192): loaded from bsdl_luts.bin beside
// liboslref.so when present, else baked in-process and saved there.
const float* luts();
// Bakes both tables into out[kLutFloats] (nthreads CPU threads).
void bake_luts(float* out, int nthreads);
bool save_luts(const char* path, const float* tables);

struct V2 {
    float x, y;
};

// Result of a lobe or composite evaluation / sample (bsdl::Sample and
// testrender BSDF::Sample).
struct Sample {
    V3 wi;
    V3 weight;
    float pdf;
    float roughness;
};

// One MaterialX closure lobe (testrender's BSDL_WRAP around libbsdl's
// DielectricLobe / SchlickLobe / SheenLobe, front facing, path roughness
// 0, reflection only), built at wo: parameters, then opaque state.
struct MxParams {
    int kind;  // 0 dielectric_bsdf, 1 generalized_schlick_bsdf, 2 sheen_bsdf
    V3 N, U;
    V3 tint;   // reflection tint; sheen albedo
    float rx, ry;  // roughness_x, roughness_y; sheen: rx = roughness
    float ior;     // dielectric
    V3 F0, F90;    // Schlick
    float exponent;
};
struct MxLobe {
    alignas(16) unsigned char state[384];
};
void mx_build(MxLobe* L, const MxParams& p, V3 wo);
Sample mx_eval(const MxLobe& L, V3 wo, V3 wi);              // world directions
Sample mx_sample(const MxLobe& L, V3 wo, float rx, float ry);  // world wi
V3 mx_albedo(const MxLobe& L);
float mx_filter_o(const MxLobe& L);

// Per-point parameters of the layered material
// layer(sheen_bsdf, layer(dielectric_bsdf, generalized_schlick_bsdf)).
struct LayeredParams {
    V3 N;
    float coat_ior, coat_roughness;   // dielectric_bsdf: ior, roughness_x = roughness_y
    V3 coat_refl_tint, coat_refr_tint;
    float base_roughness;             // generalized_schlick_bsdf roughness_x = roughness_y
    V3 base_f0, base_f90;
    float base_exponent;
    V3 base_refl_tint, base_refr_tint;
    float sheen_roughness;            // sheen_bsdf
    V3 sheen_albedo;
    int sheen_mode;                   // 0: Conty-Kulla (the only mode restated here)
};

// The flattened closure of one shading point (testrender
// process_bsdf_closure + CompositeBSDF::prepare), ready for eval/sample.
struct LayeredBsdf;
LayeredBsdf* layered_new();
void layered_free(LayeredBsdf*);
void layered_build(LayeredBsdf*, const LayeredParams& p, V3 wo);
Sample layered_eval(const LayeredBsdf*, V3 wo, V3 wi);
Sample layered_sample(const LayeredBsdf*, V3 wo, float rx, float ry, float rz);
// lobe weights (3 RGB) and selection pdfs (3), for tests
void layered_weights(const LayeredBsdf*, float* w9, float* pdf3);

// ---- building blocks exposed for unit tests --------------------------
float ggx_D(float ax, float ay, V3 m);
float ggx_G1(float ax, float ay, V3 w);
float ggx_G2_G1(float ax, float ay, V3 wi, V3 wo);
float ggx_D_refl_D(float ax, float ay, V3 wo, V3 m);
V3 ggx_sample_vndf(float ax, float ay, V3 wo, float u, float v);
V3 ggx_sample_bvndf(float ax, float ay, V3 wo, float u, float v, bool libm_trig);
V2 square_to_unit_disc(float u, float v);
V3 sample_uniform_hemisphere(float u, float v);
float dielectric_fresnel(float eta, float c);
float emiss_eval(int table, float roughness, float fresnel_index, float c);
float conty_kulla_albedo(float a, float cosNO);
float conty_kulla_D(float a, V3 m);
// Sample::update (bsdf_impl.h:16-38) on 3 channels.
void sample_update(V3* w, float* pdf, V3 ow, float opdf, float cpdf);

}  // namespace mtx
}  // namespace oref
