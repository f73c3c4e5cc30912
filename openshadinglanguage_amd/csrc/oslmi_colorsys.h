// Host-side construction of the color system tables and the Gabor uniform
// parameters (run once per shading system / per noise call site, never per
// point). Restates ColorSystem::set_colorspace (src/liboslexec/opcolor.cpp:
// 142-211) and GaborUniformParams (src/liboslnoise/sfm_gabornoise.h:
// 100-135) in single precision, operation for operation.
#pragma once
#include <string>

#include "oslmi_color.h"
#include "oslmi_gabor.h"

namespace oslmi {

// Fills cs for a named color space; returns false for an unknown name.
bool build_colorsys(const std::string& name, ColorSysDev& cs);

// NoiseParams fields used by Gabor.
struct GaborOptions {
    int anisotropic = 0;
    int do_filter   = 1;
    float direction[3] = { 1.0f, 0.0f, 0.0f };
    float bandwidth = 1.0f;
    float impulses  = 16.0f;
};
GaborUniform build_gabor_uniform(const GaborOptions& o);

}  // namespace oslmi
