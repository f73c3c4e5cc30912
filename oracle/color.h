// CPU oracle: color system (blackbody, color-space transforms).
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
#pragma once
#include "oracle_core.h"

namespace oref {

// ColorSystem (src/liboslexec/opcolor.h:40-137, opcolor.cpp:142-211)
struct ColorSys {
    float xyz2rgb[3][3];
    float rgb2xyz[3][3];
    V3 lum_scale;
    V3 bb_table[317];
    int bb_entries;
};

// set_colorspace for one of the reference's chroma entries
// ("Rec709", "sRGB", "NTSC", "EBU", "PAL", "SECAM", "SMPTE", "HDTV", "CIE",
// "AdobeRGB", "XYZ", "ACES2065-1", "ACEScg"); returns false if unknown.
bool colorsys_init(ColorSys& cs, const char* name);

V3 blackbody_rgb(const ColorSys& cs, float T);       // opcolor_impl.h:568
V3 xyz_to_rgb(const ColorSys& cs, V3 c);             // C * m_XYZ2RGB
D2<V3> xyz_to_rgb(const ColorSys& cs, const D2<V3>& c);
float luminance(const ColorSys& cs, V3 c);            // opcolor.h:58

// Color-space names known without OpenColorIO (opcolor.cpp:275-380):
// 0 rgb (also "RGB", the system's own name and, for transformc, "linear"),
// 1 hsv, 2 hsl, 3 YIQ, 4 XYZ, 5 xyY, 6 sRGB (transformc only); -1 otherwise.
int color_space_code(const char* name, bool transformc);
// ColorSystem::transformc (opcolor.cpp:324-380) and to_rgb (:275-296)
V3 transformc(const ColorSys& cs, int from, int to, V3 c);

}  // namespace oref
