// Kernel argument block shared by the host executor and emitted group
// kernels (plain data; identical layout on both sides).
#pragma once

#if !defined(__HIPCC_RTC__)
#include <cstdint>
#endif

namespace oslmi {

// Shader globals, the varying part of BatchedShaderGlobals
// (src/include/OSL/batched_shaderglobals.h:77-140) as SoA planes.
enum GlobalIndex {
    kG_P = 0, kG_dPdx, kG_dPdy, kG_N, kG_Ng, kG_I, kG_u, kG_v, kG_dudx, kG_dudy, kG_dvdx, kG_dvdy,
    kG_time, kG_dPdu, kG_dPdv, kG_Ps, kG_COUNT
};

#define OSLMI_MAX_OUTPUT_PLANES 32

struct GroupArgs {
    int64_t n;
    // planes per global (3 for triples, 1 for floats); NULL = uniform value
    const float* g[kG_COUNT];
    float gu[kG_COUNT][3];
    const void* colorsys;  // ColorSysDev*
    float* out[OSLMI_MAX_OUTPUT_PLANES];
    // optional fused closure evaluation: wo/wi 3 planes, eval 4 planes
    const float* wo;
    const float* wi;
    float* eval;
};

}  // namespace oslmi
