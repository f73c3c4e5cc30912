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
    // the shading system's transform table (33 floats per space pair)
    const float* xf;
    // optional closure sampling (kernels emitted with sampling): rnd 3
    // planes (rx, ry, rz), samp 7 planes (wi xyz, weight rgb, pdf)
    const float* rnd;
    float* samp;
    // libbsdl energy tables (kernels with MaterialX closures)
    const float* bsdl;
    // optional closure export (kernels emitted with export): 1 +
    // clo_max * OSLMI_CLOSURE_RECORD planes (count, then per lobe id,
    // weight rgb, parameters)
    float* clo;
    int32_t clo_max;
    int32_t clo_pad;
};

// planes per exported lobe: id, weight rgb, 21 parameter floats (the
// largest set, generalized_schlick_bsdf's)
#define OSLMI_CLOSURE_PARAMS 21
#define OSLMI_CLOSURE_RECORD (4 + OSLMI_CLOSURE_PARAMS)

// One unit of work of the persistent batch kernel: points [begin, end) of
// batch item `item` (which is also the group's case in the kernel).
struct BatchTile {
    int32_t item;
    int32_t pad;
    int64_t begin, end;
};

struct BatchArgs {
    const GroupArgs* args;   // one per item, device memory
    const BatchTile* tiles;  // device memory
    int32_t tile_begin, tile_end;
    int32_t* counter;        // tile fetch counter, zero between launches
    int32_t ppt;             // passes of blockDim.x points per tile
    int32_t bb_lds;          // stage the blackbody table in LDS
    const float* bb;         // blackbody table (global memory)
};

}  // namespace oslmi
