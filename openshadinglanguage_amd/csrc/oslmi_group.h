// Shading system, shader groups and the group emitter (host side).
//
// A ShadingSys owns shader masters (parsed OSO) and groups. A group is an
// ordered list of layers (shader instances with parameter values) plus
// connections between layer outputs and inputs (ShaderGroupBegin / Parameter
// / Shader / ConnectShaders / ShaderGroupEnd, src/include/OSL/oslexec.h:
// 634-742). jit() lowers the whole group to one HIP kernel (emit.cpp) and
// compiles it with hiprtc for gfx950; execute() runs it over n points.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "oslmi.h"
#include "oslmi_groupargs.h"
#include "oslmi_matrix.h"
#include "oslmi_oso.h"

namespace oslmi {

struct ParamValue {
    OType type;
    std::vector<float> f;
    std::vector<int> i;
    std::vector<std::string> s;
};

struct Layer {
    std::string name;
    const ShaderMaster* master = nullptr;
    std::map<std::string, ParamValue> inst;  // instance parameter values
};

struct Connection {
    int src_layer, src_sym, dst_layer, dst_sym;
};

// Where an output lands in the caller's plane array.
struct OutputSlot {
    std::string name;
    int layer, sym;
    int planes;  // 1 (int/float) or 3 (triple)
    int first_plane;
};

// A renderer-side location of an output symbol (SymLocationDesc in the
// output arena): point i's value goes to base + offset + shadeindex * stride.
struct SymLoc {
    std::string name;
    int comps  = 1;  // 1 or 3 values
    int is_int = 0;
    int64_t offset = 0, stride = 0;
};

struct Group {
    std::string name;
    bool ended = false;
    std::vector<Layer> layers;
    std::vector<Connection> conns;
    std::vector<std::string> output_names;
    // compiled state
    bool jitted = false;
    std::string source;
    std::vector<OutputSlot> outputs;
    int output_planes        = 0;
    bool surface_closure     = false;  // some layer has the Ci global (shared by all layers)
    bool uses_colorsys       = false;
    bool uses_mx             = false;  // MaterialX closures: reads the libbsdl energy tables
    bool uses_xforms         = false;  // reads the transform / attribute table
    bool can_sample          = false;  // kernel samples Ci when A.samp is set
    bool can_export          = false;  // kernel exports Ci's lobes when A.clo is set
    std::vector<unsigned> globals_used;  // GlobalIndex values read
    hipModule_t module       = nullptr;
    hipFunction_t kernel     = nullptr;
    double jit_seconds       = 0.0;
    // oslmi_execute_wide16: symbol locations and the device workspace of
    // gathered globals and output planes (grown on demand)
    std::vector<SymLoc> symlocs;
    float* wide_ws        = nullptr;
    size_t wide_ws_floats = 0;
    std::mutex mu;
    int find_layer(const std::string& n) const;
};

// A named coordinate system the renderer registers: the matrix to "common"
// space (RendererServices::get_matrix) and, optionally, its inverse
// (get_inverse_matrix); without one the inverse is computed.
struct NamedXform {
    float m[16];
    float inv[16];
    bool has_inv = false;
};

// Transform table: one slot per (from, to) space pair some emitted kernel
// uses, kXformSlotFloats floats each: the from-to matrix (16), the
// transposed inverse for normals (16) and 1/0 for whether both spaces
// were known (1).
constexpr int kXformSlotFloats = 33;
constexpr int kMaxXformSlots   = 256;

// A renderer attribute (RendererServices::get_attribute): base type
// kInt / kFloat / kTriple / kMatrix and its value (an int in v[0]'s bits).
struct RendererAttr {
    int base = 0;
    float v[16] = {};
};
// Attribute table: one slot per attribute name some kernel reads,
// kAttrSlotFloats floats each: the value (16), the base type + 1 when the
// renderer has the attribute, else 0 (1).
constexpr int kAttrSlotFloats = 17;
constexpr int kMaxAttrSlots   = 256;

struct XformTable {
    std::mutex mu;
    std::map<std::string, NamedXform> named;
    std::vector<std::pair<std::string, std::string>> pairs;
    float* device = nullptr;  // kMaxXformSlots slots, allocated on first use
    bool dirty    = true;     // device copy out of date
    // renderer attributes, in the same table after the transform slots
    std::map<std::string, RendererAttr> attrs;
    std::vector<std::string> attr_names;
};

struct ShadingSys {
    std::map<std::string, std::unique_ptr<ShaderMaster>> masters;
    std::vector<std::unique_ptr<Group>> groups;
    std::string colorspace = "Rec709";
    std::string commonspace = "world";  // synonym of "common"
    std::string cache_dir;   // code-object cache ("" = none)
    std::string include_dir; // device library headers for hiprtc
    int waves_per_eu = 6;    // occupancy target of emitted kernels (0: compiler default)
    bool closure_sample = false;  // emit CompositeBSDF::sample into surface kernels
    bool closure_export = false;  // emit the closure export into surface kernels
    mutable XformTable xf;
    std::mutex mu;
};

// Slot of the (from, to) pair in the shading system's transform table,
// added on first use; -1 when the table is full.
int xform_slot(const ShadingSys& ss, const std::string& from, const std::string& to);

// Slot of a renderer attribute name in the same table (offset in floats from
// the table start), added on first use; -1 when the table is full.
int attr_slot(const ShadingSys& ss, const std::string& name);

// Brings the device copy of the transform table up to date (allocating it
// on first use); a no-op when nothing changed.
int xform_table_sync(const ShadingSys& ss);

// Emits the HIP source of the group's fused kernel; fills g.outputs,
// g.output_planes, g.surface_closure, g.uses_colorsys, g.globals_used.
bool emit_group(const ShadingSys& ss, Group& g, std::string& src, std::string& err);

// Emits the HIP source of the persistent batch kernel oslmi_batch_main
// (BatchArgs) over the given groups; groups[j] is batch item ids[j].
bool emit_batch(const ShadingSys& ss, const std::vector<Group*>& groups,
                const std::vector<int>& ids, std::string& src, std::string& err);

// hiprtc compile of src for gfx950 through the shading system's code-object
// cache (key: source, options and device library headers).
int build_code_object(const ShadingSys& ss, const std::string& src, const std::string& name,
                      std::vector<char>& code, double* seconds);

// The colour system (matrices, blackbody table) in device memory, built on
// first use of a device; with allocate false it only reports an existing
// copy (an error if this device has none yet).
struct ColorSysDev;
int device_colorsys(const ColorSysDev** out, bool allocate = true);
// libbsdl's energy tables in device memory (oslmi_layered.hip), uploaded on
// first use of a device unless allocate is false.
namespace mtx {
int device_tables(const float** out, bool allocate = true);
}

// Kernel argument block of one execution of group g (validates the
// arguments like BatchedExecutor::execute; `who` prefixes error messages).
// With allocate false, tables the group reads (colour system, transform
// table) must already be in device memory.
int make_group_args(const ShadingSys& ss, const Group& g, const oslmi_batched_globals* bg,
                    float* d_outputs, const float* d_wo, const float* d_wi, float* d_eval,
                    const char* who, GroupArgs& A, bool allocate = true);

// The no-alloc workspace (oslmi_noalloc.hip): one fixed device array per
// GPU, laid out as tile counters, argument blocks, tiles.
constexpr int kNoAllocMaxCounters   = 256;
constexpr int kNoAllocMaxItems      = 256;
constexpr int kNoAllocMaxTiles      = 32768;
constexpr size_t kNoAllocArgsOffset = 1024;
constexpr size_t kNoAllocArgsSlot    = 768;  // bytes per argument block
constexpr size_t kNoAllocTilesOffset = kNoAllocArgsOffset + kNoAllocMaxItems * kNoAllocArgsSlot;
constexpr size_t kNoAllocScratchBytes = 1 << 20;
// looks up this device's workspace address (with the device current)
int noalloc_scratch_init();
// the workspace of the current device (oslmi_init must have run on it)
int noalloc_scratch(unsigned char** out);
// zero-fills n tile counters on the stream
int noalloc_clear(int32_t* p, int n, hipStream_t stream);
}  // namespace oslmi

// Opaque handles of the C ABI.
struct oslmi_shadingsys {
    oslmi::ShadingSys ss;
};
struct oslmi_group {
    oslmi::Group* g;
};
