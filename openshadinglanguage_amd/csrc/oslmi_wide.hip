// Execution from the reference's own batch layout: BatchedShaderGlobals<16>
// blocks in, renderer outputs placed by symbol locations out.
//
// The reference's BatchedExecutor<16>::execute (src/include/OSL/oslexec.h:
// 1000-1004) shades one 16-lane batch per call from a
// BatchedShaderGlobals<16> (batched_shaderglobals.h:21-140) and writes the
// symbols a renderer placed with add_symlocs (oslexec.h:69-104, :1072-1076)
// to output_base + offset + shadeindex * stride. Here all batches of a
// frame go in one call: a gather kernel turns the 16-lane blocks into the
// structure-of-arrays planes the group kernel reads (only the globals the
// group uses), the group's emitted kernel runs unchanged, and a scatter
// kernel places each output symbol per point by its shade index.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "oslmi.h"
#include "oslmi_group.h"
#include "oslmi_groupargs.h"
#include "oslmi_internal.h"

namespace oslmi {

namespace {

constexpr int kMaxSymlocs = 16;

// byte offset of each global the emitted kernels read within one block
const size_t kBsgOffset[kG_COUNT] = {
    offsetof(oslmi_bsg16, P),    offsetof(oslmi_bsg16, dPdx), offsetof(oslmi_bsg16, dPdy),
    offsetof(oslmi_bsg16, N),    offsetof(oslmi_bsg16, Ng),   offsetof(oslmi_bsg16, I),
    offsetof(oslmi_bsg16, u),    offsetof(oslmi_bsg16, v),    offsetof(oslmi_bsg16, dudx),
    offsetof(oslmi_bsg16, dudy), offsetof(oslmi_bsg16, dvdx), offsetof(oslmi_bsg16, dvdy),
    offsetof(oslmi_bsg16, time), offsetof(oslmi_bsg16, dPdu), offsetof(oslmi_bsg16, dPdv),
    offsetof(oslmi_bsg16, Ps)
};
const int kGlobalComps[kG_COUNT] = { 3, 3, 3, 3, 3, 3, 1, 1, 1, 1, 1, 1, 1, 3, 3, 3 };

static_assert(sizeof(oslmi_bsg16) == 4032, "BatchedShaderGlobals<16> is 4032 bytes");
static_assert(offsetof(oslmi_bsg16, P) == 64, "VaryingShaderGlobals<16> follows a 64-byte block");
static_assert(offsetof(oslmi_bsg16, backfacing) == 3968, "layout of VaryingShaderGlobals<16>");

// One plane copy of the gather: component planes of one global.
struct GatherPlan {
    int nplanes;
    int32_t src_off[3 * kG_COUNT];  // byte offset of component 0 of lane 0 in a block
    float* dst[3 * kG_COUNT];
};

// point i = 16 b + lane; component c of a global sits at offset + 64 c +
// 4 lane in block b
__global__ __launch_bounds__(256) void k_bsg_gather(const unsigned char* __restrict__ bsg,
                                                    GatherPlan plan, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const unsigned char* blk = bsg + (i >> 4) * (int64_t)sizeof(oslmi_bsg16) + 4 * (i & 15);
    for (int p = 0; p < plan.nplanes; ++p)
        plan.dst[p][i] = *(const float*)(blk + plan.src_off[p]);
}

struct ScatterLoc {
    int64_t offset, stride;
    int first_plane, comps, is_int, pad;
};
struct ScatterPlan {
    int nlocs;
    ScatterLoc loc[kMaxSymlocs];
};

__global__ __launch_bounds__(256) void k_symloc_scatter(const float* __restrict__ planes,
                                                        const int32_t* __restrict__ shadeindex,
                                                        unsigned char* __restrict__ base,
                                                        ScatterPlan plan, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t si = shadeindex[i];
    for (int k = 0; k < plan.nlocs; ++k) {
        const ScatterLoc& L = plan.loc[k];
        unsigned char* dst  = base + L.offset + si * L.stride;
        for (int c = 0; c < L.comps; ++c) {
            const float v = planes[(int64_t)(L.first_plane + c) * n + i];
            if (L.is_int)
                ((int32_t*)dst)[c] = (int32_t)v;
            else
                ((float*)dst)[c] = v;
        }
    }
}

int type_comps(const std::string& t, int& comps, int& is_int)
{
    is_int = 0;
    if (t == "float")
        comps = 1;
    else if (t == "int")
        comps = 1, is_int = 1;
    else if (t == "color" || t == "point" || t == "vector" || t == "normal")
        comps = 3;
    else
        return -1;
    return 0;
}

}  // namespace

}  // namespace oslmi

using namespace oslmi;

extern "C" {

int oslmi_add_symlocs(oslmi_shadingsys* s, oslmi_group* gh, int n, const oslmi_symloc* locs)
{
    if (!s || !gh || n < 0 || (n > 0 && !locs))
        return set_error(OSLMI_ERR_INVALID, "oslmi_add_symlocs: bad arguments");
    Group& g = *gh->g;
    std::vector<SymLoc> add;
    for (int k = 0; k < n; ++k) {
        const oslmi_symloc& L = locs[k];
        if (!L.name || !L.type)
            return set_error(OSLMI_ERR_INVALID, "oslmi_add_symlocs: null name or type");
        if (L.arena != OSLMI_ARENA_OUTPUTS)
            return set_error(OSLMI_ERR_UNSUPPORTED,
                             std::string("oslmi_add_symlocs: ") + L.name
                                 + ": only the output arena is supported");
        if (L.derivs)
            return set_error(OSLMI_ERR_UNSUPPORTED,
                             std::string("oslmi_add_symlocs: ") + L.name + ": derivatives are not placed");
        int comps = 0, is_int = 0;
        if (type_comps(L.type, comps, is_int))
            return set_error(OSLMI_ERR_INVALID, std::string("oslmi_add_symlocs: ") + L.name
                                                    + ": unsupported type " + L.type);
        if (L.offset < 0 || L.stride < comps * 4)
            return set_error(OSLMI_ERR_INVALID, std::string("oslmi_add_symlocs: ") + L.name
                                                    + ": negative offset or stride below the value size");
        SymLoc sl;
        sl.name   = L.name;
        sl.comps  = comps;
        sl.is_int = is_int;
        sl.offset = L.offset;
        sl.stride = L.stride;
        add.push_back(sl);
    }
    std::lock_guard<std::mutex> lock(g.mu);
    if (g.symlocs.size() + add.size() > (size_t)kMaxSymlocs)
        return set_error(OSLMI_ERR_INVALID, "oslmi_add_symlocs: more than 16 symbol locations");
    for (const SymLoc& sl : add) {
        bool known = false;
        for (const std::string& o : g.output_names)
            known = known || o == sl.name;
        if (!known) {
            if (g.jitted)
                return set_error(OSLMI_ERR_INVALID, "oslmi_add_symlocs: " + sl.name
                                                        + " is not an output of the compiled group");
            g.output_names.push_back(sl.name);
        }
        g.symlocs.push_back(sl);
    }
    return OSLMI_OK;
}

int oslmi_clear_symlocs(oslmi_shadingsys* s, oslmi_group* gh)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_clear_symlocs: null argument");
    std::lock_guard<std::mutex> lock(gh->g->mu);
    gh->g->symlocs.clear();
    return OSLMI_OK;
}

int oslmi_execute_wide16(oslmi_shadingsys* s, oslmi_group* gh, int64_t nbatch, int last_batch_size,
                         const oslmi_bsg16* d_bsg, const int32_t* d_shadeindex,
                         void* d_output_base, void* stream)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_wide16: null argument");
    if (nbatch < 0 || (nbatch > 0 && (last_batch_size < 1 || last_batch_size > 16)))
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_wide16: nbatch < 0 or batch size not in 1..16");
    if (nbatch == 0)
        return OSLMI_OK;
    if (!d_bsg || !d_shadeindex)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_wide16: null globals or shade indices");
    Group& g = *gh->g;
    if (!g.jitted)
        if (int rc = oslmi_jit_group(s, gh))
            return rc;
    const int64_t n = (nbatch - 1) * 16 + last_batch_size;
    // symbol locations -> output planes
    ScatterPlan sp;
    memset(&sp, 0, sizeof(sp));
    for (const SymLoc& sl : g.symlocs) {
        const OutputSlot* slot = nullptr;
        for (const OutputSlot& o : g.outputs)
            if (o.name == sl.name)
                slot = &o;
        if (!slot)
            return set_error(OSLMI_ERR_INVALID, "oslmi_execute_wide16: symbol location " + sl.name
                                                    + " is not an output of the group");
        if (slot->planes != sl.comps)
            return set_error(OSLMI_ERR_INVALID, "oslmi_execute_wide16: " + sl.name
                                                    + ": the symbol location's type does not match the symbol");
        ScatterLoc& L  = sp.loc[sp.nlocs++];
        L.offset       = sl.offset;
        L.stride       = sl.stride;
        L.first_plane  = slot->first_plane;
        L.comps        = sl.comps;
        L.is_int       = sl.is_int;
    }
    if (sp.nlocs > 0 && !d_output_base)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_wide16: symbol locations but no output base");
    // workspace: gathered globals, then the output planes
    int gplanes = 0;
    for (unsigned k : g.globals_used)
        gplanes += kGlobalComps[k];
    const size_t need = (size_t)n * (size_t)(gplanes + g.output_planes);
    if (need > g.wide_ws_floats) {
        if (g.wide_ws)
            (void)hipFree(g.wide_ws);
        g.wide_ws        = nullptr;
        g.wide_ws_floats = 0;
        if (int rc = check_hip(hipMalloc((void**)&g.wide_ws, need * sizeof(float)),
                               "hipMalloc(wide workspace)"))
            return rc;
        g.wide_ws_floats = need;
    }
    oslmi_batched_globals bg;
    memset(&bg, 0, sizeof(bg));
    bg.n = n;
    GatherPlan gp;
    memset(&gp, 0, sizeof(gp));
    float* next = g.wide_ws;
    for (unsigned k : g.globals_used) {
        bg.planes[k] = next;
        for (int c = 0; c < kGlobalComps[k]; ++c) {
            gp.src_off[gp.nplanes] = (int32_t)(kBsgOffset[k] + 64 * c);
            gp.dst[gp.nplanes++]   = next;
            next += n;
        }
    }
    float* outputs = next;
    hipStream_t st = (hipStream_t)stream;
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (gp.nplanes > 0) {
        hipLaunchKernelGGL(k_bsg_gather, dim3(grid), dim3(256), 0, st,
                           (const unsigned char*)d_bsg, gp, n);
        if (int rc = check_launch("oslmi_execute_wide16 (gather)"))
            return rc;
    }
    GroupArgs A;
    if (int rc = make_group_args(s->ss, g, &bg, outputs, nullptr, nullptr, nullptr,
                                 "oslmi_execute_wide16", A))
        return rc;
    void* args[] = { &A };
    if (int rc = check_hip(hipModuleLaunchKernel(g.kernel, grid_for(n), 1, 1, 256, 1, 1, 0, st, args,
                                                 nullptr),
                           "hipModuleLaunchKernel"))
        return rc;
    if (sp.nlocs > 0) {
        hipLaunchKernelGGL(k_symloc_scatter, dim3(grid), dim3(256), 0, st, outputs, d_shadeindex,
                           (unsigned char*)d_output_base, sp, n);
        if (int rc = check_launch("oslmi_execute_wide16 (scatter)"))
            return rc;
    }
    return OSLMI_OK;
}

}  // extern "C"
