// C++ facade over the C ABI (include/oslmi.h) in the shape of the
// reference's batched API, for a renderer that today builds groups with
// OSL::ShadingSystem and shades with OSL::ShadingSystem::BatchedExecutor<16>
// (src/include/OSL/oslexec.h:634-742, :980-1027). Header only; link with
// liboslmi.so and the HIP runtime.
//
//   ShadingSystem        LoadMemoryCompiledShader, ShaderGroupBegin,
//                        Parameter, Shader, ConnectShaders, ShaderGroupEnd,
//                        add_symlocs / clear_symlocs, name_transform
//   BatchedShaderGlobals16  BatchedShaderGlobals<16> in the reference's
//                        layout (oslmi_bsg16)
//   BatchedExecutor16    execute(group, batch_size, wide_shadeindex, bsg,
//                        userdata, output): one 16-lane batch in host
//                        memory, as the reference's call; execute_frame():
//                        every batch of a frame in device memory, one call
//
// Errors are reported by exceptions carrying oslmi_last_error(); the
// reference returns bool and reports through its ErrorHandler.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "oslmi.h"

namespace oslmi_adapter {

using BatchedShaderGlobals16 = oslmi_bsg16;

inline void check(int rc, const char* what)
{
    if (rc != OSLMI_OK)
        throw std::runtime_error(std::string(what) + ": " + oslmi_last_error());
}
inline void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// SymLocationDesc (oslexec.h:69-104) in the output arena
struct SymLocationDesc {
    std::string name;
    std::string type;  // "float", "int", "color", "point", "vector", "normal"
    int64_t offset;
    int64_t stride;
};

class ShadingSystem;

class ShaderGroup {
public:
    oslmi_group* handle() const { return g_; }

private:
    friend class ShadingSystem;
    oslmi_group* g_ = nullptr;
};

class ShadingSystem {
public:
    // csrc_dir: the library's device headers (openshadinglanguage_amd/csrc);
    // jit_cache: code-object cache directory or "" for none
    ShadingSystem(const std::string& csrc_dir, const std::string& jit_cache)
        : ss_(oslmi_shadingsys_create(csrc_dir.c_str(), jit_cache.empty() ? nullptr : jit_cache.c_str()))
    {
        if (!ss_)
            throw std::runtime_error(std::string("oslmi_shadingsys_create: ") + oslmi_last_error());
        check(oslmi_init(), "oslmi_init");
    }
    ~ShadingSystem()
    {
        for (oslmi_group* g : groups_)
            oslmi_group_release(g);
        oslmi_shadingsys_destroy(ss_);
    }
    ShadingSystem(const ShadingSystem&) = delete;
    ShadingSystem& operator=(const ShadingSystem&) = delete;

    oslmi_shadingsys* handle() const { return ss_; }

    bool attribute(const std::string& name, const std::string& value)
    {
        return oslmi_shadingsys_attribute(ss_, name.c_str(), value.c_str()) == OSLMI_OK;
    }
    void LoadMemoryCompiledShader(const std::string& name, const std::string& oso_text)
    {
        check(oslmi_load_memory_compiled_shader(ss_, name.c_str(), oso_text.c_str()),
              "LoadMemoryCompiledShader");
    }
    ShaderGroup ShaderGroupBegin(const std::string& name)
    {
        ShaderGroup r;
        r.g_ = oslmi_shader_group_begin(ss_, name.c_str());
        if (!r.g_)
            throw std::runtime_error(std::string("ShaderGroupBegin: ") + oslmi_last_error());
        groups_.push_back(r.g_);
        return r;
    }
    // type: "int", "float", "color", "point", "vector", "normal", "matrix", "string"
    void Parameter(ShaderGroup& g, const std::string& name, const std::string& type,
                   const void* value, int count = 1)
    {
        check(oslmi_parameter(ss_, g.g_, name.c_str(), type.c_str(), value, count), "Parameter");
    }
    void Shader(ShaderGroup& g, const std::string& usage, const std::string& shadername,
                const std::string& layername)
    {
        check(oslmi_shader(ss_, g.g_, usage.c_str(), shadername.c_str(), layername.c_str()),
              "Shader");
    }
    void ConnectShaders(ShaderGroup& g, const std::string& srclayer, const std::string& srcparam,
                        const std::string& dstlayer, const std::string& dstparam)
    {
        check(oslmi_connect_shaders(ss_, g.g_, srclayer.c_str(), srcparam.c_str(),
                                    dstlayer.c_str(), dstparam.c_str()),
              "ConnectShaders");
    }
    void ShaderGroupEnd(ShaderGroup& g) { check(oslmi_shader_group_end(ss_, g.g_), "ShaderGroupEnd"); }
    void add_symlocs(ShaderGroup& g, const std::vector<SymLocationDesc>& locs)
    {
        std::vector<oslmi_symloc> c(locs.size());
        for (size_t k = 0; k < locs.size(); ++k)
            c[k] = oslmi_symloc{ locs[k].name.c_str(), locs[k].type.c_str(), OSLMI_ARENA_OUTPUTS, 0,
                                 locs[k].offset, locs[k].stride };
        check(oslmi_add_symlocs(ss_, g.g_, (int)c.size(), c.data()), "add_symlocs");
    }
    void clear_symlocs(ShaderGroup& g) { check(oslmi_clear_symlocs(ss_, g.g_), "clear_symlocs"); }
    void name_transform(const std::string& name, const float m[16], const float* inv = nullptr)
    {
        check(oslmi_renderer_name_transform(ss_, name.c_str(), m, inv), "name_transform");
    }
    // BatchedExecutor::jit_group (oslexec.h:993-998)
    void jit_group(ShaderGroup& g) { check(oslmi_jit_group(ss_, g.g_), "jit_group"); }

private:
    oslmi_shadingsys* ss_;
    std::vector<oslmi_group*> groups_;
};

// BatchedExecutor<16> (oslexec.h:980-1027)
class BatchedExecutor16 {
public:
    explicit BatchedExecutor16(ShadingSystem& ss) : ss_(ss) {}
    ~BatchedExecutor16()
    {
        (void)hipFree(d_bsg_);
        (void)hipFree(d_index_);
        (void)hipFree(d_out_);
    }

    // One batch from host memory, the reference's call: batch_size lanes of
    // bsg, shade indices wide_shadeindex[0..batch_size), outputs placed by
    // the group's symlocs relative to output_base (host). userdata is not
    // used (no userdata arena). Synchronous, like the CPU call it replaces.
    bool execute(ShaderGroup& g, int batch_size, const int* wide_shadeindex,
                 const BatchedShaderGlobals16& bsg, void* /*userdata*/, void* output_base,
                 size_t output_bytes)
    {
        if (batch_size < 1 || batch_size > 16)
            return false;
        // the device copy of the renderer's output block, so bytes the
        // shader does not write keep their values
        grow((void**)&d_bsg_, bsg_cap_, sizeof(BatchedShaderGlobals16));
        grow((void**)&d_index_, index_cap_, 16 * sizeof(int));
        grow((void**)&d_out_, out_cap_, output_bytes);
        int idx[16];
        for (int k = 0; k < 16; ++k)
            idx[k] = wide_shadeindex[k < batch_size ? k : 0];
        check_hip(hipMemcpy(d_bsg_, &bsg, sizeof(bsg), hipMemcpyHostToDevice), "hipMemcpy(bsg)");
        check_hip(hipMemcpy(d_index_, idx, sizeof(idx), hipMemcpyHostToDevice), "hipMemcpy(index)");
        check_hip(hipMemcpy(d_out_, output_base, output_bytes, hipMemcpyHostToDevice), "hipMemcpy(out)");
        int rc = oslmi_execute_wide16(ss_.handle(), g.handle(), 1, batch_size, d_bsg_, d_index_,
                                      d_out_, nullptr);
        if (rc != OSLMI_OK)
            return false;
        check_hip(hipMemcpy(output_base, d_out_, output_bytes, hipMemcpyDeviceToHost), "hipMemcpy(out)");
        return true;
    }

    // Every batch of a frame at once, all in device memory: nbatch blocks,
    // the last holding last_batch_size lanes; nbatch * 16 shade indices.
    void execute_frame(ShaderGroup& g, int64_t nbatch, int last_batch_size,
                       const BatchedShaderGlobals16* d_bsg, const int32_t* d_shadeindex,
                       void* d_output_base, hipStream_t stream = nullptr)
    {
        check(oslmi_execute_wide16(ss_.handle(), g.handle(), nbatch, last_batch_size, d_bsg,
                                   d_shadeindex, d_output_base, (void*)stream),
              "execute_frame");
    }

private:
    static void grow(void** p, size_t& cap, size_t bytes)
    {
        if (bytes <= cap)
            return;
        (void)hipFree(*p);
        *p = nullptr;
        cap = 0;
        check_hip(hipMalloc(p, bytes), "hipMalloc");
        cap = bytes;
    }
    ShadingSystem& ss_;
    BatchedShaderGlobals16* d_bsg_ = nullptr;
    int* d_index_                  = nullptr;
    void* d_out_                   = nullptr;
    size_t bsg_cap_ = 0, index_cap_ = 0, out_cap_ = 0;
};

}  // namespace oslmi_adapter
