// ShadingSystem / ShaderGroup / BatchedExecutor behind the C ABI.
//
//   oslmi_shadingsys_create        ShadingSystem::ShadingSystem (oslexec.h:170)
//   oslmi_shadingsys_attribute     ShadingSystem::attribute (oslexec.h:177-324)
//   oslmi_load_memory_compiled_shader  LoadMemoryCompiledShader (oslexec.h)
//   oslmi_shader_group_begin/end   ShaderGroupBegin / ShaderGroupEnd (:634-742)
//   oslmi_parameter                Parameter(group, name, TypeDesc, val)
//   oslmi_shader                   Shader(group, usage, shadername, layername)
//   oslmi_connect_shaders          ConnectShaders(group, srclayer, srcparam, ...)
//   oslmi_group_outputs            attribute(group, "renderer_outputs", ...)
//   oslmi_jit_group                BatchedExecutor<W>::jit_group (:993-998)
//   oslmi_execute_batch            BatchedExecutor<W>::execute (:1000-1004)
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

#include "oslmi.h"
#include "oslmi_color.h"
#include "oslmi_colorsys.h"
#include "oslmi_group.h"
#include "oslmi_internal.h"

namespace oslmi {


namespace {

bool parse_type(const char* t, OType& out)
{
    std::string s(t ? t : "");
    if (s == "int") out.base = kInt;
    else if (s == "float") out.base = kFloat;
    else if (s == "color" || s == "point" || s == "vector" || s == "normal") {
        out.base        = kTriple;
        out.triple_kind = s;
    } else if (s == "string") out.base = kString;
    else if (s == "matrix") out.base = kMatrix;
    else return false;
    return true;
}

std::string hex64(uint64_t h)
{
    char b[17];
    snprintf(b, sizeof(b), "%016llx", (unsigned long long)h);
    return b;
}

uint64_t fnv1a(const std::string& s)
{
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

bool read_file(const std::string& path, std::vector<char>& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !out.empty();
}

void mkdirs(const std::string& d)
{
    std::string cur;
    for (size_t i = 0; i < d.size(); ++i) {
        cur += d[i];
        if (d[i] == '/' || i + 1 == d.size())
            mkdir(cur.c_str(), 0755);
    }
}

const char* kCompileOpts[] = { "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops" };

}  // namespace

int build_code_object(const ShadingSys& ss, const std::string& src, const std::string& name,
                      std::vector<char>& code, double* seconds)
{
    std::string opts_key;
    for (const char* o : kCompileOpts)
        opts_key += std::string(o) + " ";
    // include the device library headers in the cache key
    std::string hdrs;
    for (const char* h : { "oslmi_rt.h", "oslmi_types.h", "oslmi_noise.h", "oslmi_gabor.h",
                           "oslmi_math.h", "oslmi_color.h", "oslmi_closure.h", "oslmi_groupargs.h",
                           "oslmi_matrix.h", "oslmi_mtx.h" }) {
        std::vector<char> c;
        read_file(ss.include_dir + "/" + h, c);
        hdrs.append(c.begin(), c.end());
    }
    std::string key = hex64(fnv1a(src + "|" + opts_key + "|" + hdrs));
    std::string cache_path = ss.cache_dir.empty() ? "" : ss.cache_dir + "/" + key + ".hsaco";
    auto t0 = std::chrono::steady_clock::now();
    if (cache_path.empty() || !read_file(cache_path, code)) {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr)
            != HIPRTC_SUCCESS)
            return set_error(OSLMI_ERR_COMPILE, "hiprtcCreateProgram failed");
        std::vector<const char*> opts(std::begin(kCompileOpts), std::end(kCompileOpts));
        std::string inc = "-I" + ss.include_dir;
        opts.push_back(inc.c_str());
        hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
        size_t logsz    = 0;
        hiprtcGetProgramLogSize(prog, &logsz);
        std::string log(logsz, '\0');
        if (logsz)
            hiprtcGetProgramLog(prog, &log[0]);
        if (rc != HIPRTC_SUCCESS) {
            hiprtcDestroyProgram(&prog);
            return set_error(OSLMI_ERR_COMPILE, "hiprtc compile of " + name + " failed:\n" + log);
        }
        size_t sz = 0;
        hiprtcGetCodeSize(prog, &sz);
        code.resize(sz);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        if (!cache_path.empty()) {
            mkdirs(ss.cache_dir);
            // unique per process and per call (compiles may run in threads)
            static std::atomic<long> seq { 0 };
            std::string tmp = cache_path + ".tmp" + std::to_string((long)getpid()) + "_"
                              + std::to_string(seq.fetch_add(1));
            std::ofstream f(tmp, std::ios::binary);
            f.write(code.data(), (std::streamsize)code.size());
            f.close();
            rename(tmp.c_str(), cache_path.c_str());
        }
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return OSLMI_OK;
}

int make_group_args(const ShadingSys& ss, const Group& g, const oslmi_batched_globals* bg,
                    float* d_outputs, const float* d_wo, const float* d_wi, float* d_eval,
                    const char* who, GroupArgs& A, bool allocate)
{
    std::string w(who);
    if (bg->n < 0)
        return set_error(OSLMI_ERR_INVALID, w + ": n < 0");
    if (bg->n > 0 && g.output_planes > 0 && !d_outputs)
        return set_error(OSLMI_ERR_INVALID, w + ": group has outputs but d_outputs is NULL");
    if (d_eval && (!g.surface_closure || !d_wo || !d_wi))
        return set_error(OSLMI_ERR_INVALID,
                         w + ": closure evaluation needs a surface closure and wo/wi");
    memset(&A, 0, sizeof(A));
    A.n = bg->n;
    for (int k = 0; k < kG_COUNT; ++k) {
        A.g[k] = bg->planes[k];
        for (int c = 0; c < 3; ++c)
            A.gu[k][c] = bg->uniform[k][c];
    }
    for (int p = 0; p < g.output_planes; ++p)
        A.out[p] = d_outputs + (int64_t)p * bg->n;
    A.wo   = d_wo;
    A.wi   = d_wi;
    A.eval = d_eval;
    if (g.uses_colorsys) {
        const ColorSysDev* cs = nullptr;
        if (int rc = device_colorsys(&cs, allocate))
            return rc;
        A.colorsys = cs;
    }
    if (g.uses_mx) {
        const float* lut = nullptr;
        if (int rc = mtx::device_tables(&lut, allocate))
            return rc;
        A.bsdl = lut;
    }
    if (g.uses_xforms) {
        if (!allocate && !ss.xf.device)
            return set_error(OSLMI_ERR_INVALID,
                             w + ": the transform table is not yet in device memory (run one "
                                 "oslmi_execute_batch or oslmi_batch_bind of the shading system "
                                 "first)");
        if (int rc = xform_table_sync(ss))
            return rc;
        A.xf = ss.xf.device;
    }
    return OSLMI_OK;
}

// ---- renderer services: named coordinate systems ---------------------------
namespace {

bool is_common(const ShadingSys& ss, const std::string& n)
{
    return n == "common" || n == ss.commonspace;
}
// RendererServices::get_matrix(from): to-common matrix; identity and false
// for an unknown space (wide_opmatrix.cpp:533-568)
bool get_matrix(const ShadingSys& ss, const std::string& from, M44& M)
{
    M = m44_identity();
    if (is_common(ss, from))
        return true;
    auto it = ss.xf.named.find(from);
    if (it == ss.xf.named.end())
        return false;
    memcpy(M.m, it->second.m, sizeof(M.m));
    return true;
}
// get_inverse_matrix(to): the renderer's inverse, or the matrix inverted as
// the default implementation does (wide_opmatrix.cpp:134-150, 573-610)
bool get_inverse_matrix(const ShadingSys& ss, const std::string& to, M44& M)
{
    M = m44_identity();
    if (is_common(ss, to))
        return true;
    auto it = ss.xf.named.find(to);
    if (it == ss.xf.named.end())
        return false;
    if (it->second.has_inv) {
        memcpy(M.m, it->second.inv, sizeof(M.m));
    } else {
        M44 f;
        memcpy(f.m, it->second.m, sizeof(f.m));
        M = m44_inverse(f);
    }
    return true;
}
// the from-to matrix of osl_transform_triple / osl_get_from_to_matrix
// (src/liboslexec/opmatrix.cpp:232-262)
bool from_to_matrix(const ShadingSys& ss, const std::string& from, const std::string& to, M44& M)
{
    M44 mf, mt;
    if (is_common(ss, from))
        return get_inverse_matrix(ss, to, M);
    if (is_common(ss, to))
        return get_matrix(ss, from, M);
    bool ok = get_matrix(ss, from, mf);
    ok      = get_inverse_matrix(ss, to, mt) && ok;
    M       = m44_mul(mf, mt);
    return ok;
}

}  // namespace

int xform_slot(const ShadingSys& ss, const std::string& from, const std::string& to)
{
    std::lock_guard<std::mutex> lock(ss.xf.mu);
    for (size_t k = 0; k < ss.xf.pairs.size(); ++k)
        if (ss.xf.pairs[k].first == from && ss.xf.pairs[k].second == to)
            return (int)k;
    if ((int)ss.xf.pairs.size() >= kMaxXformSlots)
        return -1;
    ss.xf.pairs.emplace_back(from, to);
    ss.xf.dirty = true;
    return (int)ss.xf.pairs.size() - 1;
}

int attr_slot(const ShadingSys& ss, const std::string& name)
{
    std::lock_guard<std::mutex> lock(ss.xf.mu);
    size_t k = 0;
    for (; k < ss.xf.attr_names.size(); ++k)
        if (ss.xf.attr_names[k] == name)
            break;
    if (k == ss.xf.attr_names.size()) {
        if ((int)k >= kMaxAttrSlots)
            return -1;
        ss.xf.attr_names.push_back(name);
        ss.xf.dirty = true;
    }
    return kXformSlotFloats * kMaxXformSlots + kAttrSlotFloats * (int)k;
}

int xform_table_sync(const ShadingSys& ss)
{
    std::lock_guard<std::mutex> lock(ss.xf.mu);
    if (!ss.xf.dirty && ss.xf.device)
        return OSLMI_OK;
    if (!ss.xf.device) {
        if (int rc = check_hip(hipMalloc((void**)&ss.xf.device,
                                         sizeof(float) * (kXformSlotFloats * kMaxXformSlots
                                                          + kAttrSlotFloats * kMaxAttrSlots)),
                               "hipMalloc(transform table)"))
            return rc;
    }
    std::vector<float> host(kXformSlotFloats * ss.xf.pairs.size());
    for (size_t k = 0; k < ss.xf.pairs.size(); ++k) {
        M44 M;
        const bool ok = from_to_matrix(ss, ss.xf.pairs[k].first, ss.xf.pairs[k].second, M);
        const M44 N   = m44_transposed(m44_inverse(M));
        float* slot   = host.data() + kXformSlotFloats * k;
        memcpy(slot, M.m, sizeof(M.m));
        memcpy(slot + 16, N.m, sizeof(N.m));
        slot[32] = ok ? 1.0f : 0.0f;
    }
    // kernels of earlier launches may still read the table
    if (int rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return rc;
    if (!host.empty()) {
        if (int rc = check_hip(hipMemcpy(ss.xf.device, host.data(), sizeof(float) * host.size(),
                                         hipMemcpyHostToDevice),
                               "hipMemcpy(transform table)"))
            return rc;
    }
    std::vector<float> attrs(kAttrSlotFloats * ss.xf.attr_names.size(), 0.0f);
    for (size_t k = 0; k < ss.xf.attr_names.size(); ++k) {
        auto it = ss.xf.attrs.find(ss.xf.attr_names[k]);
        if (it == ss.xf.attrs.end())
            continue;
        float* slot = attrs.data() + kAttrSlotFloats * k;
        memcpy(slot, it->second.v, sizeof(it->second.v));
        slot[16] = (float)(it->second.base + 1);
    }
    if (!attrs.empty()) {
        if (int rc = check_hip(hipMemcpy(ss.xf.device + kXformSlotFloats * kMaxXformSlots,
                                         attrs.data(), sizeof(float) * attrs.size(),
                                         hipMemcpyHostToDevice),
                               "hipMemcpy(attribute table)"))
            return rc;
    }
    ss.xf.dirty = false;
    return OSLMI_OK;
}

namespace {

int compile_group(ShadingSys& ss, Group& g, bool load = true)
{
    std::string src, err;
    if (!emit_group(ss, g, src, err))
        return set_error(OSLMI_ERR_COMPILE, "emit " + g.name + ": " + err);
    g.source = src;
    std::vector<char> code;
    if (int rc = build_code_object(ss, src, g.name, code, &g.jit_seconds))
        return rc;
    if (!load)
        return OSLMI_OK;
    if (g.module) {
        (void)hipModuleUnload(g.module);
        g.module = nullptr;
    }
    if (int rc = check_hip(hipModuleLoadData(&g.module, code.data()), "hipModuleLoadData"))
        return rc;
    if (int rc = check_hip(hipModuleGetFunction(&g.kernel, g.module, "oslmi_group_main"),
                           "hipModuleGetFunction"))
        return rc;
    g.jitted = true;
    return OSLMI_OK;
}

// Compiles without a device: emits and runs hiprtc, filling the cache.
int precompile_group(ShadingSys& ss, Group& g)
{
    std::string src, err;
    if (!emit_group(ss, g, src, err))
        return set_error(OSLMI_ERR_COMPILE, "emit " + g.name + ": " + err);
    g.source = src;
    return OSLMI_OK;
}

}  // namespace
}  // namespace oslmi

using namespace oslmi;

extern "C" {

oslmi_shadingsys* oslmi_shadingsys_create(const char* include_dir, const char* cache_dir)
{
    auto* s           = new oslmi_shadingsys;
    s->ss.include_dir = include_dir ? include_dir : "";
    s->ss.cache_dir   = cache_dir ? cache_dir : "";
    return s;
}

void oslmi_shadingsys_destroy(oslmi_shadingsys* s)
{
    if (!s)
        return;
    for (auto& g : s->ss.groups) {
        if (g->module)
            (void)hipModuleUnload(g->module);
        if (g->wide_ws)
            (void)hipFree(g->wide_ws);
    }
    if (s->ss.xf.device)
        (void)hipFree(s->ss.xf.device);
    delete s;
}

int oslmi_shadingsys_attribute(oslmi_shadingsys* s, const char* name, const char* value)
{
    if (!s || !name || !value)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shadingsys_attribute: null argument");
    std::string n(name);
    if (n == "colorspace") {
        if (std::string(value) != "Rec709")
            return set_error(OSLMI_ERR_UNSUPPORTED, "only the Rec709 color space is built in");
        s->ss.colorspace = value;
    } else if (n == "jit_cache") {
        s->ss.cache_dir = value;
    } else if (n == "include_dir") {
        s->ss.include_dir = value;
    } else if (n == "commonspace") {
        // the name that is a synonym of "common" (ShadingSystem attribute
        // "commonspace", default "world")
        if (!*value)
            return set_error(OSLMI_ERR_INVALID, "commonspace must be a name");
        std::lock_guard<std::mutex> lock(s->ss.xf.mu);
        s->ss.commonspace = value;
        s->ss.xf.dirty    = true;
    } else if (n == "closure_sample") {
        // surface kernels JIT-compiled from now on also carry
        // CompositeBSDF::sample (oslmi_execute_batch_sample)
        std::string v(value);
        if (v != "0" && v != "1")
            return set_error(OSLMI_ERR_INVALID, "closure_sample must be 0 or 1");
        s->ss.closure_sample = v == "1";
    } else if (n == "closure_export") {
        // surface kernels JIT-compiled from now on also export Ci's lobes
        // (oslmi_execute_batch_closures)
        std::string v(value);
        if (v != "0" && v != "1")
            return set_error(OSLMI_ERR_INVALID, "closure_export must be 0 or 1");
        s->ss.closure_export = v == "1";
    } else if (n == "gpu_waves_per_eu") {
        // occupancy target (waves per SIMD) the emitted kernels are
        // register-allocated for; 0 leaves it to the compiler
        char* end = nullptr;
        long w    = strtol(value, &end, 10);
        if (!end || end == value || *end || w < 0 || w > 8)
            return set_error(OSLMI_ERR_INVALID, "gpu_waves_per_eu must be 0..8");
        s->ss.waves_per_eu = (int)w;
    } else {
        return set_error(OSLMI_ERR_UNSUPPORTED, "unknown attribute " + n);
    }
    return OSLMI_OK;
}

int oslmi_renderer_name_transform(oslmi_shadingsys* s, const char* name, const float* m,
                                  const float* inv)
{
    if (!s || !name || !m)
        return set_error(OSLMI_ERR_INVALID, "oslmi_renderer_name_transform: null argument");
    std::string n(name);
    if (n.empty() || n == "common" || n == s->ss.commonspace)
        return set_error(OSLMI_ERR_INVALID,
                         "oslmi_renderer_name_transform: \"" + n + "\" is the common space");
    std::lock_guard<std::mutex> lock(s->ss.xf.mu);
    NamedXform& x = s->ss.xf.named[n];
    memcpy(x.m, m, sizeof(x.m));
    x.has_inv = inv != nullptr;
    if (inv)
        memcpy(x.inv, inv, sizeof(x.inv));
    s->ss.xf.dirty = true;
    return OSLMI_OK;
}

int oslmi_renderer_set_attribute(oslmi_shadingsys* s, const char* name, const char* type,
                                 const void* value)
{
    if (!s || !name || !type || !value || !*name)
        return set_error(OSLMI_ERR_INVALID, "oslmi_renderer_set_attribute: null argument");
    OType t;
    if (!parse_type(type, t) || t.base == kString)
        return set_error(OSLMI_ERR_UNSUPPORTED, std::string("oslmi_renderer_set_attribute: type ")
                                                    + type + " is not supported");
    RendererAttr a;
    a.base = t.base;
    const int count = t.base == kMatrix ? 16 : (t.base == kTriple ? 3 : 1);
    memcpy(a.v, value, sizeof(float) * count);  // an int keeps its bits in v[0]
    std::lock_guard<std::mutex> lock(s->ss.xf.mu);
    s->ss.xf.attrs[name] = a;
    s->ss.xf.dirty       = true;
    return OSLMI_OK;
}

int oslmi_load_memory_compiled_shader(oslmi_shadingsys* s, const char* shadername,
                                      const char* oso_text)
{
    if (!s || !shadername || !oso_text)
        return set_error(OSLMI_ERR_INVALID, "oslmi_load_memory_compiled_shader: null argument");
    auto m = std::make_unique<ShaderMaster>();
    std::string err;
    if (!parse_oso(oso_text, *m, err))
        return set_error(OSLMI_ERR_INVALID, std::string(shadername) + ": " + err);
    std::lock_guard<std::mutex> lock(s->ss.mu);
    s->ss.masters[shadername] = std::move(m);
    return OSLMI_OK;
}

oslmi_group* oslmi_shader_group_begin(oslmi_shadingsys* s, const char* groupname)
{
    if (!s)
        return nullptr;
    std::lock_guard<std::mutex> lock(s->ss.mu);
    s->ss.groups.push_back(std::make_unique<Group>());
    Group* g = s->ss.groups.back().get();
    g->name  = groupname ? groupname : "group";
    return new oslmi_group { g };
}

int oslmi_parameter(oslmi_shadingsys* s, oslmi_group* gh, const char* name, const char* type,
                    const void* value, int count)
{
    if (!s || !gh || !name || !type || !value || count < 1)
        return set_error(OSLMI_ERR_INVALID, "oslmi_parameter: bad argument");
    Group& g = *gh->g;
    if (g.ended)
        return set_error(OSLMI_ERR_INVALID, "oslmi_parameter: group already ended");
    ParamValue pv;
    if (!parse_type(type, pv.type))
        return set_error(OSLMI_ERR_UNSUPPORTED, std::string("oslmi_parameter: type ") + type);
    int ncomp = pv.type.base == kTriple ? 3 : pv.type.base == kMatrix ? 16 : 1;
    if (pv.type.base == kString) {
        const char* const* sv = (const char* const*)value;
        for (int i = 0; i < count; ++i)
            pv.s.push_back(sv[i]);
    } else if (pv.type.base == kInt) {
        const int* iv = (const int*)value;
        for (int i = 0; i < count; ++i)
            pv.i.push_back(iv[i]);
    } else {
        const float* fv = (const float*)value;
        for (int i = 0; i < count * ncomp; ++i)
            pv.f.push_back(fv[i]);
    }
    // A parameter applies to the next Shader() call (ShadingSystem
    // semantics): values collect on a trailing layer that has no shader yet.
    if (g.layers.empty() || g.layers.back().master != nullptr)
        g.layers.push_back(Layer());
    g.layers.back().inst[name] = pv;
    return OSLMI_OK;
}

int oslmi_shader(oslmi_shadingsys* s, oslmi_group* gh, const char* usage, const char* shadername,
                 const char* layername)
{
    if (!s || !gh || !shadername)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shader: null argument");
    Group& g = *gh->g;
    auto it  = s->ss.masters.find(shadername);
    if (it == s->ss.masters.end())
        return set_error(OSLMI_ERR_INVALID, std::string("oslmi_shader: unknown shader ") + shadername);
    (void)usage;
    if (g.layers.empty() || g.layers.back().master != nullptr)
        g.layers.push_back(Layer());
    Layer& L = g.layers.back();
    L.master = it->second.get();
    L.name   = layername && *layername ? layername : ("layer" + std::to_string(g.layers.size() - 1));
    for (auto& kv : L.inst)
        if (L.master->find(kv.first) < 0)
            return set_error(OSLMI_ERR_INVALID, "oslmi_shader: shader " + std::string(shadername)
                                                    + " has no parameter " + kv.first);
    return OSLMI_OK;
}

int oslmi_connect_shaders(oslmi_shadingsys* s, oslmi_group* gh, const char* srclayer,
                          const char* srcparam, const char* dstlayer, const char* dstparam)
{
    if (!s || !gh || !srclayer || !srcparam || !dstlayer || !dstparam)
        return set_error(OSLMI_ERR_INVALID, "oslmi_connect_shaders: null argument");
    Group& g = *gh->g;
    int sl = g.find_layer(srclayer), dl = g.find_layer(dstlayer);
    if (sl < 0 || dl < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_connect_shaders: unknown layer");
    int ss_ = g.layers[sl].master->find(srcparam), ds = g.layers[dl].master->find(dstparam);
    if (ss_ < 0 || ds < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_connect_shaders: unknown parameter");
    if (sl >= dl)
        return set_error(OSLMI_ERR_INVALID, "oslmi_connect_shaders: source layer must come first");
    g.conns.push_back({ sl, ss_, dl, ds });
    return OSLMI_OK;
}

int oslmi_shader_group_end(oslmi_shadingsys* s, oslmi_group* gh)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shader_group_end: null argument");
    Group& g = *gh->g;
    if (!g.layers.empty() && g.layers.back().master == nullptr)
        return set_error(OSLMI_ERR_INVALID, "parameters given after the last shader");
    g.ended = true;
    return OSLMI_OK;
}

int oslmi_group_outputs(oslmi_shadingsys* s, oslmi_group* gh, int n, const char* const* names)
{
    if (!s || !gh || n < 0 || (n > 0 && !names))
        return set_error(OSLMI_ERR_INVALID, "oslmi_group_outputs: bad argument");
    gh->g->output_names.assign(names, names + n);
    gh->g->jitted = false;
    return OSLMI_OK;
}

int oslmi_jit_group(oslmi_shadingsys* s, oslmi_group* gh)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_jit_group: null argument");
    Group& g = *gh->g;
    std::lock_guard<std::mutex> lock(g.mu);
    if (!g.ended)
        return set_error(OSLMI_ERR_INVALID, "oslmi_jit_group: group not ended");
    return compile_group(s->ss, g);
}

int oslmi_compile_group(oslmi_shadingsys* s, oslmi_group* gh)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_compile_group: null argument");
    Group& g = *gh->g;
    std::lock_guard<std::mutex> lock(g.mu);
    if (!g.ended)
        return set_error(OSLMI_ERR_INVALID, "oslmi_compile_group: group not ended");
    return compile_group(s->ss, g, false);
}

int oslmi_group_emit(oslmi_shadingsys* s, oslmi_group* gh, char* buf, int64_t bufsize)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_group_emit: null argument");
    Group& g = *gh->g;
    std::lock_guard<std::mutex> lock(g.mu);
    if (int rc = precompile_group(s->ss, g))
        return rc;
    if (buf && bufsize > 0) {
        size_t n = std::min((size_t)bufsize - 1, g.source.size());
        memcpy(buf, g.source.data(), n);
        buf[n] = 0;
    }
    return (int)std::min<size_t>(g.source.size(), 0x7fffffff);
}

int oslmi_group_output_planes(oslmi_shadingsys* s, oslmi_group* gh)
{
    if (!s || !gh)
        return set_error(OSLMI_ERR_INVALID, "oslmi_group_output_planes: null argument");
    return gh->g->jitted ? gh->g->output_planes : set_error(OSLMI_ERR_INVALID, "group not jitted");
}

double oslmi_group_jit_seconds(oslmi_group* gh) { return gh ? gh->g->jit_seconds : 0.0; }

void oslmi_group_release(oslmi_group* gh) { delete gh; }

int oslmi_execute_batch(oslmi_shadingsys* s, oslmi_group* gh, const oslmi_batched_globals* bg,
                        float* d_outputs, const float* d_wo, const float* d_wi, float* d_eval,
                        void* stream)
{
    if (!s || !gh || !bg)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_batch: null argument");
    Group& g = *gh->g;
    if (!g.jitted) {
        if (int rc = oslmi_jit_group(s, gh))
            return rc;
    }
    GroupArgs A;
    if (int rc = make_group_args(s->ss, g, bg, d_outputs, d_wo, d_wi, d_eval,
                                 "oslmi_execute_batch", A))
        return rc;
    if (bg->n == 0)
        return OSLMI_OK;
    void* args[] = { &A };
    if (int rc = check_hip(hipModuleLaunchKernel(g.kernel, grid_for(bg->n), 1, 1, 256, 1, 1, 0,
                                                 (hipStream_t)stream, args, nullptr),
                           "hipModuleLaunchKernel"))
        return rc;
    return OSLMI_OK;
}

int oslmi_execute_batch_sample(oslmi_shadingsys* s, oslmi_group* gh,
                               const oslmi_batched_globals* bg, float* d_outputs,
                               const float* d_wo, const float* d_rnd, float* d_samples,
                               void* stream)
{
    if (!s || !gh || !bg || !d_wo || !d_rnd || !d_samples)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_batch_sample: null argument");
    Group& g = *gh->g;
    if (!g.jitted) {
        if (int rc = oslmi_jit_group(s, gh))
            return rc;
    }
    if (!g.can_sample)
        return set_error(OSLMI_ERR_INVALID,
                         "oslmi_execute_batch_sample: group " + g.name
                             + " has no surface closure or was compiled without the"
                               " closure_sample attribute");
    GroupArgs A;
    if (int rc = make_group_args(s->ss, g, bg, d_outputs, nullptr, nullptr, nullptr,
                                 "oslmi_execute_batch_sample", A))
        return rc;
    A.wo   = d_wo;
    A.rnd  = d_rnd;
    A.samp = d_samples;
    if (bg->n == 0)
        return OSLMI_OK;
    void* args[] = { &A };
    return check_hip(hipModuleLaunchKernel(g.kernel, grid_for(bg->n), 1, 1, 256, 1, 1, 0,
                                           (hipStream_t)stream, args, nullptr),
                     "hipModuleLaunchKernel");
}

int oslmi_execute_batch_closures(oslmi_shadingsys* s, oslmi_group* gh,
                                 const oslmi_batched_globals* bg, float* d_outputs,
                                 const float* d_wo, int max_lobes, float* d_closures,
                                 void* stream)
{
    if (!s || !gh || !bg || !d_closures)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_batch_closures: null argument");
    if (max_lobes < 1 || max_lobes > 8)
        return set_error(OSLMI_ERR_INVALID, "oslmi_execute_batch_closures: max_lobes must be 1..8");
    Group& g = *gh->g;
    if (!g.jitted) {
        if (int rc = oslmi_jit_group(s, gh))
            return rc;
    }
    if (!g.can_export)
        return set_error(OSLMI_ERR_INVALID,
                         "oslmi_execute_batch_closures: group " + g.name
                             + " has no surface closure or was compiled without the"
                               " closure_export attribute");
    GroupArgs A;
    if (int rc = make_group_args(s->ss, g, bg, d_outputs, nullptr, nullptr, nullptr,
                                 "oslmi_execute_batch_closures", A))
        return rc;
    A.wo      = d_wo;
    A.clo     = d_closures;
    A.clo_max = max_lobes;
    if (bg->n == 0)
        return OSLMI_OK;
    void* args[] = { &A };
    return check_hip(hipModuleLaunchKernel(g.kernel, grid_for(bg->n), 1, 1, 256, 1, 1, 0,
                                           (hipStream_t)stream, args, nullptr),
                     "hipModuleLaunchKernel");
}

}  // extern "C"
