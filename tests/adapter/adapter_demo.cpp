// A renderer-side program written against the C++ facade
// (openshadinglanguage_amd/adapter/oslmi_batched.hpp), the way a renderer
// moving from OSL's BatchedExecutor<16> would: it builds the three-layer
// layered3 group from .oso text with ShaderGroupBegin / Shader /
// ConnectShaders / ShaderGroupEnd, places three output symbols into its own
// pixel records with add_symlocs, fills BatchedShaderGlobals<16> blocks
// and shades them twice: batch by batch through execute() (host memory, the
// reference's call shape) and all at once through execute_frame() (device
// memory). Both results go to the output file for the test to check.
//
// usage: adapter_demo <csrc dir> <jit cache> <shader dir> <input> <output>
//   input:  int64 n, then float P[3][n], N[3][n], then int32 shadeindex[n]
//           (a permutation of 0..n-1)
//   output: the pixel records of the per-batch run, then those of the
//           frame run: n records of { float fbm; float Cd[3]; float gab;
//           int32 tag; float pad[2]; }
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "oslmi_batched.hpp"

using namespace oslmi_adapter;

struct Pixel {
    float fbm;
    float Cd[3];
    float gab;
    int32_t tag;  // untouched by shading: must survive both paths
    float pad[2];
};
static_assert(sizeof(Pixel) == 32, "pixel record");

static std::string slurp(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    std::stringstream s;
    s << f.rdbuf();
    if (!f)
        throw std::runtime_error("cannot read " + path);
    return s.str();
}

int main(int argc, char** argv)
{
    if (argc != 6) {
        fprintf(stderr, "usage: %s <csrc dir> <jit cache> <shader dir> <input> <output>\n", argv[0]);
        return 2;
    }
    try {
        FILE* in = fopen(argv[4], "rb");
        if (!in)
            throw std::runtime_error("cannot open input");
        int64_t n = 0;
        if (fread(&n, 8, 1, in) != 1 || n < 1)
            throw std::runtime_error("bad input header");
        std::vector<float> P(3 * n), N(3 * n);
        std::vector<int32_t> index(n);
        if (fread(P.data(), 4, 3 * n, in) != (size_t)(3 * n) || fread(N.data(), 4, 3 * n, in) != (size_t)(3 * n)
            || fread(index.data(), 4, n, in) != (size_t)n)
            throw std::runtime_error("short input");
        fclose(in);

        ShadingSystem ss(argv[1], argv[2]);
        const std::string dir = argv[3];
        ss.LoadMemoryCompiledShader("c2_pattern", slurp(dir + "/layered3_pattern.oso"));
        ss.LoadMemoryCompiledShader("c2_color", slurp(dir + "/layered3_color.oso"));
        ss.LoadMemoryCompiledShader("c2_surface", slurp(dir + "/layered3_surface.oso"));
        ShaderGroup g = ss.ShaderGroupBegin("layered3");
        ss.Shader(g, "shader", "c2_pattern", "pattern");
        ss.Shader(g, "shader", "c2_color", "color");
        ss.Shader(g, "surface", "c2_surface", "surface");
        for (const char* p : { "fbm", "vn", "gab" })
            ss.ConnectShaders(g, "pattern", p, "color", p);
        ss.ConnectShaders(g, "color", "Cd", "surface", "Cd");
        ss.ShaderGroupEnd(g);
        ss.add_symlocs(g, { { "pattern.fbm", "float", offsetof(Pixel, fbm), sizeof(Pixel) },
                            { "color.Cd", "color", offsetof(Pixel, Cd), sizeof(Pixel) },
                            { "pattern.gab", "float", offsetof(Pixel, gab), sizeof(Pixel) } });
        ss.jit_group(g);

        // the renderer's batches: 16 lanes each, the last one partial
        const int64_t nbatch = (n + 15) / 16;
        const int last       = (int)(n - 16 * (nbatch - 1));
        std::vector<BatchedShaderGlobals16> bsg(nbatch);
        std::vector<int32_t> wide_index(16 * nbatch, 0);
        std::memset(bsg.data(), 0, bsg.size() * sizeof(BatchedShaderGlobals16));
        for (int64_t i = 0; i < n; ++i) {
            BatchedShaderGlobals16& b = bsg[i / 16];
            const int l               = (int)(i % 16);
            for (int c = 0; c < 3; ++c) {
                b.P[c][l] = P[c * n + i];
                b.N[c][l] = N[c * n + i];
                b.Ng[c][l] = N[c * n + i];
            }
            b.dPdx[0][l] = 8.0f / 1024;  // the benchmark's screen derivatives
            b.dPdy[1][l] = 8.0f / 1024;
            wide_index[i] = index[i];
        }

        std::vector<Pixel> px1(n), px2(n);
        for (int64_t i = 0; i < n; ++i)
            px1[i] = px2[i] = Pixel{ -1.0f, { -1.0f, -1.0f, -1.0f }, -1.0f, (int32_t)(7 * i + 3), { 0, 0 } };

        // 1: the reference's call shape, batch by batch from host memory
        BatchedExecutor16 exec(ss);
        for (int64_t b = 0; b < nbatch; ++b) {
            const int bs = b == nbatch - 1 ? last : 16;
            if (!exec.execute(g, bs, &wide_index[16 * b], bsg[b], nullptr, px1.data(),
                              px1.size() * sizeof(Pixel)))
                throw std::runtime_error(std::string("execute: ") + oslmi_last_error());
        }

        // 2: the whole frame in device memory, one call
        BatchedShaderGlobals16* d_bsg = nullptr;
        int32_t* d_index              = nullptr;
        Pixel* d_px                   = nullptr;
        check_hip(hipMalloc((void**)&d_bsg, bsg.size() * sizeof(BatchedShaderGlobals16)), "hipMalloc");
        check_hip(hipMalloc((void**)&d_index, wide_index.size() * 4), "hipMalloc");
        check_hip(hipMalloc((void**)&d_px, px2.size() * sizeof(Pixel)), "hipMalloc");
        check_hip(hipMemcpy(d_bsg, bsg.data(), bsg.size() * sizeof(BatchedShaderGlobals16),
                            hipMemcpyHostToDevice), "hipMemcpy");
        check_hip(hipMemcpy(d_index, wide_index.data(), wide_index.size() * 4, hipMemcpyHostToDevice),
                  "hipMemcpy");
        check_hip(hipMemcpy(d_px, px2.data(), px2.size() * sizeof(Pixel), hipMemcpyHostToDevice), "hipMemcpy");
        exec.execute_frame(g, nbatch, last, d_bsg, d_index, d_px);
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipMemcpy(px2.data(), d_px, px2.size() * sizeof(Pixel), hipMemcpyDeviceToHost), "hipMemcpy");
        (void)hipFree(d_bsg);
        (void)hipFree(d_index);
        (void)hipFree(d_px);

        FILE* out = fopen(argv[5], "wb");
        if (!out || fwrite(px1.data(), sizeof(Pixel), n, out) != (size_t)n
            || fwrite(px2.data(), sizeof(Pixel), n, out) != (size_t)n)
            throw std::runtime_error("cannot write output");
        fclose(out);
        printf("adapter_demo: %lld points in %lld batches (last %d) shaded both ways\n",
               (long long)n, (long long)nbatch, last);
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "adapter_demo: %s\n", e.what());
        return 1;
    }
}
