// Layered MaterialX closure layer(sheen_bsdf, layer(dielectric_bsdf,
// generalized_schlick_bsdf)) on gfx950: one thread per shading point builds
// the three lobes and their layer weights, evaluates the composite at one
// direction and draws four samples, as testrender's process_bsdf_closure +
// CompositeBSDF::prepare/eval/sample do per shading point
// (src/testrender/shading.cpp:1448-1667, shading.h:319-393).
//
// Energy tables: baked on the host at first use by the same lobe code
// (libbsdl/src/genluts.cpp:50-192 with libm trigonometry), cached next to
// the library in bsdl_luts.bin, uploaded once per device and staged into
// LDS by every workgroup.
#include <dlfcn.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "oslmi.h"
#include "oslmi_group.h"
#include "oslmi_internal.h"
#include "oslmi_mtx.h"

// Workgroup sizes of the layered-closure kernels. Each workgroup stages the
// 33 KB energy tables in LDS, so at most four workgroups fit on a CU. The
// eval+sample kernel needs 119 registers (four waves per SIMD), which 256
// work-items per workgroup already reach; the eval-only kernels need about
// 84, and 512 work-items per workgroup let five or six waves per SIMD share
// one copy of the tables (6 % faster over 8M queries than 256).
constexpr int kLayeredBlock = 256;
constexpr int kLocalBlock   = 512;

namespace oslmi {
namespace mtx {

// ---- host: table bake --------------------------------------------------
namespace {

struct TrigLibm {
    static void sc(float x, float* s, float* c)
    {
        *c = cosf(x);
        *s = sinf(x);
    }
};

// Larcher-Pillichshammer points with Owen-style scrambles (genluts.cpp:52-100)
uint32_t lp_radical(uint32_t i)
{
    uint32_t r = 0;
    for (uint32_t v = 1U << 31; i; i >>= 1, v |= v >> 1)
        if (i & 1)
            r ^= v;
    return r;
}
uint32_t lp_radical_inv(uint32_t i)
{
    uint32_t r = 0;
    for (uint32_t v = 3U << 30; i; i >>= 1, v >>= 1)
        if (i & 1)
            r ^= v;
    return r;
}
V3 lp_point(int si, int AA, uint32_t sx, uint32_t sy, uint32_t sz)
{
    const uint32_t ex    = si % AA;
    const uint32_t ey    = si / AA;
    const uint32_t upper = (ex ^ (sx >> 16)) << 16;
    const uint32_t lpu   = lp_radical(upper) ^ sy;
    const uint32_t delta = (ey << 16) ^ (lpu & 0xFFFF0000u);
    const uint32_t index = upper | lp_radical_inv(delta);
    const uint32_t x     = index ^ sx;
    const uint32_t y     = lpu ^ delta;
    const float jx       = (x & 65535) * (1 / 65536.0f);
    const float jy       = (y & 65535) * (1 / 65536.0f);
    uint32_t rz = sz, ii = index;
    for (uint64_t v2 = uint64_t(3) << 62; ii; ii >>= 1, v2 ^= v2 >> 1)
        if (ii & 1)
            rz ^= uint32_t(v2 >> 31);
    return V3((ex + jx) / AA, (ey + jy) / AA, rz * 2.3283063e-10f);
}
// fasthash64 of {x << 32 + y, z} (genluts.cpp:105-123)
uint32_t entry_hash(uint32_t x, uint32_t y, uint32_t z)
{
    auto mix = [](uint64_t h) {
        h ^= h >> 23;
        h *= 0x2127599bf4325c37ULL;
        h ^= h >> 47;
        return h;
    };
    const uint64_t m    = 0x880355f21e6d1965ULL;
    const uint64_t v[2] = { (uint64_t(x) << 32) + y, uint64_t(z) };
    uint64_t h          = 16 * m;
    for (int k = 0; k < 2; k++) {
        h ^= mix(v[k]);
        h *= m;
    }
    return (uint32_t)mix(h);
}
// 1 - mean sampled weight over 128 x 128 points (genluts.cpp:125-152)
template<class W> float table_entry(float cos_theta, W weight, uint32_t f, uint32_t r)
{
    const V3 wo(sqrtf(1 - SQR(cos_theta)), 0, cos_theta);
    const uint32_t sx = entry_hash(f, r, 0), sy = entry_hash(f, r, 1), sz = entry_hash(f, r, 2);
    float E = 0;
    for (int i = 0; i < 128 * 128; i++)
        E = LERP(1.0f / (1.0f + i), E, weight(wo, lp_point(i, 128, sx, sy, sz)));
    return 1 - smin(E, 1.0f);
}
template<class F> void run_parallel(int n, F fn)
{
    unsigned hc = std::thread::hardware_concurrency();
    int nt      = hc ? (int)(hc < 16 ? hc : 16) : 4;
    std::atomic<int> next { 0 };
    auto work = [&]() {
        for (int i = next++; i < n; i = next++)
            fn(i);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++)
        th.emplace_back(work);
    work();
    for (auto& t : th)
        t.join();
}

void bake_tables(float* out)
{
    float* mini = out + kLutMiniGGX;
    // MiniMicrofacetGGX: sample weight G2/G1 of a VNDF-sampled reflection
    run_parallel(kNr, [&](int r) {
        const GGX d = ggx_make(float(r) * (1.0f / (kNr - 1)), 0.0f, false);
        for (int c = 0; c < kNc; c++) {
            auto w = [&](V3 wo, V3 rnd) {
                const V3 m = ggx_sample_vndf(d, wo, rnd.x, rnd.y);
                if (vdot(m, wo) > 0) {
                    const V3 wi = reflect(wo, m);
                    if (wi.z > 0)
                        return ggx_G2_G1(d, wi, wo);
                }
                return 0.0f;
            };
            mini[r * kNc + c] = table_entry(get_cosine(kLutMiniGGX, c), w, 0u, (uint32_t)r);
        }
    });
    // DielectricReflFront: Turquin reflection of DielectricFresnel
    float* refl = out + kLutReflFront;
    run_parallel(kNfRefl * kNr, [&](int fr) {
        const int f         = fr / kNr;
        const int r         = fr % kNr;
        const float findex  = float(f) * (1.0f / (kNfRefl - 1));
        const float rough   = float(r) * (1.0f / (kNr - 1));
        SpecDielectric sp;
        sp.d   = ggx_make(rough, 0.0f, false);
        sp.eta = dfresnel_eta(LERP(SQR(findex), kIOR_MIN, kIOR_MAX), false);
        for (int c = 0; c < kNc; c++) {
            const float cth = get_cosine(kLutReflFront, c);
            sp.E_ms         = emiss_eval(out, kLutMiniGGX, rough, 0.0f, cth);
            auto w          = [&](V3 wo, V3 rnd) {
                if (wo.z <= 0)
                    return 0.0f;
                const V3 m  = ggx_sample_bvndf<TrigLibm>(sp.d, wo, rnd.x, rnd.y);
                const V3 wi = reflect(wo, m);
                if (wi.z <= 0)
                    return 0.0f;
                return pmax(eval_turquin(sp, wo, wi).weight);
            };
            refl[(f * kNr + r) * kNc + c] = table_entry(cth, w, (uint32_t)f, (uint32_t)r);
        }
    });
}

const char kMagic[8] = { 'B', 'S', 'D', 'L', 'L', 'U', 'T', '1' };

std::string table_path()
{
    Dl_info info;
    if (dladdr((void*)&bake_tables, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.find_last_of('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/bsdl_luts.bin";
    }
    return "bsdl_luts.bin";
}
bool load_tables(const std::string& path, float* out)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f)
        return false;
    char magic[8];
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kMagic, 8) == 0
              && fread(out, sizeof(float), kLutFloats, f) == (size_t)kLutFloats;
    fclose(f);
    return ok;
}
void save_tables(const std::string& path, const float* t)
{
    std::string tmp = path + ".tmp";
    FILE* f         = fopen(tmp.c_str(), "wb");
    if (!f)
        return;
    bool ok = fwrite(kMagic, 1, 8, f) == 8
              && fwrite(t, sizeof(float), kLutFloats, f) == (size_t)kLutFloats;
    fclose(f);
    if (ok)
        rename(tmp.c_str(), path.c_str());
    else
        remove(tmp.c_str());
}

std::once_flag g_host_once;
std::vector<float> g_host_tables;
std::mutex g_dev_mutex;
std::vector<float*> g_dev_tables;

}  // namespace

const float* host_tables()
{
    std::call_once(g_host_once, []() {
        g_host_tables.assign(kLutFloats, 0.0f);
        const std::string path = table_path();
        if (!load_tables(path, g_host_tables.data())) {
            bake_tables(g_host_tables.data());
            save_tables(path, g_host_tables.data());
        }
    });
    return g_host_tables.data();
}

int device_tables(const float** out, bool allocate)
{
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))
        return rc;
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    if ((int)g_dev_tables.size() <= dev)
        g_dev_tables.resize(dev + 1, nullptr);
    if (!g_dev_tables[dev] && !allocate)
        return set_error(OSLMI_ERR_INVALID,
                         "libbsdl energy tables not yet in this device's memory (run one "
                         "allocating execution of a MaterialX closure group first)");
    if (!g_dev_tables[dev]) {
        const float* host = host_tables();
        float* d = nullptr;
        if (int rc = check_hip(hipMalloc((void**)&d, kLutFloats * sizeof(float)),
                               "hipMalloc(bsdl tables)"))
            return rc;
        if (int rc = check_hip(hipMemcpy(d, host, kLutFloats * sizeof(float),
                                         hipMemcpyHostToDevice),
                               "hipMemcpy(bsdl tables)"))
            return rc;
        g_dev_tables[dev] = d;
    }
    *out = g_dev_tables[dev];
    return OSLMI_OK;
}

// ---- device: the layered material --------------------------------------
struct Material {
    V3 coat_refl, base_refl;
    float base_exponent, base_roughness;
    V3 sheen_albedo;
};

struct MicroLobe {
    Spec sp;
    V3 tint;  // reflection tint (transmission tint is zero)
    float E_ms;
    bool dorefl;
};

struct Layered {
    Frame fs, fm;   // sheen frame (X = wo); coat and base frame (X from N)
    V3 wo_s, wo_m;  // wo in the two frames
    V3 s_tint;
    float s_a, s_Emiss;
    MicroLobe coat, base;
    int n;  // lobes present: sheen, then coat, then base
    V3 w[3];
    float pdf[3];
};

// SheenLobe::eval_impl (front facing)
OSLMI_DEV Sample sheen_eval(const Layered& L, V3 wo, V3 wi)
{
    if (!(wi.z > 0 && wo.z >= 0))
        return zero_sample();
    return sheen_mf_eval(L.s_a, L.s_tint, wo, wi);
}
// DielectricLobe / SchlickLobe ::eval_impl, reflection only
OSLMI_DEV Sample micro_eval(const MicroLobe& M, V3 wo, V3 wi)
{
    if (!M.dorefl)
        return zero_sample();
    Sample s = eval_turquin(M.sp, wo, wi);
    s.weight = s.weight * (s.wi.z > 0 ? M.tint : V3(0.0f));
    return s;
}

OSLMI_DEV void micro_setup(MicroLobe* M, float roughness_xy, float cosNO, const float* T)
{
    const float rx    = CLAMP(roughness_xy, kEPSILON, 2.0f);
    const float ry    = rx;
    const float ax    = smax(rx, ry);
    const float ay    = smin(rx, ry);
    const float b     = ay / ax;
    const float aniso = (1 - b) / (1 + b);
    const float rough = regularize_roughness(sqrtf(ax / (1 + aniso)));
    M->sp.d           = ggx_make(rough, aniso, rx < ry);
    M->sp.E_ms        = emiss_eval(T, kLutMiniGGX, rough, 0.0f, cosNO);
    M->E_ms           = M->dorefl ? emiss_eval(T, kLutReflFront, rough,
                                               dfresnel_table_index(M->sp.eta), cosNO)
                                  : 0.0f;
}

OSLMI_DEV void layered_build(Layered* L, const Material& mat, V3 N, V3 wo, float coat_ior,
                             float coat_rough, float sheen_rough, V3 F0, V3 F90, const float* T)
{
    const V3 Z = visible_normal(wo, N, N);
    L->fs      = frame_zx(Z, wo);
    L->fm      = frame_zx(Z, V3(0.0f));
    L->wo_s    = L->fs.local(wo);
    L->wo_m    = L->fm.local(wo);
    // sheen_bsdf(N, albedo, roughness), Conty-Kulla mode
    const float s_rough = regularize_roughness(CLAMP(sheen_rough, 0.0f, 1.0f));
    L->s_tint           = mat.sheen_albedo;
    L->s_a              = CLAMP(smax(0.06f, s_rough), 0.06f, 1.0f);
    const float cosNOs  = CLAMP(vdot(L->fs.Z, wo), 0.0f, 1.0f);
    L->s_Emiss          = 1 - smin(conty_kulla_albedo(L->s_a, cosNOs) * pmax(L->s_tint), 1.0f);
    const float cosNO = vdot(L->fm.Z, wo);
    // dielectric_bsdf coat
    L->coat.tint    = mat.coat_refl;
    L->coat.dorefl  = MAX_RGB(mat.coat_refl) > 0;
    L->coat.sp.kind = 0;
    L->coat.sp.eta  = dfresnel_eta(CLAMP(coat_ior, kIOR_MIN, kIOR_MAX) / 1.0f, false);
    L->coat.sp.F0 = L->coat.sp.F90 = V3(0.0f);
    L->coat.sp.exponent = L->coat.sp.tir_cos = 0.0f;
    micro_setup(&L->coat, coat_rough, cosNO, T);
    // generalized_schlick_bsdf base
    const float avg_F0 = CLAMP((F0.x + F0.y + F0.z) * (1.0f / 3), 0.0f, 0.99f);
    const float sqF0   = sqrtf(avg_F0);
    L->base.tint       = mat.base_refl;
    L->base.dorefl     = MAX_RGB(mat.base_refl) > 0;
    L->base.sp.kind    = 1;
    L->base.sp.eta     = dfresnel_eta((1 + sqF0) / (1 - sqF0), false);
    L->base.sp.F0 = V3(smax(0.0f, smin(1.0f, F0.x)), smax(0.0f, smin(1.0f, F0.y)),
                       smax(0.0f, smin(1.0f, F0.z)));
    L->base.sp.F90 = V3(smax(0.0f, smin(1.0f, F90.x)), smax(0.0f, smin(1.0f, F90.y)),
                        smax(0.0f, smin(1.0f, F90.z)));
    L->base.sp.exponent = mat.base_exponent;
    L->base.sp.tir_cos  = L->base.sp.eta >= 1 ? 0 : sqrtf(1 - SQR(L->base.sp.eta));
    micro_setup(&L->base, mat.base_roughness < 0 ? coat_rough : mat.base_roughness, cosNO, T);
    // layer weights: top opacity = 1 - filter_o, clamped (closure weights 1)
    const V3 one(1.0f);
    auto oclamp = [](float a) {
        float v = a;
        if (!(0.0f <= v))
            v = 0.0f;
        if (!(v <= 1.0f))
            v = 1.0f;
        return v;
    };
    auto black = [](V3 c) { return c.x == 0 && c.y == 0 && c.z == 0; };
    L->n          = 1;
    L->w[0]       = one * one;
    V3 op         = one * (one * (one - V3(L->s_Emiss)));
    const V3 bw   = one * (one - V3(oclamp(op.x), oclamp(op.y), oclamp(op.z)));
    if (!black(bw)) {
        L->n       = 2;
        L->w[1]    = (bw * one) * one;
        V3 opc     = one * (one * (one - V3(L->coat.E_ms)));
        const V3 bw2 = bw * (one - V3(oclamp(opc.x), oclamp(opc.y), oclamp(opc.z)));
        if (!black(bw2)) {
            L->n    = 3;
            L->w[2] = bw2 * one;
        }
    }
    // CompositeBSDF::prepare(wo, path_weight = 1, absorb = false)
    const V3 alb[3] = { L->s_tint * (1 - L->s_Emiss), L->coat.tint * (1 - L->coat.E_ms),
                        L->base.tint * (1 - L->base.E_ms) };
    float total = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i < L->n) {
            L->pdf[i] = vdot(L->w[i], one * alb[i]) / (1.0f + 1.0f + 1.0f);
            total += L->pdf[i];
        } else {
            L->pdf[i] = 0.0f;
        }
    }
    if (total > 0 || total > 1) {
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (i < L->n)
                L->pdf[i] = L->pdf[i] / total;
    }
}

OSLMI_DEV void mis_update(V3* w, float* pdf, V3 ow, float opdf, float b)
{
    if (b > kFLT_MIN) {
        opdf *= b;
        ow = ow * (1 / b);
        float mis;
        if (*pdf < opdf)
            mis = 1 / (1 + *pdf / opdf);
        else if (opdf < *pdf)
            mis = 1 - 1 / (1 + opdf / *pdf);
        else
            mis = 0.5f;
        *w = *w * (1 - mis) + ow * mis;
        *pdf += opdf;
    }
}

// lobe i evaluated at the local directions of its frame
OSLMI_DEV Sample lobe_eval_local(const Layered& L, int i, V3 wi_local)
{
    if (i == 0)
        return sheen_eval(L, L.wo_s, wi_local);
    return micro_eval(i == 1 ? L.coat : L.base, L.wo_m, wi_local);
}

// CompositeBSDF::eval at world direction wi: weight rgb, pdf
OSLMI_DEV void layered_eval(const Layered& L, V3 wi, V3* weight, float* pdf)
{
    V3 sw(0.0f);
    float sp        = 0.0f;
    const V3 wi_s   = L.fs.local(wi);
    const V3 wi_m   = L.fm.local(wi);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i < L.n) {
            Sample b = lobe_eval_local(L, i, i == 0 ? wi_s : wi_m);
            mis_update(&sw, &sp, b.weight * L.w[i], b.pdf, L.pdf[i]);
        }
    }
    *weight = sw;
    *pdf    = sp;
}

// CompositeBSDF::sample(wo, rx, ry, rz): the chosen lobe draws the
// direction in its frame; every present lobe is then evaluated at it (the
// chosen one at its local direction, as the lobe's own sample does) and
// the others are MIS-combined in closure order.
OSLMI_DEV void layered_sample(const Layered& L, float rx, float ry, V3* wi_out, V3* weight,
                              float* pdf)
{
    *wi_out  = V3(0.0f);
    *weight  = V3(0.0f);
    *pdf     = 0.0f;
    int pick = -1;
    float accum = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (pick < 0 && i < L.n) {
            if (rx < (L.pdf[i] + accum)) {
                pick = i;
                rx   = (rx - accum) / L.pdf[i];
                rx   = smin(rx, 0.99999994f);
            } else {
                accum += L.pdf[i];
            }
        }
    }
    if (pick < 0)
        return;
    V3 wl;
    if (pick == 0) {
        wl = sample_uniform_hemisphere(rx, ry);
    } else {
        const MicroLobe& M = pick == 1 ? L.coat : L.base;
        if (!M.dorefl || L.wo_m.z <= 0)
            return;
        const V3 m = ggx_sample_bvndf<TrigFast>(M.sp.d, L.wo_m, rx, ry);
        wl         = reflect(L.wo_m, m);
        if (wl.z <= 0 || MAX_ABS_XYZ(wl) < kEPSILON)
            return;
    }
    const Frame& f = pick == 0 ? L.fs : L.fm;
    const V3 wi    = f.world(wl);
    Sample b[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (j < L.n) {
            V3 d = j == pick ? wl : (j == 0 ? L.fs.local(wi) : L.fm.local(wi));
            b[j] = lobe_eval_local(L, j, d);
        }
    }
    V3 sw    = V3(0.0f);
    float sp = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (j == pick) {
            sw = b[j].weight * (L.w[j] * (1 / L.pdf[j]));
            sp = b[j].pdf * L.pdf[j];
        }
    }
    if (sp == 0.0f)
        return;
#pragma unroll
    for (int j = 0; j < 3; j++)
        if (j < L.n && j != pick)
            mis_update(&sw, &sp, b[j].weight * L.w[j], b[j].pdf, L.pdf[j]);
    *wi_out = wi;
    *weight = sw;
    *pdf    = sp;
}

// ---- kernels ---------------------------------------------------------------
// Tables staged in LDS: 8448 floats (33 KB) per workgroup.
__global__ __launch_bounds__(kLayeredBlock) void k_layered(Material mat, const float* __restrict__ lut,
                                                 const float* __restrict__ Np,
                                                 const float* __restrict__ Vp,
                                                 const float* __restrict__ Lp,
                                                 const float* __restrict__ prm,
                                                 const float* __restrict__ rnd,
                                                 float* __restrict__ eval,
                                                 float* __restrict__ samp, int64_t n)
{
    __shared__ float T[kLutFloats];
    for (int k = threadIdx.x; k < kLutFloats; k += blockDim.x)
        T[k] = lut[k];
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const V3 N(Np[i], Np[n + i], Np[2 * n + i]);
        const V3 wo(Vp[i], Vp[n + i], Vp[2 * n + i]);
        Layered L;
        layered_build(&L, mat, N, wo, prm[i], prm[n + i], prm[2 * n + i],
                      V3(prm[3 * n + i], prm[4 * n + i], prm[5 * n + i]),
                      V3(prm[6 * n + i], prm[7 * n + i], prm[8 * n + i]), T);
        if (eval) {
            V3 w;
            float p;
            layered_eval(L, V3(Lp[i], Lp[n + i], Lp[2 * n + i]), &w, &p);
            eval[i]         = w.x;
            eval[n + i]     = w.y;
            eval[2 * n + i] = w.z;
            eval[3 * n + i] = p;
        }
        if (samp) {
#pragma unroll 1
            for (int k = 0; k < 4; k++) {
                V3 wi, w;
                float p;
                layered_sample(L, rnd[(2 * k) * n + i], rnd[(2 * k + 1) * n + i], &wi, &w, &p);
                float* o     = samp + (int64_t)(7 * k) * n;
                o[i]         = wi.x;
                o[n + i]     = wi.y;
                o[2 * n + i] = wi.z;
                o[3 * n + i] = w.x;
                o[4 * n + i] = w.y;
                o[5 * n + i] = w.z;
                o[6 * n + i] = p;
            }
        }
    }
}

// Exact evaluation of surrogate queries in the local frame (N = +z,
// wo = (x, 0, z)): q 14 planes = wi xyz, wo x z, coat ior, coat roughness,
// sheen roughness, F0 rgb, F90 rgb; out 4 planes = weight rgb, pdf.
OSLMI_DEV void eval_query(const Material& mat, const float* T, const float* __restrict__ q,
                          int64_t n, int64_t i, float* __restrict__ out)
{
    const float lz = q[2 * n + i], vz = q[4 * n + i];
    if (vz > 0.0f && lz <= 0.0f) {
        // N = +z is the visible normal and every lobe is reflection-only:
        // the composite's weight and pdf are exactly zero
        out[i] = out[n + i] = out[2 * n + i] = out[3 * n + i] = 0.0f;
        return;
    }
    const float* p = q + 5 * n;
    Layered L;
    layered_build(&L, mat, V3(0.0f, 0.0f, 1.0f), V3(q[3 * n + i], 0.0f, vz), p[i], p[n + i],
                  p[2 * n + i], V3(p[3 * n + i], p[4 * n + i], p[5 * n + i]),
                  V3(p[6 * n + i], p[7 * n + i], p[8 * n + i]), T);
    V3 w;
    float pdf;
    layered_eval(L, V3(q[i], q[n + i], lz), &w, &pdf);
    out[i]         = w.x;
    out[n + i]     = w.y;
    out[2 * n + i] = w.z;
    out[3 * n + i] = pdf;
}

__global__ __launch_bounds__(kLocalBlock) void k_layered_local(Material mat,
                                                       const float* __restrict__ lut,
                                                       const float* __restrict__ q,
                                                       float* __restrict__ out, int64_t n)
{
    __shared__ float T[kLutFloats];
    for (int k = threadIdx.x; k < kLutFloats; k += blockDim.x)
        T[k] = lut[k];
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        eval_query(mat, T, q, n, i, out);
}

// The same over the query indices idx[0 .. *count) (count written on the
// device by an earlier kernel on the stream; every index < n).
__global__ __launch_bounds__(kLocalBlock) void k_layered_local_indexed(Material mat,
                                                               const float* __restrict__ lut,
                                                               const float* __restrict__ q,
                                                               int64_t n,
                                                               const int* __restrict__ idx,
                                                               const int* __restrict__ count,
                                                               float* __restrict__ out)
{
    __shared__ float T[kLutFloats];
    const int64_t cnt = *count;
    if ((int64_t)blockIdx.x * blockDim.x >= cnt)
        return;
    for (int k = threadIdx.x; k < kLutFloats; k += blockDim.x)
        T[k] = lut[k];
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += stride) {
        const int64_t qi = idx[i];
        if (qi >= 0 && qi < n)
            eval_query(mat, T, q, n, qi, out);
    }
}

}  // namespace mtx

// Persistent-style grid: enough workgroups to fill the chip several times
// over while amortising the LDS table fill.
static unsigned layered_grid(int64_t n, int block)
{
    int64_t blocks    = (n + block - 1) / block;
    const int64_t cap = 256 * 8 * 256 / block;
    if (blocks > cap)
        blocks = cap;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

static int material_of(const oslmi_layered_material* m, mtx::Material* out, const char* who)
{
    if (!m)
        return set_error(OSLMI_ERR_INVALID, std::string(who) + ": null material");
    for (int c = 0; c < 3; c++)
        if (m->coat_refr_tint[c] > 0 || m->base_refr_tint[c] > 0)
            return set_error(OSLMI_ERR_UNSUPPORTED,
                             std::string(who) + ": transmission (refr_tint > 0) is not supported");
    out->coat_refl      = V3(m->coat_refl_tint[0], m->coat_refl_tint[1], m->coat_refl_tint[2]);
    out->base_refl      = V3(m->base_refl_tint[0], m->base_refl_tint[1], m->base_refl_tint[2]);
    out->base_exponent  = m->base_exponent;
    out->base_roughness = m->base_roughness;
    out->sheen_albedo   = V3(m->sheen_albedo[0], m->sheen_albedo[1], m->sheen_albedo[2]);
    return OSLMI_OK;
}

// Exact evaluation of the queries listed in d_index[0 .. *d_count) (the
// surrogate's fallback list); the grid covers up to n listed queries.
int layered_eval_local_indexed(const oslmi_layered_material* m, int64_t n, const float* d_query,
                               const int* d_index, const int* d_count, float* d_out,
                               hipStream_t stream)
{
    mtx::Material mat;
    if (int rc = material_of(m, &mat, "surrogate fallback"))
        return rc;
    const float* lut = nullptr;
    if (int rc = mtx::device_tables(&lut))
        return rc;
    hipLaunchKernelGGL(mtx::k_layered_local_indexed, dim3(layered_grid(n, kLocalBlock)), dim3(kLocalBlock),
                       0, stream, mat, lut, d_query, n, d_index, d_count, d_out);
    return check_launch("surrogate fallback");
}

}  // namespace oslmi

using namespace oslmi;

extern "C" int oslmi_layered_closure(const oslmi_layered_material* m, int64_t n, const float* d_N,
                                     const float* d_V, const float* d_L, const float* d_params,
                                     const float* d_rnd, float* d_eval, float* d_samples,
                                     void* stream)
{
    mtx::Material mat;
    if (int rc = material_of(m, &mat, "oslmi_layered_closure"))
        return rc;
    if (n < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_layered_closure: n < 0");
    if (n == 0)
        return OSLMI_OK;
    if (!d_N || !d_V || !d_params || (d_eval && !d_L) || (d_samples && !d_rnd))
        return set_error(OSLMI_ERR_INVALID, "oslmi_layered_closure: null buffer");
    const float* lut = nullptr;
    if (int rc = mtx::device_tables(&lut))
        return rc;
    hipLaunchKernelGGL(mtx::k_layered, dim3(layered_grid(n, kLayeredBlock)), dim3(kLayeredBlock), 0,
                       (hipStream_t)stream, mat, lut, d_N, d_V, d_L, d_params, d_rnd, d_eval, d_samples, n);
    return check_launch("oslmi_layered_closure");
}

extern "C" int oslmi_layered_eval_local(const oslmi_layered_material* m, int64_t n,
                                        const float* d_query, float* d_out, void* stream)
{
    mtx::Material mat;
    if (int rc = material_of(m, &mat, "oslmi_layered_eval_local"))
        return rc;
    if (n < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_layered_eval_local: n < 0");
    if (n == 0)
        return OSLMI_OK;
    if (!d_query || !d_out)
        return set_error(OSLMI_ERR_INVALID, "oslmi_layered_eval_local: null buffer");
    const float* lut = nullptr;
    if (int rc = mtx::device_tables(&lut))
        return rc;
    hipLaunchKernelGGL(mtx::k_layered_local, dim3(layered_grid(n, kLocalBlock)), dim3(kLocalBlock), 0,
                       (hipStream_t)stream, mat, lut, d_query, d_out, n);
    return check_launch("oslmi_layered_eval_local");
}

extern "C" int oslmi_bsdl_tables(float* out, int64_t count)
{
    if (!out || count != mtx::kLutFloats)
        return set_error(OSLMI_ERR_INVALID, "oslmi_bsdl_tables: need a buffer of 8448 floats");
    memcpy(out, mtx::host_tables(), sizeof(float) * mtx::kLutFloats);
    return OSLMI_OK;
}
