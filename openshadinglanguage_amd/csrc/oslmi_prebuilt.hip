// Prebuilt lowerings of the two benchmark shader groups (the OSL sources
// are in openshadinglanguage_amd/shaders/):
//
//   fbm_cell  : one layer; 4-octave perlin fBm + 0.25*cellnoise(3P) ->
//               smoothstep -> mix of two colors.
//   layered3  : pattern (6-octave simplex fBm, vector perlin, anisotropic
//               2D Gabor; all with derivatives) -> color (filterwidth-
//               antialiased smoothstep, mix, blackbody, transformc XYZ->rgb,
//               luminance) -> surface (Cd*diffuse(N) + microfacet ggx), with
//               the closure evaluated at one (wo, wi) pair per point.
//
// Each group is one fused kernel: layer outputs stay in registers, the
// closure is a register-resident lobe list, the blackbody table is staged
// in LDS per workgroup. One work-item shades one point; globals are read
// from SoA planes (coalesced 256-byte wavefront loads).
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "oslmi.h"
#include "oslmi_closure.h"
#include "oslmi_color.h"
#include "oslmi_colorsys.h"
#include "oslmi_gabor.h"
#include "oslmi_internal.h"
#include "oslmi_noise.h"

namespace oslmi {

// ------------------------------------------------------------ fbm_cell
OSLMI_DEV V3 mixv(V3 a, V3 b, float x)
{
    return V3(a.x * (1.0f - x) + b.x * x, a.y * (1.0f - x) + b.y * x, a.z * (1.0f - x) + b.z * x);
}
OSLMI_DEV float clampf(float x, float lo, float hi) { return fmaxf(fminf(x, hi), lo); }

__global__ __launch_bounds__(256) void k_fbm_cell(const float* __restrict__ P,
                                                  float* __restrict__ Cout, int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        V3 p(P[i], P[n + i], P[2 * n + i]);
        float f   = 0.0f;
        float amp = 1.0f;
        V3 pp     = p;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f   = f + amp * perlin3<float>(pp.x, pp.y, pp.z);
            amp = amp * 0.5f;
            pp  = pp * 2.0f;
        }
        V3 pc       = p * 3.0f;
        f           = f + 0.25f * intnoise3<true>(pc.x, pc.y, pc.z);
        float t     = smoothstep(0.0f, 1.0f, f);
        V3 c        = mixv(V3(0.1f, 0.2f, 0.4f), V3(0.9f, 0.7f, 0.3f), t);
        Cout[i]         = c.x;
        Cout[n + i]     = c.y;
        Cout[2 * n + i] = c.z;
    }
}

// ------------------------------------------------------------ layered3
struct Layered3Uniform {
    float dPdx[3], dPdy[3];
    GaborUniform gabor;
};

template<bool AOV>
__global__ __launch_bounds__(256) void k_layered3(const float* __restrict__ P,
                                                  const float* __restrict__ N,
                                                  const float* __restrict__ WO,
                                                  const float* __restrict__ WI,
                                                  Layered3Uniform U,
                                                  const ColorSysDev* __restrict__ cs,
                                                  float* __restrict__ out,
                                                  float* __restrict__ aov, int64_t n)
{
    __shared__ float s_bb[OSLMI_BB_TABLE * 3];
    for (int j = threadIdx.x; j < OSLMI_BB_TABLE * 3; j += blockDim.x)
        s_bb[j] = cs->bb[j];
    __syncthreads();

    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        V3 p(P[i], P[n + i], P[2 * n + i]);
        V3 nn(N[i], N[n + i], N[2 * n + i]);
        V3 wo(WO[i], WO[n + i], WO[2 * n + i]);
        V3 wi(WI[i], WI[n + i], WI[2 * n + i]);

        // layer A "pattern"
        Dv Q   = Dv(p, V3(U.dPdx[0], U.dPdx[1], U.dPdx[2]), V3(U.dPdy[0], U.dPdy[1], U.dPdy[2])) * 8.0f;
        Df fbm = Df(0.0f);
        float amp = 1.0f;
        Dv q      = Q;
        for (int o = 0; o < 6; ++o) {
            fbm = fbm + amp * simplex3_d(comp(q, 0), comp(q, 1), comp(q, 2), 0u);
            amp = amp * 0.5f;
            q   = q * 2.0f;
        }
        Df vn[3];
        vperlin3<Df>(comp(Q, 0), comp(Q, 1), comp(Q, 2), vn);
        Dv Pg(V3(Q.v.x, Q.v.y, 0.0f), V3(Q.dx.x, Q.dx.y, 0.0f), V3(Q.dy.x, Q.dy.y, 0.0f));
        Df gab = gabor(U.gabor, Pg);

        // layer B "color"
        float fw   = sqrtf(fbm.dx * fbm.dx + fbm.dy * fbm.dy);
        float t    = smoothstep(-0.8f - fw, 0.8f + fw, fbm.v);
        V3 base    = mixv(V3(0.8f, 0.3f, 0.1f), V3(0.1f, 0.4f, 0.8f), t);
        V3 xyz     = V3(0.5f, 0.5f, 0.5f) + 0.5f * V3(vn[0].v, vn[1].v, vn[2].v);
        float vw   = sqrtf(vn[1].dx * vn[1].dx + vn[1].dy * vn[1].dy);
        V3 tc      = xyz_to_rgb(cs->xyz2rgb, xyz) * (1.0f - clampf(vw, 0.0f, 1.0f));
        float temp = 1000.0f + 9000.0f * clampf(0.5f + 0.5f * gab.v, 0.0f, 1.0f);
        V3 bb      = blackbody_rgb(s_bb, cs->xyz2rgb, cs->cie, temp);
        float lum  = luminance(cs->lum, bb);
        V3 c       = mixv(base, tc, 0.3f) * (bb / lum);
        V3 Cd(clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f));

        // layer C "surface"
        V3 pn       = p * 8.0f + V3(17.3f, 4.1f, 9.7f);
        float nz    = 0.5f * (perlin3<float>(pn.x, pn.y, pn.z) + 1.0f);
        float alpha = 0.05f + 0.9f * nz;
        LobeList<2> L;
        Lobe l0;
        l0.id     = kClosureDiffuse;
        l0.weight = Cd;
        l0.N      = nn;
        l0.alpha  = 0.0f;
        l0.eta    = 0.0f;
        L.add(l0);
        Lobe l1;
        l1.id     = kClosureMicrofacetGGX;
        l1.weight = V3(1.0f);
        l1.N      = nn;
        l1.alpha  = alpha;
        l1.eta    = 1.5f;
        L.add(l1);
        composite_prepare(L, wo);
        BsdfEval e = composite_eval(L, wo, wi);

        out[i]         = e.weight.x;
        out[n + i]     = e.weight.y;
        out[2 * n + i] = e.weight.z;
        out[3 * n + i] = e.pdf;
        if (AOV) {
            const float v[20] = { fbm.v,    fbm.dx,   fbm.dy,   vn[0].v, vn[1].v, vn[2].v, vn[0].dx,
                                  vn[1].dx, vn[2].dx, vn[0].dy, vn[1].dy, vn[2].dy, gab.v,   gab.dx,
                                  gab.dy,   Cd.x,     Cd.y,     Cd.z,     alpha,   0.0f };
#pragma unroll
            for (int c2 = 0; c2 < 20; ++c2)
                aov[c2 * n + i] = v[c2];
        }
    }
}

// Per-device color system in HBM (built on first use, kept for the
// process lifetime).
static std::mutex g_cs_mutex;
static std::vector<ColorSysDev*> g_cs_dev;

int device_colorsys(const ColorSysDev** out, bool allocate)
{
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))
        return rc;
    std::lock_guard<std::mutex> lock(g_cs_mutex);
    if ((int)g_cs_dev.size() <= dev)
        g_cs_dev.resize(dev + 1, nullptr);
    if (!g_cs_dev[dev] && !allocate)
        return set_error(OSLMI_ERR_INVALID,
                         "colour system not yet in this device's memory: call oslmi_init first");
    if (!g_cs_dev[dev]) {
        ColorSysDev host;
        if (!build_colorsys("Rec709", host))
            return set_error(OSLMI_ERR_INVALID, "color system Rec709 unavailable");
        ColorSysDev* d = nullptr;
        if (int rc = check_hip(hipMalloc((void**)&d, sizeof(ColorSysDev)), "hipMalloc(colorsys)"))
            return rc;
        if (int rc = check_hip(hipMemcpy(d, &host, sizeof(ColorSysDev), hipMemcpyHostToDevice),
                               "hipMemcpy(colorsys)"))
            return rc;
        g_cs_dev[dev] = d;
    }
    *out = g_cs_dev[dev];
    return OSLMI_OK;
}

}  // namespace oslmi

using namespace oslmi;

extern "C" int oslmi_shade_fbm_cell(int64_t n, const float* d_P, float* d_Cout, void* stream)
{
    if (n < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shade_fbm_cell: n < 0");
    if (n == 0)
        return OSLMI_OK;
    if (!d_P || !d_Cout)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shade_fbm_cell: null buffer");
    hipLaunchKernelGGL(k_fbm_cell, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, d_P,
                       d_Cout, n);
    return check_launch("oslmi_shade_fbm_cell");
}

extern "C" int oslmi_shade_layered3(int64_t n, const float* d_P, const float* d_N,
                                    const float* d_wo, const float* d_wi, const float* dPdx,
                                    const float* dPdy, float* d_out, float* d_aov, void* stream)
{
    if (n < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shade_layered3: n < 0");
    if (n == 0)
        return OSLMI_OK;
    if (!d_P || !d_N || !d_wo || !d_wi || !dPdx || !dPdy || !d_out)
        return set_error(OSLMI_ERR_INVALID, "oslmi_shade_layered3: null argument");
    const ColorSysDev* cs = nullptr;
    if (int rc = device_colorsys(&cs, true))
        return rc;
    Layered3Uniform U;
    for (int c = 0; c < 3; ++c) {
        U.dPdx[c] = dPdx[c];
        U.dPdy[c] = dPdy[c];
    }
    GaborOptions go;
    go.anisotropic = 1;
    go.do_filter   = 1;
    U.gabor        = build_gabor_uniform(go);
    dim3 grid(grid_for(n)), block(256);
    if (d_aov)
        hipLaunchKernelGGL(k_layered3<true>, grid, block, 0, (hipStream_t)stream, d_P, d_N, d_wo,
                           d_wi, U, cs, d_out, d_aov, n);
    else
        hipLaunchKernelGGL(k_layered3<false>, grid, block, 0, (hipStream_t)stream, d_P, d_N, d_wo,
                           d_wi, U, cs, d_out, d_aov, n);
    return check_launch("oslmi_shade_layered3");
}
