// CPU oracle: C entry points for the layered-closure restatement (mtx.cpp).
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// Plane layouts match the library's oslmi_layered_closure:
//   mat[17]  coat_refl_tint[3], coat_refr_tint[3], base_refl_tint[3],
//            base_refr_tint[3], base_exponent, base_roughness (< 0: the
//            point's coat roughness), sheen_albedo[3]
//   N, V, L  3 planes each (V: towards the viewer, -I; L: eval direction)
//   prm      9 planes: coat ior, coat roughness, sheen roughness, F0 rgb, F90 rgb
//   rnd      8 planes: (rx, ry) of samples 0..3
//   eval     4 planes: weight rgb, pdf
//   samp     28 planes: per sample k, planes 7k..7k+6 = wi xyz, weight rgb, pdf
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "mtx.h"
#include "omath.h"

namespace {

template<class F> void mtx_parallel(int64_t n, int nthreads, F f)
{
    if (nthreads <= 1 || n < 1024) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t grain = 2048;
    std::atomic<int64_t> next(0);
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&]() {
            for (;;) {
                int64_t b = next.fetch_add(grain);
                if (b >= n)
                    break;
                f(b, std::min(n, b + grain));
            }
        });
    for (auto& t : th)
        t.join();
}

using oref::V3;
namespace M = oref::mtx;

V3 ld3m(const float* p, int64_t n, int64_t i) { return V3(p[i], p[n + i], p[2 * n + i]); }

M::LayeredParams params_of(const float* mat, const float* prm, int64_t n, int64_t i, V3 N)
{
    M::LayeredParams p;
    p.N               = N;
    p.coat_ior        = prm[i];
    p.coat_roughness  = prm[n + i];
    p.sheen_roughness = prm[2 * n + i];
    p.base_f0         = V3(prm[3 * n + i], prm[4 * n + i], prm[5 * n + i]);
    p.base_f90        = V3(prm[6 * n + i], prm[7 * n + i], prm[8 * n + i]);
    p.coat_refl_tint  = V3(mat[0], mat[1], mat[2]);
    p.coat_refr_tint  = V3(mat[3], mat[4], mat[5]);
    p.base_refl_tint  = V3(mat[6], mat[7], mat[8]);
    p.base_refr_tint  = V3(mat[9], mat[10], mat[11]);
    p.base_exponent   = mat[12];
    p.base_roughness  = mat[13] < 0 ? p.coat_roughness : mat[13];
    p.sheen_albedo    = V3(mat[14], mat[15], mat[16]);
    p.sheen_mode      = 0;
    return p;
}

}  // namespace

extern "C" {

void oref_mtx_layered(const float* mat, const float* N, const float* V, const float* L,
                      const float* prm, const float* rnd, float* eval, float* samp, int64_t n,
                      int nthreads)
{
    mtx_parallel(n, nthreads, [&](int64_t b, int64_t e) {
        M::LayeredBsdf* B = M::layered_new();
        for (int64_t i = b; i < e; ++i) {
            const V3 wo = ld3m(V, n, i);
            M::layered_build(B, params_of(mat, prm, n, i, ld3m(N, n, i)), wo);
            if (eval) {
                M::Sample s     = M::layered_eval(B, wo, ld3m(L, n, i));
                eval[i]         = s.weight.x;
                eval[n + i]     = s.weight.y;
                eval[2 * n + i] = s.weight.z;
                eval[3 * n + i] = s.pdf;
            }
            if (samp) {
                for (int k = 0; k < 4; k++) {
                    M::Sample s = M::layered_sample(B, wo, rnd[(2 * k) * n + i],
                                                    rnd[(2 * k + 1) * n + i], 0.0f);
                    float* o    = samp + (int64_t)(7 * k) * n;
                    o[i]        = s.wi.x;
                    o[n + i]     = s.wi.y;
                    o[2 * n + i] = s.wi.z;
                    o[3 * n + i] = s.weight.x;
                    o[4 * n + i] = s.weight.y;
                    o[5 * n + i] = s.weight.z;
                    o[6 * n + i] = s.pdf;
                }
            }
        }
        M::layered_free(B);
    });
}

// Exact evaluation of surrogate queries: q 14 planes = wi xyz (local
// frame), wo x and z (wo.y = 0), coat ior, coat roughness, sheen
// roughness, F0 rgb, F90 rgb; N = (0, 0, 1). out: weight rgb, pdf.
void oref_mtx_eval_local(const float* mat, const float* q, float* out, int64_t n, int nthreads)
{
    mtx_parallel(n, nthreads, [&](int64_t b, int64_t e) {
        M::LayeredBsdf* B = M::layered_new();
        for (int64_t i = b; i < e; ++i) {
            const V3 wi(q[i], q[n + i], q[2 * n + i]);
            const V3 wo(q[3 * n + i], 0.0f, q[4 * n + i]);
            M::layered_build(B, params_of(mat, q + 5 * n, n, i, V3(0.0f, 0.0f, 1.0f)), wo);
            M::Sample s    = M::layered_eval(B, wo, wi);
            out[i]         = s.weight.x;
            out[n + i]     = s.weight.y;
            out[2 * n + i] = s.weight.z;
            out[3 * n + i] = s.pdf;
        }
        M::layered_free(B);
    });
}

// lobe weights (9: sheen rgb, coat rgb, base rgb) and selection pdfs (3)
void oref_mtx_weights(const float* mat, const float* N, const float* V, const float* prm,
                      float* w9, float* pdf3, int64_t n)
{
    M::LayeredBsdf* B = M::layered_new();
    for (int64_t i = 0; i < n; ++i) {
        M::layered_build(B, params_of(mat, prm, n, i, ld3m(N, n, i)), ld3m(V, n, i));
        float w[9], p[3];
        M::layered_weights(B, w, p);
        for (int k = 0; k < 9; k++)
            w9[k * n + i] = w[k];
        for (int k = 0; k < 3; k++)
            pdf3[k * n + i] = p[k];
    }
    M::layered_free(B);
}

void oref_mtx_luts(float* out)
{
    const float* t = M::luts();
    for (int i = 0; i < M::kLutFloats; i++)
        out[i] = t[i];
}

#if !defined(OREF_COUNTING) && !defined(OREF_F64)
void oref_mtx_bake(float* out, int nthreads) { M::bake_luts(out, nthreads); }
#endif

// Building blocks on AoS rows (tests): fn, row width in, width out
//  0 ggx_D(ax, ay, m3)            1 G1(ax, ay, w3)        2 G2_G1(ax, ay, wi3, wo3)
//  3 D_refl_D(ax, ay, wo3)        4 vndf(ax, ay, wo3, u, v) -> 3
//  5 bvndf(ax, ay, wo3, u, v) -> 3                       6 unit_disc(u, v) -> 2
//  7 uniform_hemisphere(u, v) -> 3                       8 fresnel(eta, c)
//  9 emiss(table, rough, fidx, c) 10 ck_albedo(a, c)      11 ck_D(a, m3)
// 12 pow(x, y)                    13 log2(x)              14 exp2(x)
int oref_mtx_fn(int fn, const float* in, float* out, int64_t n)
{
    static const int win[]  = { 5, 5, 8, 5, 7, 7, 2, 2, 2, 4, 2, 4, 2, 1, 1 };
    static const int wout[] = { 1, 1, 1, 1, 3, 3, 2, 3, 1, 1, 1, 1, 1, 1, 1 };
    if (fn < 0 || fn > 14)
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        const float* a = in + i * win[fn];
        float* o       = out + i * wout[fn];
        switch (fn) {
        case 0: o[0] = M::ggx_D(a[0], a[1], V3(a[2], a[3], a[4])); break;
        case 1: o[0] = M::ggx_G1(a[0], a[1], V3(a[2], a[3], a[4])); break;
        case 2:
            o[0] = M::ggx_G2_G1(a[0], a[1], V3(a[2], a[3], a[4]), V3(a[5], a[6], a[7]));
            break;
        case 3: o[0] = M::ggx_D_refl_D(a[0], a[1], V3(a[2], a[3], a[4]), V3(0.0f)); break;
        case 4:
        case 5: {
            V3 m = fn == 4 ? M::ggx_sample_vndf(a[0], a[1], V3(a[2], a[3], a[4]), a[5], a[6])
                           : M::ggx_sample_bvndf(a[0], a[1], V3(a[2], a[3], a[4]), a[5], a[6],
                                                 false);
            o[0] = m.x;
            o[1] = m.y;
            o[2] = m.z;
            break;
        }
        case 6: {
            M::V2 p = M::square_to_unit_disc(a[0], a[1]);
            o[0]    = p.x;
            o[1]    = p.y;
            break;
        }
        case 7: {
            V3 w = M::sample_uniform_hemisphere(a[0], a[1]);
            o[0] = w.x;
            o[1] = w.y;
            o[2] = w.z;
            break;
        }
        case 8: o[0] = M::dielectric_fresnel(a[0], a[1]); break;
        case 9: o[0] = M::emiss_eval((int)a[0] == 0 ? M::kLutMiniGGX : M::kLutReflFront, a[1],
                                     a[2], a[3]);
            break;
        case 10: o[0] = M::conty_kulla_albedo(a[0], a[1]); break;
        case 11: o[0] = M::conty_kulla_D(a[0], V3(a[1], a[2], a[3])); break;
        case 12: o[0] = oref::ofast_safe_powf(a[0], a[1]); break;
        case 13: o[0] = oref::ofast_log2f(a[0]); break;
        case 14: o[0] = oref::ofast_exp2f(a[0]); break;
        }
    }
    return 0;
}

}  // extern "C"
