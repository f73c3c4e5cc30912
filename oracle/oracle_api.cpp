// CPU oracle: C entry points used by tests/, the smoke check and the
// CPU-baseline leg of bench.py (through ctypes).
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// Array layout matches the GPU library: structure-of-arrays planes of n
// floats. For a Dual2 operand of k components the planes are
// [val_0..val_{k-1}, dx_0..dx_{k-1}, dy_0..dy_{k-1}].
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "closure.h"
#include "noise.h"
#include "omath.h"

using namespace oref;

namespace {

enum NoiseKind { kPerlin, kUPerlin, kCell, kHash, kSimplex, kUSimplex, kBad };

// Name mapping of llvm_gen_noise, src/liboslexec/batched_llvm_gen.cpp:6096-6134
NoiseKind noise_kind(const char* name)
{
    std::string s(name);
    if (s == "perlin" || s == "snoise")
        return kPerlin;
    if (s == "uperlin" || s == "noise")
        return kUPerlin;
    if (s == "cell" || s == "cellnoise")
        return kCell;
    if (s == "hash" || s == "hashnoise")
        return kHash;
    if (s == "simplex")
        return kSimplex;
    if (s == "usimplex")
        return kUSimplex;
    return kBad;
}

template<class F> void parallel_for(int64_t n, int nthreads, F f)
{
    if (nthreads <= 1 || n < 4096) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e)
            break;
        th.emplace_back([=]() { f(b, e); });
    }
    for (auto& t : th)
        t.join();
}

}  // namespace

extern "C" {

uint32_t oref_inthash(const uint32_t* k, int n) { return inthash_n(k, n); }
uint32_t oref_bjfinal(uint32_t a, uint32_t b, uint32_t c) { return bjfinal(a, b, c); }

// Signed perlin in the CGDefault (SIMD) interpolation order over SoA planes
// (outdim 1 or 3, indim 1..4, Dual2 planes when derivs).
int oref_perlin_cgdefault(int outdim, int indim, int derivs, const float* in, float* out,
                          int64_t n)
{
    if (indim < 1 || indim > 4 || (outdim != 1 && outdim != 3))
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        if (!derivs) {
            float p[4];
            for (int c = 0; c < indim; ++c)
                p[c] = in[c * n + i];
            for (int c = 0; c < outdim; ++c)
                out[c * n + i] = snoise_f_cgdefault(indim, p, c);
        } else {
            D2<float> p[4];
            for (int c = 0; c < indim; ++c)
                p[c] = D2<float>(in[c * n + i], in[(indim + c) * n + i], in[(2 * indim + c) * n + i]);
            for (int c = 0; c < outdim; ++c) {
                D2<float> r                   = snoise_df_cgdefault(indim, p, c);
                out[c * n + i]                = r.v;
                out[(outdim + c) * n + i]     = r.dx;
                out[(2 * outdim + c) * n + i] = r.dy;
            }
        }
    }
    return 0;
}

// Generic noise over SoA planes. Returns 0 on success, -1 on a bad request.
int oref_noise(const char* name, int outdim, int indim, int derivs,
               const float* in, float* out, int64_t n)
{
    NoiseKind kind = noise_kind(name);
    if (kind == kBad || indim < 1 || indim > 4 || (outdim != 1 && outdim != 3))
        return -1;
    // cellnoise and hashnoise have no derivatives (forced off, :6109-6116)
    bool dv = derivs && kind != kCell && kind != kHash;
    for (int64_t i = 0; i < n; ++i) {
        if (!dv) {
            float p[4];
            for (int c = 0; c < indim; ++c)
                p[c] = in[c * n + i];
            float r[3] = { 0, 0, 0 };
            if (outdim == 1) {
                switch (kind) {
                case kPerlin: r[0] = snoise_f(indim, p); break;
                case kUPerlin: r[0] = unoise_f(indim, p); break;
                case kCell: r[0] = cellnoise_f(indim, p); break;
                case kHash: r[0] = hashnoise_f(indim, p); break;
                case kSimplex: r[0] = simplex_f(indim, p, 0, nullptr); break;
                case kUSimplex:
                    r[0] = 0.5f * (simplex_f(indim, p, 0, nullptr) + 1.0f);
                    break;
                default: break;
                }
            } else {
                V3 v;
                switch (kind) {
                case kPerlin: v = snoise_v(indim, p); break;
                case kUPerlin: v = unoise_v(indim, p); break;
                case kCell: v = cellnoise_v(indim, p); break;
                case kHash: v = hashnoise_v(indim, p); break;
                case kSimplex:
                case kUSimplex:
                    v = V3(simplex_f(indim, p, 0, nullptr),
                           simplex_f(indim, p, 1, nullptr),
                           simplex_f(indim, p, 2, nullptr));
                    if (kind == kUSimplex)
                        v = V3(0.5f * (v.x + 1.0f), 0.5f * (v.y + 1.0f),
                               0.5f * (v.z + 1.0f));
                    break;
                default: break;
                }
                r[0] = v.x;
                r[1] = v.y;
                r[2] = v.z;
            }
            for (int c = 0; c < outdim; ++c)
                out[c * n + i] = r[c];
            if (derivs)
                for (int c = 0; c < 2 * outdim; ++c)
                    out[(outdim + c) * n + i] = 0.0f;
        } else {
            D2<float> p[4];
            for (int c = 0; c < indim; ++c)
                p[c] = D2<float>(in[c * n + i], in[(indim + c) * n + i],
                                 in[(2 * indim + c) * n + i]);
            if (outdim == 1) {
                D2<float> r;
                switch (kind) {
                case kPerlin: r = snoise_df(indim, p); break;
                case kUPerlin: r = unoise_df(indim, p); break;
                case kSimplex: r = simplex_df(indim, p, 0); break;
                case kUSimplex: {
                    // USimplexNoiseScalar: 0.5*(r+1) on val, 0.5*d on partials
                    D2<float> s = simplex_df(indim, p, 0);
                    r = D2<float>(0.5f * (s.v + 1.0f), 0.5f * s.dx, 0.5f * s.dy);
                    break;
                }
                default: break;
                }
                out[i]         = r.v;
                out[n + i]     = r.dx;
                out[2 * n + i] = r.dy;
            } else {
                D2<V3> r;
                switch (kind) {
                case kPerlin: r = snoise_dv(indim, p); break;
                case kUPerlin: r = unoise_dv(indim, p); break;
                case kSimplex: r = simplex_dv(indim, p); break;
                case kUSimplex: {
                    D2<V3> s = simplex_dv(indim, p);
                    r = D2<V3>(V3(0.5f * (s.v.x + 1.0f), 0.5f * (s.v.y + 1.0f),
                                  0.5f * (s.v.z + 1.0f)),
                               0.5f * s.dx, 0.5f * s.dy);
                    break;
                }
                default: break;
                }
                const float vals[9] = { r.v.x,  r.v.y,  r.v.z,  r.dx.x, r.dx.y,
                                        r.dx.z, r.dy.x, r.dy.y, r.dy.z };
                for (int c = 0; c < 9; ++c)
                    out[c * n + i] = vals[c];
            }
        }
    }
    return 0;
}

// Periodic noise over SoA planes: period holds indim planes of n floats
// (no derivatives). Names map as the periodic branch of llvm_gen_noise
// (batched_llvm_gen.cpp:6096-6134): perlin/snoise/psnoise -> psnoise,
// uperlin/noise/pnoise -> pnoise, cell -> pcellnoise, hash -> phashnoise;
// simplex has no periodic form (-1).
int oref_pnoise(const char* name, int outdim, int indim, int derivs, const float* in,
                const float* period, float* out, int64_t n)
{
    std::string s(name);
    NoiseKind kind = (s == "psnoise") ? kPerlin : (s == "pnoise") ? kUPerlin : noise_kind(name);
    if (kind == kBad || kind == kSimplex || kind == kUSimplex || indim < 1 || indim > 4
        || (outdim != 1 && outdim != 3))
        return -1;
    bool dv = derivs && kind != kCell && kind != kHash;
    for (int64_t i = 0; i < n; ++i) {
        float per[4];
        for (int c = 0; c < indim; ++c)
            per[c] = period[c * n + i];
        if (!dv) {
            float p[4];
            for (int c = 0; c < indim; ++c)
                p[c] = in[c * n + i];
            V3 v(0.0f);
            if (outdim == 1) {
                switch (kind) {
                case kPerlin: v.x = psnoise_f(indim, p, per); break;
                case kUPerlin: v.x = pnoise_f(indim, p, per); break;
                case kCell: v.x = pcellnoise_f(indim, p, per); break;
                default: v.x = phashnoise_f(indim, p, per); break;
                }
            } else {
                switch (kind) {
                case kPerlin: v = psnoise_v(indim, p, per); break;
                case kUPerlin: v = pnoise_v(indim, p, per); break;
                case kCell: v = pcellnoise_v(indim, p, per); break;
                default: v = phashnoise_v(indim, p, per); break;
                }
            }
            const float r[3] = { v.x, v.y, v.z };
            for (int c = 0; c < outdim; ++c)
                out[c * n + i] = r[c];
            if (derivs)
                for (int c = 0; c < 2 * outdim; ++c)
                    out[(outdim + c) * n + i] = 0.0f;
        } else {
            D2<float> p[4];
            for (int c = 0; c < indim; ++c)
                p[c] = D2<float>(in[c * n + i], in[(indim + c) * n + i],
                                 in[(2 * indim + c) * n + i]);
            if (outdim == 1) {
                D2<float> r = kind == kPerlin ? psnoise_df(indim, p, per) : pnoise_df(indim, p, per);
                out[i]         = r.v;
                out[n + i]     = r.dx;
                out[2 * n + i] = r.dy;
            } else {
                D2<V3> r = kind == kPerlin ? psnoise_dv(indim, p, per) : pnoise_dv(indim, p, per);
                const float vals[9] = { r.v.x,  r.v.y,  r.v.z,  r.dx.x, r.dx.y,
                                        r.dx.z, r.dy.x, r.dy.y, r.dy.z };
                for (int c = 0; c < 9; ++c)
                    out[c * n + i] = vals[c];
            }
        }
    }
    return 0;
}

int oref_hash_op(int indim, const float* in, int32_t* out, int64_t n)
{
    if (indim < 1 || indim > 4)
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        float p[4];
        for (int c = 0; c < indim; ++c)
            p[c] = in[c * n + i];
        out[i] = hash_op_f(indim, p);
    }
    return 0;
}

void oref_smoothstep(const float* e0, const float* e1, const float* x,
                     float* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        out[i] = smoothstep_f(e0[i], e1[i], x[i]);
}

// The oracle's fixed transcendentals: fn 0 exp, 1 sin, 2 cos, 3 cbrt.
void oref_math(int fn, const float* in, float* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        float s, c;
        switch (fn) {
        case 0: out[i] = ofast_expf(in[i]); break;
        case 1: ofast_sincosf(in[i], &s, &c); out[i] = s; break;
        case 2: ofast_sincosf(in[i], &s, &c); out[i] = c; break;
        case 4: out[i] = ofast_logf(in[i]); break;
        case 5: out[i] = ofast_erff(in[i]); break;
        case 6: out[i] = ofast_ierff(in[i]); break;
        default: out[i] = ofast_cbrtf(in[i]); break;
        }
    }
}

// A lobe from a row of 9 floats (kind, weight rgb, N xyz, alpha, eta), 13
// (the same, then the tangent U xyz and yalpha of an anisotropic microfacet
// lobe) or 14 (then the refract mode 0 / 1 / 2).
static Lobe lobe_row(const float* r, int cols)
{
    Lobe l;
    l.kind   = (int)r[0];
    l.weight = V3(r[1], r[2], r[3]);
    l.N      = V3(r[4], r[5], r[6]);
    l.alpha  = r[7];
    l.eta    = r[8];
    if (cols >= 13) {
        l.aniso = true;
        l.U     = V3(r[9], r[10], r[11]);
        l.ay    = r[12];
    }
    if (cols >= 14)
        l.refract = (int)r[13];
    return l;
}

// CompositeBSDF prepare + eval over one lobe list for n direction pairs.
// lobes: nlobes rows of cols (9 or 13) floats, see lobe_row; wo, wi: 3
// planes of n; out: 4 planes (weight rgb, pdf).
void oref_composite_eval(int nlobes, const float* lobes, const float* wo, const float* wi,
                         float* out, int64_t n, int cols)
{
    Composite c;
    for (int k = 0; k < nlobes; ++k)
        c.add(lobe_row(lobes + cols * k, cols));
    for (int64_t i = 0; i < n; ++i) {
        Composite ci = c;
        const V3 o(wo[i], wo[n + i], wo[2 * n + i]);
        const V3 w(wi[i], wi[n + i], wi[2 * n + i]);
        ci.prepare(o);
        const BsdfSample s = ci.eval(o, w);
        out[i]             = s.weight.x;
        out[n + i]         = s.weight.y;
        out[2 * n + i]     = s.weight.z;
        out[3 * n + i]     = s.pdf;
    }
}

// CompositeBSDF prepare + sample over one lobe list: wo 3 planes, rnd 3
// planes (rx, ry, rz); out 7 planes (wi xyz, weight rgb, pdf).
void oref_composite_sample(int nlobes, const float* lobes, const float* wo, const float* rnd,
                           float* out, int64_t n, int cols)
{
    Composite c;
    for (int k = 0; k < nlobes; ++k)
        c.add(lobe_row(lobes + cols * k, cols));
    for (int64_t i = 0; i < n; ++i) {
        Composite ci = c;
        const V3 o(wo[i], wo[n + i], wo[2 * n + i]);
        ci.prepare(o);
        const BsdfDirSample s = ci.sample(o, rnd[i], rnd[n + i], rnd[2 * n + i]);
        const float v[7]      = { s.wi.x, s.wi.y, s.wi.z, s.weight.x, s.weight.y, s.weight.z,
                                  s.pdf };
        for (int c7 = 0; c7 < 7; ++c7)
            out[c7 * n + i] = v[c7];
    }
}

}  // extern "C"
