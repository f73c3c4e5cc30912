// Device runtime for emitted group kernels.
//
// The group emitter lowers each OSO op to a call of one of the functions
// below (or of the shadeop library: oslmi_noise.h, oslmi_gabor.h,
// oslmi_color.h, oslmi_closure.h). Symbols are registers of type int,
// float, V3 (point/vector/normal/color), their Dual2 forms Df/Dv when the
// derivative analysis says they carry derivatives, M44 (matrix) or Closure
// (a flattened lobe list). Mixed-type operands are promoted by the emitter
// before the call (float -> triple broadcast, value -> Dual2 with zero
// derivatives), matching the reference's batched code generation
// (src/liboslexec/batched_llvm_gen.cpp).
#pragma once
#include "oslmi_closure.h"
#include "oslmi_color.h"
#include "oslmi_gabor.h"
#include "oslmi_math.h"
#include "oslmi_matrix.h"
#include "oslmi_noise.h"

namespace oslmi {
namespace rt {

// ---------------------------------------------------------- promotion
OSLMI_DEV float val(float x) { return x; }
OSLMI_DEV float val(const Df& x) { return x.v; }
OSLMI_DEV V3 val(V3 x) { return x; }
OSLMI_DEV V3 val(const Dv& x) { return x.v; }
OSLMI_DEV Df dual(float x) { return Df(x); }
OSLMI_DEV Df dual(const Df& x) { return x; }
OSLMI_DEV Dv dual(V3 x) { return Dv(x); }
OSLMI_DEV Dv dual(const Dv& x) { return x; }
OSLMI_DEV V3 tri(float x) { return V3(x, x, x); }
OSLMI_DEV V3 tri(V3 x) { return x; }
OSLMI_DEV Dv tri(const Df& x) { return Dv(V3(x.v), V3(x.dx), V3(x.dy)); }
OSLMI_DEV Dv tri(const Dv& x) { return x; }
OSLMI_DEV float tofloat(int x) { return (float)x; }
OSLMI_DEV float tofloat(float x) { return x; }

// ---------------------------------------------------------- arithmetic
OSLMI_DEV float add(float a, float b) { return a + b; }
OSLMI_DEV V3 add(V3 a, V3 b) { return a + b; }
OSLMI_DEV Df add(const Df& a, const Df& b) { return a + b; }
OSLMI_DEV Dv add(const Dv& a, const Dv& b) { return a + b; }
OSLMI_DEV int add(int a, int b) { return a + b; }
OSLMI_DEV float sub(float a, float b) { return a - b; }
OSLMI_DEV V3 sub(V3 a, V3 b) { return a - b; }
OSLMI_DEV Df sub(const Df& a, const Df& b) { return a - b; }
OSLMI_DEV Dv sub(const Dv& a, const Dv& b) { return a - b; }
OSLMI_DEV int sub(int a, int b) { return a - b; }
OSLMI_DEV float mul(float a, float b) { return a * b; }
OSLMI_DEV V3 mul(V3 a, V3 b) { return a * b; }
OSLMI_DEV int mul(int a, int b) { return a * b; }
OSLMI_DEV Df mul(const Df& a, const Df& b) { return a * b; }
// componentwise product rule for triples with derivatives
OSLMI_DEV Dv mul(const Dv& a, const Dv& b)
{
    return Dv(a.v * b.v, a.v * b.dx + a.dx * b.v, a.v * b.dy + a.dy * b.v);
}
OSLMI_DEV float div(float a, float b) { return a / b; }
OSLMI_DEV V3 div(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
OSLMI_DEV int div(int a, int b) { return b == 0 ? 0 : a / b; }
OSLMI_DEV Df div(const Df& a, const Df& b) { return a / b; }
OSLMI_DEV Dv div(const Dv& a, const Dv& b)
{
    Df x = comp(a, 0) / comp(b, 0), y = comp(a, 1) / comp(b, 1), z = comp(a, 2) / comp(b, 2);
    return make_dv(x, y, z);
}
OSLMI_DEV float neg(float a) { return -a; }
OSLMI_DEV V3 neg(V3 a) { return -a; }
OSLMI_DEV Df neg(const Df& a) { return -a; }
OSLMI_DEV Dv neg(const Dv& a) { return -a; }
OSLMI_DEV int neg(int a) { return -a; }

// OSL mod: a - b*floor(a/b) (safe: 0 when b == 0)
// the mod op on floats: the exact remainder of a / b with the sign of a
// (LLVM frem, batched_llvm_gen.cpp:1683-1750), 0 when b is 0
OSLMI_DEV float fmod_osl(float a, float b) { return b != 0.0f ? fmodf(a, b) : 0.0f; }

// ---------------------------------------------------------- comparisons
template<class A, class B> OSLMI_DEV int lt(A a, B b) { return val(a) < val(b); }
template<class A, class B> OSLMI_DEV int gt(A a, B b) { return val(a) > val(b); }
template<class A, class B> OSLMI_DEV int le(A a, B b) { return val(a) <= val(b); }
template<class A, class B> OSLMI_DEV int ge(A a, B b) { return val(a) >= val(b); }
OSLMI_DEV int lt(int a, int b) { return a < b; }
OSLMI_DEV int gt(int a, int b) { return a > b; }
OSLMI_DEV int le(int a, int b) { return a <= b; }
OSLMI_DEV int ge(int a, int b) { return a >= b; }
OSLMI_DEV int eq(float a, float b) { return a == b; }
OSLMI_DEV int eq(int a, int b) { return a == b; }
OSLMI_DEV int neq(float a, float b) { return a != b; }
OSLMI_DEV int neq(int a, int b) { return a != b; }

// ---------------------------------------------------------- mix / smoothstep
// llvm_gen_mix (batched_llvm_gen.cpp:1868-1960)
OSLMI_DEV float mix(float a, float b, float x) { return a * (1.0f - x) + b * x; }
OSLMI_DEV V3 mix(V3 a, V3 b, V3 x)
{
    return V3(mix(a.x, b.x, x.x), mix(a.y, b.y, x.y), mix(a.z, b.z, x.z));
}
OSLMI_DEV Df mix(const Df& a, const Df& b, const Df& x)
{
    float omx = 1.0f - x.v;
    float rx  = ((a.dx * omx - a.v * x.dx) + b.v * x.dx) + b.dx * x.v;
    float ry  = ((a.dy * omx - a.v * x.dy) + b.v * x.dy) + b.dy * x.v;
    return Df(a.v * omx + b.v * x.v, rx, ry);
}
OSLMI_DEV Dv mix(const Dv& a, const Dv& b, const Dv& x)
{
    return make_dv(mix(comp(a, 0), comp(b, 0), comp(x, 0)), mix(comp(a, 1), comp(b, 1), comp(x, 1)),
                   mix(comp(a, 2), comp(b, 2), comp(x, 2)));
}
// smoothstep on floats / duals: oslmi_noise.h; triples componentwise
OSLMI_DEV V3 smoothstep(V3 e0, V3 e1, V3 x)
{
    return V3(oslmi::smoothstep(e0.x, e1.x, x.x), oslmi::smoothstep(e0.y, e1.y, x.y),
              oslmi::smoothstep(e0.z, e1.z, x.z));
}

// min/max/clamp: max(min(x, hi), lo) (stdosl.h:140-142); duals select whole
OSLMI_DEV float fmin_(float a, float b) { return b < a ? b : a; }
OSLMI_DEV float fmax_(float a, float b) { return a < b ? b : a; }
OSLMI_DEV float min_(float a, float b) { return fminf(a, b); }
OSLMI_DEV float max_(float a, float b) { return fmaxf(a, b); }
OSLMI_DEV int min_(int a, int b) { return a < b ? a : b; }
OSLMI_DEV int max_(int a, int b) { return a > b ? a : b; }
OSLMI_DEV V3 min_(V3 a, V3 b) { return V3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
OSLMI_DEV V3 max_(V3 a, V3 b) { return V3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
OSLMI_DEV Df min_(const Df& a, const Df& b) { return b.v < a.v ? b : a; }
OSLMI_DEV Df max_(const Df& a, const Df& b) { return a.v < b.v ? b : a; }
template<class T> OSLMI_DEV T clamp(const T& x, const T& lo, const T& hi) { return max_(min_(x, hi), lo); }

// ---------------------------------------------------------- math
OSLMI_DEV float floor_(float x) { return floorf(x); }
OSLMI_DEV V3 floor_(V3 x) { return V3(floorf(x.x), floorf(x.y), floorf(x.z)); }
OSLMI_DEV float abs_(float x) { return fabsf(x); }
OSLMI_DEV int abs_(int x) { return x < 0 ? -x : x; }
OSLMI_DEV float sqrt_(float x) { return x > 0.0f ? sqrtf(x) : 0.0f; }
OSLMI_DEV Df sqrt_(const Df& a)
{
    if (a.v > 0.0f) {
        float f = sqrtf(a.v), df = 0.5f / f;
        return Df(f, df * a.dx, df * a.dy);
    }
    return Df(0.0f);
}
OSLMI_DEV float exp_(float x) { return fexp(x); }
OSLMI_DEV Df exp_(const Df& a) { return dexp(a); }
OSLMI_DEV float sin_(float x) { float s, c; fsincos(x, &s, &c); return s; }
OSLMI_DEV float cos_(float x) { float s, c; fsincos(x, &s, &c); return c; }
OSLMI_DEV Df cos_(const Df& a) { return dcos(a); }
OSLMI_DEV Df sin_(const Df& a)
{
    float s, c;
    fsincos(a.v, &s, &c);
    return Df(s, c * a.dx, c * a.dy);
}
OSLMI_DEV float dot_(V3 a, V3 b) { return dot(a, b); }
OSLMI_DEV Df dot_(const Dv& a, const Dv& b) { return ddot(a, b); }
OSLMI_DEV float length_(V3 a) { return sqrtf(dot(a, a)); }
OSLMI_DEV V3 normalize_(V3 a) { return vnormalized(a); }
OSLMI_DEV V3 cross_(V3 a, V3 b) { return cross(a, b); }
// filterwidth (llvm_ops.cpp:946-956): sqrt(dx^2 + dy^2)
OSLMI_DEV float filterwidth(const Df& x) { return sqrtf(x.dx * x.dx + x.dy * x.dy); }
OSLMI_DEV V3 filterwidth(const Dv& x)
{
    return V3(sqrtf(x.dx.x * x.dx.x + x.dy.x * x.dy.x), sqrtf(x.dx.y * x.dx.y + x.dy.y * x.dy.y),
              sqrtf(x.dx.z * x.dx.z + x.dy.z * x.dy.z));
}
OSLMI_DEV float Dx(const Df& x) { return x.dx; }
OSLMI_DEV float Dy(const Df& x) { return x.dy; }
OSLMI_DEV V3 Dx(const Dv& x) { return x.dx; }
OSLMI_DEV V3 Dy(const Dv& x) { return x.dy; }

// ---------------------------------------------------------- components
OSLMI_DEV float compref(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
OSLMI_DEV Df compref(const Dv& a, int i) { return comp(a, i); }
OSLMI_DEV void compassign(V3& a, int i, float x)
{
    if (i == 0) a.x = x; else if (i == 1) a.y = x; else a.z = x;
}
OSLMI_DEV void compassign(Dv& a, int i, const Df& x)
{
    if (i == 0) { a.v.x = x.v; a.dx.x = x.dx; a.dy.x = x.dy; }
    else if (i == 1) { a.v.y = x.v; a.dx.y = x.dx; a.dy.y = x.dy; }
    else { a.v.z = x.v; a.dx.z = x.dx; a.dy.z = x.dy; }
}
OSLMI_DEV V3 triple(float a, float b, float c) { return V3(a, b, c); }
OSLMI_DEV Dv triple(const Df& a, const Df& b, const Df& c) { return make_dv(a, b, c); }

// ---------------------------------------------------------- matrices
// M44 and the point / vector / normal transforms: oslmi_matrix.h

// ---------------------------------------------------------- closures
// A closure value is a LobeList whose capacity the emitter bounds per
// group (the number of closure constructors it can reach, at most 8);
// primitives make one-lobe lists, ADD concatenates, MUL scales.
typedef LobeList<8> Closure;
#ifdef OSLMI_CLOSURE_EXPORT
// the exported parameters of a lobe, in testrender's registration order
// (src/testrender/shading.cpp:185-296; MTX/bsdf_*_decl.h entry()),
// zero-padded
OSLMI_DEV void set_xp(Lobe& l, const float* v, int k)
{
#pragma unroll
    for (int j = 0; j < OSLMI_CLOSURE_PARAMS; j++)
        l.xp[j] = j < k ? v[j] : 0.0f;
}
#endif
#ifdef OSLMI_MX_CLOSURES
// a component (not the null closure); as the top of a layer, a lobe other
// than a MaterialX one is opaque (evaluate_layer_opacity's default case)
OSLMI_DEV void mark_component(LobeList<1>& c)
{
    c.nn = 1;
    c.op = V3(1.0f);
}
#endif
OSLMI_DEV LobeList<1> closure_diffuse(V3 N)
{
    LobeList<1> c;
    Lobe l;
    l.id = kClosureDiffuse; l.weight = V3(1.0f); l.N = N; l.alpha = 0.0f; l.eta = 0.0f;
#ifdef OSLMI_CLOSURE_EXPORT
    {
        const float v[3] = { N.x, N.y, N.z };  // DiffuseParams: N
        set_xp(l, v, 3);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
// translucent(N): testrender's Diffuse<1>, a diffuse lobe about -N
OSLMI_DEV LobeList<1> closure_translucent(V3 N)
{
    LobeList<1> c;
    Lobe l;
    l.id = kClosureTranslucent; l.weight = V3(1.0f); l.N = -N; l.alpha = 0.0f; l.eta = 0.0f;
#ifdef OSLMI_CLOSURE_EXPORT
    {
        const float v[3] = { N.x, N.y, N.z };  // DiffuseParams: N as given
        set_xp(l, v, 3);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
// phong(N, exponent): testrender's Phong
OSLMI_DEV LobeList<1> closure_phong(V3 N, float exponent)
{
    LobeList<1> c;
    Lobe l;
    l.id = kClosurePhong; l.weight = V3(1.0f); l.N = N; l.alpha = exponent; l.eta = 0.0f;
#ifdef OSLMI_CLOSURE_EXPORT
    {
        const float v[4] = { N.x, N.y, N.z, exponent };  // PhongParams: N, exponent
        set_xp(l, v, 4);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
// reflection(N) (eta 0: a perfect mirror), reflection(N, eta),
// refraction(N, eta) and transparent(): testrender's delta lobes; also
// emission() (no parameters)
OSLMI_DEV LobeList<1> closure_delta(int id, V3 N, float eta)
{
    LobeList<1> c;
    Lobe l;
    l.id = id; l.weight = V3(1.0f); l.N = N; l.alpha = 0.0f; l.eta = eta;
#ifdef OSLMI_CLOSURE_EXPORT
    {
        // ReflectionParams / RefractionParams: N, eta (none for transparent)
        const float v[4] = { N.x, N.y, N.z, eta };
        set_xp(l, v, (id == kClosureTransparent || id == kClosureEmission) ? 0
                     : id == kClosureReflection                              ? 3
                                                                             : 4);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
OSLMI_DEV LobeList<1> closure_microfacet_ggx(V3 N, float alpha, float eta)
{
    LobeList<1> c;
    Lobe l;
    l.id = kClosureMicrofacetGGX; l.weight = V3(1.0f); l.N = N; l.alpha = alpha; l.eta = eta;
#ifdef OSLMI_FULL_MICROFACET
    l.U       = V3(0.0f);
    l.ay      = alpha;
    l.refract = 0;
#endif
#ifdef OSLMI_CLOSURE_EXPORT
    {
        // MicrofacetParams: dist (0 ggx, 1 beckmann), N, U, xalpha, yalpha,
        // eta, refract
        const float v[11] = { 0.0f, N.x, N.y, N.z, 0.0f, 0.0f, 0.0f, alpha, alpha, eta, 0.0f };
        set_xp(l, v, 11);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
OSLMI_DEV LobeList<1> closure_microfacet_beckmann(V3 N, float alpha, float eta)
{
    LobeList<1> c;
    Lobe l;
    l.id = kClosureMicrofacetBeckmann; l.weight = V3(1.0f); l.N = N; l.alpha = alpha; l.eta = eta;
#ifdef OSLMI_FULL_MICROFACET
    l.U       = V3(0.0f);
    l.ay      = alpha;
    l.refract = 0;
#endif
#ifdef OSLMI_CLOSURE_EXPORT
    {
        // MicrofacetParams: dist (0 ggx, 1 beckmann), N, U, xalpha, yalpha,
        // eta, refract
        const float v[11] = { 1.0f, N.x, N.y, N.z, 0.0f, 0.0f, 0.0f, alpha, alpha, eta, 0.0f };
        set_xp(l, v, 11);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
#ifdef OSLMI_FULL_MICROFACET
// microfacet(dist, N, U, xalpha, yalpha, eta, refract) in general: a
// tangent, two roughnesses, refraction (1) or both reflection and
// refraction (2): testrender's Microfacet<Dist, refract> in the frame
// from_normal_and_tangent(N, U) (src/testrender/shading.cpp:570-800)
OSLMI_DEV LobeList<1> closure_microfacet_full(int id, V3 N, V3 U, float xalpha, float yalpha,
                                              float eta, int refract)
{
    LobeList<1> c;
    Lobe l;
    l.id = id; l.weight = V3(1.0f); l.N = N; l.alpha = xalpha; l.eta = eta;
    l.U       = U;
    l.ay      = yalpha;
    l.refract = refract;
#ifdef OSLMI_CLOSURE_EXPORT
    {
        const float v[11] = { id == kClosureMicrofacetBeckmann ? 1.0f : 0.0f, N.x, N.y, N.z, U.x,
                              U.y, U.z, xalpha, yalpha, eta, (float)refract };
        set_xp(l, v, 11);
    }
#endif
    c.add(l);
#ifdef OSLMI_MX_CLOSURES
    mark_component(c);
#endif
    return c;
}
#endif
// color * closure (a MUL node): scales every lobe weight. A zero weight
// gives the null closure (osl_mul_closure_color, opclosure.cpp:40-60); the
// layer walks see the difference, lobe lists without MaterialX closures
// do not (a zero-weight lobe adds nothing to an eval or a sample), so only
// kernels with MaterialX closures pay for the test.
template<int K> OSLMI_DEV LobeList<K> mul(const LobeList<K>& c, V3 w)
{
    LobeList<K> r = c;
#ifdef OSLMI_MX_CLOSURES
    if (w.x == 0.0f && w.y == 0.0f && w.z == 0.0f)
        return LobeList<K>();
#endif
#pragma unroll
    for (int i = 0; i < K; ++i)
        if (i < r.n)
            r.l[i].weight = r.l[i].weight * w;
#ifdef OSLMI_MX_CLOSURES
    r.op = w * r.op;
#endif
    return r;
}
template<int K> OSLMI_DEV LobeList<K> mul(V3 w, const LobeList<K>& c) { return mul(c, w); }
// ADD node: lobes of a then b (depth-first order of process_bsdf_closure);
// a null operand gives the other (osl_add_closure_closure)
template<int KA, int KB>
OSLMI_DEV LobeList<(KA + KB < 8 ? KA + KB : 8)> add(const LobeList<KA>& a, const LobeList<KB>& b)
{
    LobeList<(KA + KB < 8 ? KA + KB : 8)> r(a);
#ifdef OSLMI_MX_CLOSURES
    // the walks visit A, then B (queued), unless they end inside A
    if (!a.nn)
        return LobeList<(KA + KB < 8 ? KA + KB : 8)>(b);
    if (!b.nn)
        return r;
    if (!a.tr) {
#pragma unroll
        for (int i = 0; i < KB; ++i)
            if (i < b.n)
                r.add(b.l[i]);
        r.tr = b.tr;
    }
    if (!a.opend) {
        r.op    = b.op;
        r.opend = b.opend;
    }
#else
#pragma unroll
    for (int i = 0; i < KB; ++i)
        if (i < b.n)
            r.add(b.l[i]);
#endif
    return r;
}

#ifdef OSLMI_MX_CLOSURES
// MaterialX closures (stdosl.h dielectric_bsdf, generalized_schlick_bsdf,
// sheen_bsdf), built at wo with the energy tables T. Their opacity as the
// top of a layer is 1 - filter_o (evaluate_layer_opacity,
// src/testrender/shading.cpp:1242-1268).
OSLMI_DEV LobeList<1> closure_mx_lobe(int id, V3 N)
{
    LobeList<1> c;
    Lobe l;
    l.id     = id;
    l.weight = V3(1.0f);
    l.N      = N;
    l.alpha  = 0.0f;
    l.eta    = 0.0f;
    c.add(l);
    c.nn = 1;
    return c;
}
OSLMI_DEV LobeList<1> closure_mx_dielectric(V3 N, V3 U, V3 refl_tint, float rx, float ry, float ior,
                                            V3 wo, const float* T)
{
    LobeList<1> c = closure_mx_lobe(kClosureMxDielectric, N);
    mtx::mx_micro(&c.l[0].mx, mtx::kMxDielectric, N, U, refl_tint, rx, ry, ior, V3(0.0f), V3(0.0f),
                  0.0f, wo, T);
#ifdef OSLMI_CLOSURE_EXPORT
    {
        // dielectric_bsdf: N, U, reflection tint, transmission tint (0),
        // roughness_x, roughness_y, ior
        const float v[15] = { N.x, N.y, N.z, U.x, U.y, U.z, refl_tint.x, refl_tint.y, refl_tint.z,
                              0.0f, 0.0f, 0.0f, rx, ry, ior };
        set_xp(c.l[0], v, 15);
    }
#endif
    c.op = V3(1.0f) * (V3(1.0f) - V3(mtx::mx_filter_o(c.l[0].mx)));
    return c;
}
OSLMI_DEV LobeList<1> closure_mx_schlick(V3 N, V3 U, V3 refl_tint, float rx, float ry, V3 F0, V3 F90,
                                         float exponent, V3 wo, const float* T)
{
    LobeList<1> c = closure_mx_lobe(kClosureMxSchlick, N);
    mtx::mx_micro(&c.l[0].mx, mtx::kMxSchlick, N, U, refl_tint, rx, ry, 0.0f, F0, F90, exponent, wo,
                  T);
#ifdef OSLMI_CLOSURE_EXPORT
    {
        // generalized_schlick_bsdf: N, U, reflection tint, transmission
        // tint (0), roughness_x, roughness_y, f0, f90, exponent
        const float v[21] = { N.x,  N.y,  N.z,  U.x,  U.y,   U.z,   refl_tint.x, refl_tint.y,
                              refl_tint.z, 0.0f, 0.0f, 0.0f, rx, ry, F0.x, F0.y, F0.z, F90.x,
                              F90.y, F90.z, exponent };
        set_xp(c.l[0], v, 21);
    }
#endif
    c.op = V3(1.0f) * (V3(1.0f) - V3(mtx::mx_filter_o(c.l[0].mx)));
    return c;
}
OSLMI_DEV LobeList<1> closure_mx_sheen(V3 N, V3 albedo, float roughness, V3 wo)
{
    LobeList<1> c = closure_mx_lobe(kClosureMxSheen, N);
    mtx::mx_sheen(&c.l[0].mx, N, albedo, roughness, wo);
#ifdef OSLMI_CLOSURE_EXPORT
    {
        // sheen_bsdf: N, albedo, roughness, mode (0)
        const float v[8] = { N.x, N.y, N.z, albedo.x, albedo.y, albedo.z, roughness, 0.0f };
        set_xp(c.l[0], v, 8);
    }
#endif
    c.op = V3(1.0f) * (V3(1.0f) - V3(mtx::mx_filter_o(c.l[0].mx)));
    return c;
}
// layer(top, base): process_bsdf_closure's MX_LAYER_ID
// (src/testrender/shading.cpp:1645-1662): the top's lobes as they are, the
// base's scaled by 1 - clamp(opacity(top), 0, 1) and left out when that is
// black. As the top of another layer, its opacity is the base's (walked
// last), or the layer's own weight when the base is null.
OSLMI_DEV float clamp01_nan0(float v)
{
    if (!(0.0f <= v))
        v = 0.0f;
    if (!(v <= 1.0f))
        v = 1.0f;
    return v;
}
template<int KT, int KB>
OSLMI_DEV LobeList<(KT + KB < 8 ? KT + KB : 8)> layer(const LobeList<KT>& top, const LobeList<KB>& base)
{
    LobeList<(KT + KB < 8 ? KT + KB : 8)> r(top);
    r.nn = 1;
    // process_bsdf_closure: the base is queued (when its weight is not
    // black) and visited after the top, unless the walk ended in the top
    const V3 o  = top.nn ? top.op : V3(0.0f);
    const V3 bw = V3(1.0f) - V3(clamp01_nan0(o.x), clamp01_nan0(o.y), clamp01_nan0(o.z));
    r.tr        = 0;
    if (!top.tr && !(bw.x == 0.0f && bw.y == 0.0f && bw.z == 0.0f)) {
        if (!base.nn)
            r.tr = 1;  // a queued null base ends the walk
#pragma unroll
        for (int i = 0; i < KB; ++i)
            if (i < base.n) {
                Lobe x   = base.l[i];
                x.weight = bw * x.weight;
                r.add(x);
            }
        if (base.nn)
            r.tr = base.tr;
    } else if (top.tr) {
        r.tr = 1;
    }
    // evaluate_layer_opacity: the base is always queued; its opacity, or
    // the layer's own weight when it is null (the walk ends there)
    if (top.nn && top.opend) {
        r.op    = top.op;
        r.opend = 1;
    } else if (base.nn) {
        r.op    = base.op;
        r.opend = base.opend;
    } else {
        r.op    = V3(1.0f);
        r.opend = 1;
    }
    return r;
}
#endif

}  // namespace rt
}  // namespace oslmi
