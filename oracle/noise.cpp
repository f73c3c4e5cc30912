// CPU oracle: integer hashing and lattice noises.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
#include "noise.h"

namespace oref {

uint32_t inthash_n(const uint32_t* k, int n)
{
    // reference_inthash<N>, src/include/OSL/oslnoise.h:192-215
    uint32_t a, b, c;
    const unsigned seed = 0xdeadbeefu + ((unsigned)n << 2) + 13u;  // a constant per N
    uint32_t len        = (uint32_t)n;
    a = b = c = seed;
    while (len > 3) {
        a += k[0];
        b += k[1];
        c += k[2];
        bjmix(a, b, c);
        len -= 3;
        k += 3;
    }
    switch (len) {
    case 3: c += k[2];  // fallthrough
    case 2: b += k[1];  // fallthrough
    case 1:
        a += k[0];
        c = bjfinal(a, b, c);
        break;
    default: break;
    }
    return c;
}

// ---------------------------------------------------------------------------
// cell / hash noise: oslnoise.h:324-480 (IntHashNoiseBase, CellNoise,
// HashNoise). The Vec3 variant hashes (keys..., 0/1/2); for 3 and 4 keys the
// reference shares the bjmix (inthashVec overloads at :412-450), which gives
// the same values as the generic form used here.
static uint32_t cell_key(float v) { return (uint32_t)ifloor(v); }
static uint32_t hash_key(float v) { return f2u(v); }

template<uint32_t (*KEY)(float)>
static float intnoise_f(int dim, const float* p)
{
    uint32_t k[4];
    for (int i = 0; i < dim; ++i)
        k[i] = KEY(p[i]);
    return bits_to_01(inthash_n(k, dim));
}

template<uint32_t (*KEY)(float)>
static V3 intnoise_v(int dim, const float* p)
{
    uint32_t k[5];
    for (int i = 0; i < dim; ++i)
        k[i] = KEY(p[i]);
    float r[3];
    for (uint32_t c = 0; c < 3; ++c) {
        k[dim] = c;
        r[c]   = bits_to_01(inthash_n(k, dim + 1));
    }
    return V3(r[0], r[1], r[2]);
}

float cellnoise_f(int dim, const float* p) { return intnoise_f<cell_key>(dim, p); }
V3 cellnoise_v(int dim, const float* p) { return intnoise_v<cell_key>(dim, p); }
float hashnoise_f(int dim, const float* p) { return intnoise_f<hash_key>(dim, p); }
V3 hashnoise_v(int dim, const float* p) { return intnoise_v<hash_key>(dim, p); }

int hash_op_f(int dim, const float* p)
{
    // inthashf overloads, oslnoise.h:565-610
    uint32_t k[4];
    for (int i = 0; i < dim; ++i)
        k[i] = f2u(p[i]);
    return (int)inthash_n(k, dim);
}

int hash_op_i(int x) { return (int)inthash((uint32_t)x); }  // inthashi :555

// ---------------------------------------------------------------------------
// perlin, CGScalar branches
//
// Helpers follow oslnoise.h: floorfrac :872-904, fade :909-912,
// grad :926-1032, scale1..4 :1034-1041, HashScalar :1044-1090,
// sliceup/HashVector :1092-1155. Interpolation uses OIIO's lerp/bilerp/
// trilerp forms  v0*(1-x)+v1*x  and the nested (1-t)*(...)+t*(...)
// expansion; OIIO's source is not part of the reference tree.

template<class T> struct Ops;
template<> struct Ops<float> {
    static float one() { return 1.0f; }
    static float c(float x) { return x; }
    static float neg_if(float v, bool b)
    {
        OREF_COUNT(1);
        return b ? -v : v;
    }
    static float sel(bool b, float t, float f)
    {
        OREF_COUNT(1);
        return b ? t : f;
    }
};
template<> struct Ops<D2<float>> {
    static D2<float> one() { return D2<float>(1.0f); }
    static D2<float> c(float x) { return D2<float>(x); }
    static D2<float> neg_if(const D2<float>& v, bool b)
    {
        OREF_COUNT(3);
        return b ? -v : v;
    }
    static D2<float> sel(bool b, const D2<float>& t, const D2<float>& f)
    {
        OREF_COUNT(3);
        return b ? t : f;
    }
};

static inline float floorfrac(float x, int* i)
{
    *i = ifloor(x);
    return x - (float)*i;
}
static inline D2<float> floorfrac(const D2<float>& x, int* i)
{
    float f = floorfrac(x.v, i);
    return D2<float>(f, x.dx, x.dy);
}

template<class T> static inline T fade(const T& t)
{
    // t*t*t*(t*(t*6-15)+10), constants promoted to T as in the template
    return t * t * t * (t * (t * Ops<T>::c(6.0f) - Ops<T>::c(15.0f))
                        + Ops<T>::c(10.0f));
}

template<class T, class Q> static inline T lerp(const T& a, const T& b, const Q& x)
{
    return a * (Ops<Q>::one() - x) + b * x;
}
template<class T, class Q>
static inline T bilerp(const T& v0, const T& v1, const T& v2, const T& v3,
                       const Q& s, const Q& t)
{
    Q s1 = Ops<Q>::one() - s;
    return (Ops<Q>::one() - t) * (v0 * s1 + v1 * s) + t * (v2 * s1 + v3 * s);
}
template<class T, class Q>
static inline T trilerp(const T& v0, const T& v1, const T& v2, const T& v3,
                        const T& v4, const T& v5, const T& v6, const T& v7,
                        const Q& s, const Q& t, const Q& r)
{
    Q s1 = Ops<Q>::one() - s;
    Q t1 = Ops<Q>::one() - t;
    Q r1 = Ops<Q>::one() - r;
    return r1 * (t1 * (v0 * s1 + v1 * s) + t * (v2 * s1 + v3 * s))
           + r * (t1 * (v4 * s1 + v5 * s) + t * (v6 * s1 + v7 * s));
}

template<class T> static inline T grad(int hash, const T& x)
{
    OREF_COUNT(5);
    int h   = hash & 15;
    float g = (float)(1 + (h & 7));
    if (h & 8)
        g = -g;
    return g * x;
}
template<class T> static inline T grad(int hash, const T& x, const T& y)
{
    OREF_COUNT(4);
    int h = hash & 7;
    T u   = Ops<T>::sel(h < 4, x, y);
    T v   = 2.0f * Ops<T>::sel(h < 4, y, x);
    return Ops<T>::neg_if(u, h & 1) + Ops<T>::neg_if(v, h & 2);
}
template<class T> static inline T grad(int hash, const T& x, const T& y, const T& z)
{
    OREF_COUNT(8);
    int h = hash & 15;
    T u   = Ops<T>::sel(h < 8, x, y);
    T v   = Ops<T>::sel(h < 4, y, Ops<T>::sel((h == 12) || (h == 14), x, z));
    return Ops<T>::neg_if(u, h & 1) + Ops<T>::neg_if(v, h & 2);
}
template<class T>
static inline T grad(int hash, const T& x, const T& y, const T& z, const T& w)
{
    OREF_COUNT(7);
    int h = hash & 31;
    T u   = Ops<T>::sel(h < 24, x, y);
    T v   = Ops<T>::sel(h < 16, y, z);
    T s   = Ops<T>::sel(h < 8, z, w);
    return Ops<T>::neg_if(u, h & 1) + Ops<T>::neg_if(v, h & 2)
           + Ops<T>::neg_if(s, h & 4);
}

// Lattice hash for component 'comp' of the result: comp 0 is the scalar
// HashScalar value, comps 1/2 take (h>>8)&0xFF and (h>>16)&0xFF as
// HashVector/sliceup does. grad() only looks at the low 5 bits, so the
// component-0 mask (h & 0xFF) does not change anything.
//
// With per-axis periods (HashScalarPeriodic / HashVectorPeriodic,
// oslnoise.h:1157-1290) each lattice coordinate is first reduced with imod
// (oslnoise.h:847-859) to [0, period).
struct LatHash {
    int shift;
    const int* per = nullptr;  // periods (>= 1) per axis, or null
    uint32_t w(int x, int axis) const
    {
        return per ? (uint32_t)imod(x, per[axis]) : (uint32_t)x;
    }
    int operator()(int x) const { return sl(inthash(w(x, 0))); }
    int operator()(int x, int y) const
    {
        return sl(inthash(w(x, 0), w(y, 1)));
    }
    int operator()(int x, int y, int z) const
    {
        return sl(inthash(w(x, 0), w(y, 1), w(z, 2)));
    }
    int operator()(int x, int y, int z, int ww) const
    {
        return sl(inthash(w(x, 0), w(y, 1), w(z, 2), w(ww, 3)));
    }
    int sl(uint32_t h) const
    {
        return shift == 0 ? (int)h : (int)((h >> shift) & 0xFFu);
    }
};

template<class T> static T perlin1(const LatHash& hash, const T& x)
{
    int X;
    T fx = floorfrac(x, &X);
    T u  = fade(fx);
    T r  = lerp(grad(hash(X), fx), grad(hash(X + 1), fx - 1.0f), u);
    return 0.25f * r;
}

template<class T> static T perlin2(const LatHash& hash, const T& x, const T& y)
{
    int X, Y;
    T fx = floorfrac(x, &X);
    T fy = floorfrac(y, &Y);
    T u  = fade(fx);
    T v  = fade(fy);
    T r  = bilerp(grad(hash(X, Y), fx, fy), grad(hash(X + 1, Y), fx - 1.0f, fy),
                  grad(hash(X, Y + 1), fx, fy - 1.0f),
                  grad(hash(X + 1, Y + 1), fx - 1.0f, fy - 1.0f), u, v);
    return 0.6616f * r;
}

// CGDefault (OIIO SIMD) interpolation order of the 3D/4D perlin, the
// implementation the batched path switches to for lanes of a batch with
// fewer active lanes than its SIMD threshold (wide_opnoise_impl_xmacro.h:
// 57-127, thresholds wide_opnoise_perlin_float.cpp:17-26): corners packed
// as (X,Y) (X+1,Y) (X,Y+1) (X+1,Y+1) per z plane, interpolated along z
// first, then x, then y (oslnoise.h:834-841, :1384-1430, :1461-1500). 1D
// and 2D interpolate in the scalar order either way.
template<class T, class Q>
static inline T trilerp_zxy(const T& v0, const T& v1, const T& v2, const T& v3, const T& v4,
                            const T& v5, const T& v6, const T& v7, const Q& s, const Q& t,
                            const Q& r)
{
    T xy0 = lerp(v0, v4, r), xy1 = lerp(v1, v5, r), xy2 = lerp(v2, v6, r), xy3 = lerp(v3, v7, r);
    T xx0 = lerp(xy0, xy1, s), xx2 = lerp(xy2, xy3, s);
    return lerp(xx0, xx2, t);
}
template<bool SIMD, class T, class Q>
static inline T trilerp_p(const T& v0, const T& v1, const T& v2, const T& v3, const T& v4,
                          const T& v5, const T& v6, const T& v7, const Q& s, const Q& t,
                          const Q& r)
{
    return SIMD ? trilerp_zxy(v0, v1, v2, v3, v4, v5, v6, v7, s, t, r)
                : trilerp(v0, v1, v2, v3, v4, v5, v6, v7, s, t, r);
}

template<class T, bool SIMD = false>
static T perlin3(const LatHash& hash, const T& x, const T& y, const T& z)
{
    int X, Y, Z;
    T fx = floorfrac(x, &X);
    T fy = floorfrac(y, &Y);
    T fz = floorfrac(z, &Z);
    T u  = fade(fx);
    T v  = fade(fy);
    T w  = fade(fz);
    T r  = trilerp_p<SIMD>(grad(hash(X, Y, Z), fx, fy, fz),
                   grad(hash(X + 1, Y, Z), fx - 1.0f, fy, fz),
                   grad(hash(X, Y + 1, Z), fx, fy - 1.0f, fz),
                   grad(hash(X + 1, Y + 1, Z), fx - 1.0f, fy - 1.0f, fz),
                   grad(hash(X, Y, Z + 1), fx, fy, fz - 1.0f),
                   grad(hash(X + 1, Y, Z + 1), fx - 1.0f, fy, fz - 1.0f),
                   grad(hash(X, Y + 1, Z + 1), fx, fy - 1.0f, fz - 1.0f),
                   grad(hash(X + 1, Y + 1, Z + 1), fx - 1.0f, fy - 1.0f,
                        fz - 1.0f),
                   u, v, w);
    return 0.9820f * r;
}

template<class T, bool SIMD = false>
static T perlin4(const LatHash& hash, const T& x, const T& y, const T& z,
                 const T& w)
{
    int X, Y, Z, W;
    T fx = floorfrac(x, &X);
    T fy = floorfrac(y, &Y);
    T fz = floorfrac(z, &Z);
    T fw = floorfrac(w, &W);
    T u  = fade(fx);
    T v  = fade(fy);
    T t  = fade(fz);
    T s  = fade(fw);
    T a  = trilerp_p<SIMD>(grad(hash(X, Y, Z, W), fx, fy, fz, fw),
                   grad(hash(X + 1, Y, Z, W), fx - 1.0f, fy, fz, fw),
                   grad(hash(X, Y + 1, Z, W), fx, fy - 1.0f, fz, fw),
                   grad(hash(X + 1, Y + 1, Z, W), fx - 1.0f, fy - 1.0f, fz, fw),
                   grad(hash(X, Y, Z + 1, W), fx, fy, fz - 1.0f, fw),
                   grad(hash(X + 1, Y, Z + 1, W), fx - 1.0f, fy, fz - 1.0f, fw),
                   grad(hash(X, Y + 1, Z + 1, W), fx, fy - 1.0f, fz - 1.0f, fw),
                   grad(hash(X + 1, Y + 1, Z + 1, W), fx - 1.0f, fy - 1.0f,
                        fz - 1.0f, fw),
                   u, v, t);
    T b = trilerp_p<SIMD>(grad(hash(X, Y, Z, W + 1), fx, fy, fz, fw - 1.0f),
                  grad(hash(X + 1, Y, Z, W + 1), fx - 1.0f, fy, fz, fw - 1.0f),
                  grad(hash(X, Y + 1, Z, W + 1), fx, fy - 1.0f, fz, fw - 1.0f),
                  grad(hash(X + 1, Y + 1, Z, W + 1), fx - 1.0f, fy - 1.0f, fz,
                       fw - 1.0f),
                  grad(hash(X, Y, Z + 1, W + 1), fx, fy, fz - 1.0f, fw - 1.0f),
                  grad(hash(X + 1, Y, Z + 1, W + 1), fx - 1.0f, fy, fz - 1.0f,
                       fw - 1.0f),
                  grad(hash(X, Y + 1, Z + 1, W + 1), fx, fy - 1.0f, fz - 1.0f,
                       fw - 1.0f),
                  grad(hash(X + 1, Y + 1, Z + 1, W + 1), fx - 1.0f, fy - 1.0f,
                       fz - 1.0f, fw - 1.0f),
                  u, v, t);
    return 0.8344f * lerp(a, b, s);
}

template<class T, bool SIMD = false>
static T perlin_n(int dim, const T* p, int shift, const int* per = nullptr)
{
    LatHash h { shift, per };
    switch (dim) {
    case 1: return perlin1(h, p[0]);
    case 2: return perlin2(h, p[0], p[1]);
    case 3: return perlin3<T, SIMD>(h, p[0], p[1], p[2]);
    default: return perlin4<T, SIMD>(h, p[0], p[1], p[2], p[3]);
    }
}

// the CGDefault (SIMD-order) signed perlin, value and Dual2, per component
float snoise_f_cgdefault(int dim, const float* p, int comp)
{
    return perlin_n<float, true>(dim, p, 8 * comp);
}
D2<float> snoise_df_cgdefault(int dim, const D2<float>* p, int comp)
{
    return perlin_n<D2<float>, true>(dim, p, 8 * comp);
}

float snoise_f(int dim, const float* p) { return perlin_n<float>(dim, p, 0); }
V3 snoise_v(int dim, const float* p)
{
    return V3(perlin_n<float>(dim, p, 0), perlin_n<float>(dim, p, 8),
              perlin_n<float>(dim, p, 16));
}
D2<float> snoise_df(int dim, const D2<float>* p)
{
    return perlin_n<D2<float>>(dim, p, 0);
}
D2<V3> snoise_dv(int dim, const D2<float>* p)
{
    D2<float> a = perlin_n<D2<float>>(dim, p, 0);
    D2<float> b = perlin_n<D2<float>>(dim, p, 8);
    D2<float> c = perlin_n<D2<float>>(dim, p, 16);
    return D2<V3>(V3(a.v, b.v, c.v), V3(a.dx, b.dx, c.dx), V3(a.dy, b.dy, c.dy));
}

// NoiseImpl (oslnoise.h:2284-2400): 0.5f * (perlin + 1.0f)
float unoise_f(int dim, const float* p) { return 0.5f * (snoise_f(dim, p) + 1.0f); }
V3 unoise_v(int dim, const float* p)
{
    V3 r = snoise_v(dim, p);
    return V3(0.5f * (r.x + 1.0f), 0.5f * (r.y + 1.0f), 0.5f * (r.z + 1.0f));
}
D2<float> unoise_df(int dim, const D2<float>* p)
{
    return 0.5f * (snoise_df(dim, p) + 1.0f);
}
D2<V3> unoise_dv(int dim, const D2<float>* p)
{
    D2<V3> r = snoise_dv(dim, p);
    return D2<V3>(V3(0.5f * (r.v.x + 1.0f), 0.5f * (r.v.y + 1.0f),
                     0.5f * (r.v.z + 1.0f)),
                  V3(0.5f * r.dx.x, 0.5f * r.dx.y, 0.5f * r.dx.z),
                  V3(0.5f * r.dy.x, 0.5f * r.dy.y, 0.5f * r.dy.z));
}

// ---------------------------------------------------------------------------
// Periodic variants. PeriodicSNoise / PeriodicNoise (oslnoise.h:2531-2760)
// are the perlin above over the periodic lattice hash; the periods are
// plain floats (no derivatives), turned into lattice periods by
// period_of. PeriodicCellNoise / PeriodicHashNoise (PeriodicAdaptionOf,
// oslnoise.h:486-548) wrap the float coordinates with wrap_period and then
// hash as the plain versions do.
static void lattice_periods(int dim, const float* period, int per[4])
{
    for (int i = 0; i < dim; ++i)
        per[i] = period_of(period[i]);
}

float psnoise_f(int dim, const float* p, const float* period)
{
    int per[4];
    lattice_periods(dim, period, per);
    return perlin_n<float>(dim, p, 0, per);
}
V3 psnoise_v(int dim, const float* p, const float* period)
{
    int per[4];
    lattice_periods(dim, period, per);
    return V3(perlin_n<float>(dim, p, 0, per), perlin_n<float>(dim, p, 8, per),
              perlin_n<float>(dim, p, 16, per));
}
D2<float> psnoise_df(int dim, const D2<float>* p, const float* period)
{
    int per[4];
    lattice_periods(dim, period, per);
    return perlin_n<D2<float>>(dim, p, 0, per);
}
D2<V3> psnoise_dv(int dim, const D2<float>* p, const float* period)
{
    int per[4];
    lattice_periods(dim, period, per);
    D2<float> a = perlin_n<D2<float>>(dim, p, 0, per);
    D2<float> b = perlin_n<D2<float>>(dim, p, 8, per);
    D2<float> c = perlin_n<D2<float>>(dim, p, 16, per);
    return D2<V3>(V3(a.v, b.v, c.v), V3(a.dx, b.dx, c.dx), V3(a.dy, b.dy, c.dy));
}
float pnoise_f(int dim, const float* p, const float* period)
{
    return 0.5f * (psnoise_f(dim, p, period) + 1.0f);
}
V3 pnoise_v(int dim, const float* p, const float* period)
{
    V3 r = psnoise_v(dim, p, period);
    return V3(0.5f * (r.x + 1.0f), 0.5f * (r.y + 1.0f), 0.5f * (r.z + 1.0f));
}
D2<float> pnoise_df(int dim, const D2<float>* p, const float* period)
{
    return 0.5f * (psnoise_df(dim, p, period) + 1.0f);
}
D2<V3> pnoise_dv(int dim, const D2<float>* p, const float* period)
{
    D2<V3> r = psnoise_dv(dim, p, period);
    return D2<V3>(V3(0.5f * (r.v.x + 1.0f), 0.5f * (r.v.y + 1.0f),
                     0.5f * (r.v.z + 1.0f)),
                  V3(0.5f * r.dx.x, 0.5f * r.dx.y, 0.5f * r.dx.z),
                  V3(0.5f * r.dy.x, 0.5f * r.dy.y, 0.5f * r.dy.z));
}

static void wrap_coords(int dim, const float* p, const float* period, float* q)
{
    for (int i = 0; i < dim; ++i)
        q[i] = wrap_period(p[i], period[i]);
}
float pcellnoise_f(int dim, const float* p, const float* period)
{
    float q[4];
    wrap_coords(dim, p, period, q);
    return cellnoise_f(dim, q);
}
V3 pcellnoise_v(int dim, const float* p, const float* period)
{
    float q[4];
    wrap_coords(dim, p, period, q);
    return cellnoise_v(dim, q);
}
float phashnoise_f(int dim, const float* p, const float* period)
{
    float q[4];
    wrap_coords(dim, p, period, q);
    return hashnoise_f(dim, q);
}
V3 phashnoise_v(int dim, const float* p, const float* period)
{
    float q[4];
    wrap_coords(dim, p, period, q);
    return hashnoise_v(dim, q);
}

// ---------------------------------------------------------------------------
// simplex: src/include/OSL/sfm_simplex.h (scramble :356, grad1-3 :363-440,
// simplexnoise1 :559, simplexnoise2 :600, simplexnoise3 :743). Gradient
// tables: 2D :84-96 (8 entries), 3D :129-183 (12 cube edges + 4 repeats).
static inline uint32_t scramble(uint32_t v0, uint32_t v1 = 0, uint32_t v2 = 0)
{
    return bjfinal(v0, v1, v2 ^ 0xdeadbeefu);
}

static const float kGrad2[8][2] = { { -1, -1 }, { 1, 0 },  { -1, 0 }, { 1, 1 },
                                    { -1, 1 },  { 0, -1 }, { 0, 1 },  { 1, -1 } };
static const float kGrad3[16][3] = {
    { 1, 0, 1 },  { 0, 1, 1 },  { -1, 0, 1 },  { 0, -1, 1 },
    { 1, 0, -1 }, { 0, 1, -1 }, { -1, 0, -1 }, { 0, -1, -1 },
    { 1, -1, 0 }, { 1, 1, 0 },  { -1, 1, 0 },  { -1, -1, 0 },
    { 1, 0, 1 },  { -1, 0, 1 }, { 0, 1, -1 },  { 0, -1, -1 }
};

static float simplex1(float x, int seed, float* dnoise)
{
    int i0    = ifloor(x);
    int i1    = i0 + 1;
    float x0  = x - (float)i0;
    float x1  = x0 - 1.0f;
    float x20 = x0 * x0;
    float t0  = 1.0f - x20;
    float t20 = t0 * t0;
    float t40 = t20 * t20;
    int h0    = (int)scramble((uint32_t)i0, (uint32_t)seed);
    float gx0 = 1.0f + (float)(h0 & 7);
    if (h0 & 8)
        gx0 = -gx0;
    float n0  = t40 * gx0 * x0;
    float x21 = x1 * x1;
    float t1  = 1.0f - x21;
    float t21 = t1 * t1;
    float t41 = t21 * t21;
    int h1    = (int)scramble((uint32_t)i1, (uint32_t)seed);
    float gx1 = 1.0f + (float)(h1 & 7);
    if (h1 & 8)
        gx1 = -gx1;
    float n1          = t41 * gx1 * x1;
    const float scale = 0.36f;
    if (dnoise) {
        float d = t20 * t0 * gx0 * x20;
        d += t21 * t1 * gx1 * x21;
        d *= -8.0f;
        d += t40 * gx0 + t41 * gx1;
        d *= scale;
        dnoise[0] = d;
    }
    return scale * (n0 + n1);
}

static float simplex2(float x, float y, int seed, float* dnoise)
{
    const float F2 = 0.366025403f;
    const float G2 = 0.211324865f;
    float s        = (x + y) * F2;
    float xs       = x + s;
    float ys       = y + s;
    int i          = ifloor(xs);
    int j          = ifloor(ys);
    float t        = (float)(i + j) * G2;
    float X0       = (float)i - t;
    float Y0       = (float)j - t;
    float x0       = x - X0;
    float y0       = y - Y0;
    int i1         = (x0 > y0);
    int j1         = !i1;
    float x1       = x0 - (float)i1 + G2;
    float y1       = y0 - (float)j1 + G2;
    float x2       = x0 - 1.0f + 2.0f * G2;
    float y2       = y0 - 1.0f + 2.0f * G2;
    auto g2        = [&](int a, int b, float* g) {
        int h = (int)scramble((uint32_t)a, (uint32_t)b, (uint32_t)seed);
        g[0]  = kGrad2[h & 7][0];
        g[1]  = kGrad2[h & 7][1];
    };
    float g0[2], g1[2], gg2[2];
    float t0  = 0.5f - x0 * x0 - y0 * y0;
    float t20 = t0 * t0;
    float t40 = t20 * t20;
    g2(i, j, g0);
    float n0 = t40 * (g0[0] * x0 + g0[1] * y0);
    if (t0 < 0.0f)
        n0 = 0.0f;
    float t1  = 0.5f - x1 * x1 - y1 * y1;
    float t21 = t1 * t1;
    float t41 = t21 * t21;
    g2(i + i1, j + j1, g1);
    float n1 = t41 * (g1[0] * x1 + g1[1] * y1);
    if (t1 < 0.0f)
        n1 = 0.0f;
    float t2  = 0.5f - x2 * x2 - y2 * y2;
    float t22 = t2 * t2;
    float t42 = t22 * t22;
    g2(i + 1, j + 1, gg2);
    float n2 = t42 * (gg2[0] * x2 + gg2[1] * y2);
    if (t2 < 0.0f)
        n2 = 0.0f;
    const float scale = 64.0f;
    float noise       = scale * (n0 + n1 + n2);
    if (dnoise) {
        if (t0 < 0.0f)
            g0[0] = g0[1] = 0.0f;
        float temp0 = t20 * t0 * (g0[0] * x0 + g0[1] * y0);
        float dx    = temp0 * x0;
        float dy    = temp0 * y0;
        if (t1 < 0.0f)
            g1[0] = g1[1] = 0.0f;
        float temp1 = t21 * t1 * (g1[0] * x1 + g1[1] * y1);
        dx += temp1 * x1;
        dy += temp1 * y1;
        if (t2 < 0.0f)
            gg2[0] = gg2[1] = 0.0f;
        float temp2 = t22 * t2 * (gg2[0] * x2 + gg2[1] * y2);
        dx += temp2 * x2;
        dy += temp2 * y2;
        dx *= -8.0f;
        dy *= -8.0f;
        dx += t40 * g0[0] + t41 * g1[0] + t42 * gg2[0];
        dy += t40 * g0[1] + t41 * g1[1] + t42 * gg2[1];
        dx *= scale;
        dy *= scale;
        dnoise[0] = dx;
        dnoise[1] = dy;
    }
    return noise;
}

static float simplex3(float x, float y, float z, int seed, float* dnoise)
{
    const float F3 = 0.333333333f;
    const float G3 = 0.166666667f;
    float s        = (x + y + z) * F3;
    float xs       = x + s;
    float ys       = y + s;
    float zs       = z + s;
    int i          = ifloor(xs);
    int j          = ifloor(ys);
    int k          = ifloor(zs);
    float t        = (float)(i + j + k) * G3;
    float X0       = (float)i - t;
    float Y0       = (float)j - t;
    float Z0       = (float)k - t;
    float x0       = x - X0;
    float y0       = y - Y0;
    float z0       = z - Z0;
    int bg0 = (x0 >= y0), bg1 = (y0 >= z0), bg2 = (x0 >= z0);
    int nbg0 = !bg0, nbg1 = !bg1, nbg2 = !bg2;
    int i1 = bg0 & (bg1 | bg2);
    int j1 = nbg0 & bg1;
    int k1 = nbg1 & ((bg0 & nbg2) | nbg0);
    int i2 = bg0 | (bg1 & bg2);
    int j2 = bg1 | nbg0;
    int k2 = (bg0 & nbg1) | (nbg0 & (nbg1 | nbg2));
    float x1 = x0 - (float)i1 + G3;
    float y1 = y0 - (float)j1 + G3;
    float z1 = z0 - (float)k1 + G3;
    float x2 = x0 - (float)i2 + 2.0f * G3;
    float y2 = y0 - (float)j2 + 2.0f * G3;
    float z2 = z0 - (float)k2 + 2.0f * G3;
    float x3 = x0 - 1.0f + 3.0f * G3;
    float y3 = y0 - 1.0f + 3.0f * G3;
    float z3 = z0 - 1.0f + 3.0f * G3;
    OREF_COUNT(15);  // simplex corner ordering logic (bg*/i1..k2)
    auto g3  = [&](int a, int b, int c, float* g) {
        OREF_COUNT(9);  // corner index adds, table decode (3 components)
        uint32_t h = scramble((uint32_t)a, (uint32_t)b,
                              scramble((uint32_t)c, (uint32_t)seed));
        g[0]       = kGrad3[h & 15][0];
        g[1]       = kGrad3[h & 15][1];
        g[2]       = kGrad3[h & 15][2];
    };
    float g0[3], g1[3], gg2[3], gg3[3];
    float t0  = 0.5f - x0 * x0 - y0 * y0 - z0 * z0;
    float t20 = t0 * t0;
    float t40 = t20 * t20;
    g3(i, j, k, g0);
    float n0 = t40 * (g0[0] * x0 + g0[1] * y0 + g0[2] * z0);
    if (t0 < 0.0f)
        n0 = 0.0f;
    float t1  = 0.5f - x1 * x1 - y1 * y1 - z1 * z1;
    float t21 = t1 * t1;
    float t41 = t21 * t21;
    g3(i + i1, j + j1, k + k1, g1);
    float n1 = t41 * (g1[0] * x1 + g1[1] * y1 + g1[2] * z1);
    if (t1 < 0.0f)
        n1 = 0.0f;
    float t2  = 0.5f - x2 * x2 - y2 * y2 - z2 * z2;
    float t22 = t2 * t2;
    float t42 = t22 * t22;
    g3(i + i2, j + j2, k + k2, gg2);
    float n2 = t42 * (gg2[0] * x2 + gg2[1] * y2 + gg2[2] * z2);
    if (t2 < 0.0f)
        n2 = 0.0f;
    float t3  = 0.5f - x3 * x3 - y3 * y3 - z3 * z3;
    float t23 = t3 * t3;
    float t43 = t23 * t23;
    g3(i + 1, j + 1, k + 1, gg3);
    float n3 = t43 * (gg3[0] * x3 + gg3[1] * y3 + gg3[2] * z3);
    if (t3 < 0.0f)
        n3 = 0.0f;
    const float scale = 68.0f;
    float noise       = scale * (n0 + n1 + n2 + n3);
    if (dnoise) {
        if (t0 < 0.0f)
            g0[0] = g0[1] = g0[2] = 0.0f;
        float temp0 = t20 * t0 * (g0[0] * x0 + g0[1] * y0 + g0[2] * z0);
        float dx    = temp0 * x0;
        float dy    = temp0 * y0;
        float dz    = temp0 * z0;
        if (t1 < 0.0f)
            g1[0] = g1[1] = g1[2] = 0.0f;
        float temp1 = t21 * t1 * (g1[0] * x1 + g1[1] * y1 + g1[2] * z1);
        dx += temp1 * x1;
        dy += temp1 * y1;
        dz += temp1 * z1;
        if (t2 < 0.0f)
            gg2[0] = gg2[1] = gg2[2] = 0.0f;
        float temp2 = t22 * t2 * (gg2[0] * x2 + gg2[1] * y2 + gg2[2] * z2);
        dx += temp2 * x2;
        dy += temp2 * y2;
        dz += temp2 * z2;
        if (t3 < 0.0f)
            gg3[0] = gg3[1] = gg3[2] = 0.0f;
        float temp3 = t23 * t3 * (gg3[0] * x3 + gg3[1] * y3 + gg3[2] * z3);
        dx += temp3 * x3;
        dy += temp3 * y3;
        dz += temp3 * z3;
        dx *= -8.0f;
        dy *= -8.0f;
        dz *= -8.0f;
        dx += t40 * g0[0] + t41 * g1[0] + t42 * gg2[0] + t43 * gg3[0];
        dy += t40 * g0[1] + t41 * g1[1] + t42 * gg2[1] + t43 * gg3[1];
        dz += t40 * g0[2] + t41 * g1[2] + t42 * gg2[2] + t43 * gg3[2];
        dx *= scale;
        dy *= scale;
        dz *= scale;
        dnoise[0] = dx;
        dnoise[1] = dy;
        dnoise[2] = dz;
    }
    return noise;
}

// 4D: simplexnoise4 (sfm_simplex.h:970-1230) with grad4 (:440-470) over the
// 32 tesseract edges (:184-193) and the simplex corner order of the
// 64-entry rank table (:255-263; entry c holds each coordinate's rank,
// 3 = largest).
static const float kGrad4[32][4] = {
    { 0, 1, 1, 1 },  { 0, 1, 1, -1 },  { 0, 1, -1, 1 },  { 0, 1, -1, -1 },
    { 0, -1, 1, 1 }, { 0, -1, 1, -1 }, { 0, -1, -1, 1 }, { 0, -1, -1, -1 },
    { 1, 0, 1, 1 },  { 1, 0, 1, -1 },  { 1, 0, -1, 1 },  { 1, 0, -1, -1 },
    { -1, 0, 1, 1 }, { -1, 0, 1, -1 }, { -1, 0, -1, 1 }, { -1, 0, -1, -1 },
    { 1, 1, 0, 1 },  { 1, 1, 0, -1 },  { 1, -1, 0, 1 },  { 1, -1, 0, -1 },
    { -1, 1, 0, 1 }, { -1, 1, 0, -1 }, { -1, -1, 0, 1 }, { -1, -1, 0, -1 },
    { 1, 1, 1, 0 },  { 1, 1, -1, 0 },  { 1, -1, 1, 0 },  { 1, -1, -1, 0 },
    { -1, 1, 1, 0 }, { -1, 1, -1, 0 }, { -1, -1, 1, 0 }, { -1, -1, -1, 0 }
};
static const unsigned char kSimplex4[64][4] = {
    { 0, 1, 2, 3 }, { 0, 1, 3, 2 }, { 0, 0, 0, 0 }, { 0, 2, 3, 1 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 0, 0, 0, 0 }, { 1, 2, 3, 0 }, { 0, 2, 1, 3 }, { 0, 0, 0, 0 }, { 0, 3, 1, 2 }, { 0, 3, 2, 1 },
    { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 1, 3, 2, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 1, 2, 0, 3 }, { 0, 0, 0, 0 }, { 1, 3, 0, 2 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 2, 3, 0, 1 }, { 2, 3, 1, 0 }, { 1, 0, 2, 3 }, { 1, 0, 3, 2 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 0, 0, 0, 0 }, { 2, 0, 3, 1 }, { 0, 0, 0, 0 }, { 2, 1, 3, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 2, 0, 1, 3 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 3, 0, 1, 2 }, { 3, 0, 2, 1 },
    { 0, 0, 0, 0 }, { 3, 1, 2, 0 }, { 2, 1, 0, 3 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 }, { 0, 0, 0, 0 },
    { 3, 1, 0, 2 }, { 0, 0, 0, 0 }, { 3, 2, 0, 1 }, { 3, 2, 1, 0 }
};

static float simplex4(float x, float y, float z, float w, int seed, float* dnoise)
{
    const float F4 = 0.309016994f;
    const float G4 = 0.138196601f;
    float s        = (x + y + z + w) * F4;
    float xs = x + s, ys = y + s, zs = z + s, ws = w + s;
    int i = ifloor(xs), j = ifloor(ys), k = ifloor(zs), l = ifloor(ws);
    float t  = (float)(i + j + k + l) * G4;
    float x0 = x - ((float)i - t);
    float y0 = y - ((float)j - t);
    float z0 = z - ((float)k - t);
    float w0 = w - ((float)l - t);
    int c = ((x0 > y0) ? 32 : 0) | ((x0 > z0) ? 16 : 0) | ((y0 > z0) ? 8 : 0)
            | ((x0 > w0) ? 4 : 0) | ((y0 > w0) ? 2 : 0) | ((z0 > w0) ? 1 : 0);
    const unsigned char* r = kSimplex4[c];
    OREF_COUNT(30);  // corner ordering: six compares, rank table, offsets
    int off[3][4];
    for (int q = 0; q < 3; ++q)
        for (int a = 0; a < 4; ++a)
            off[q][a] = r[a] >= 3 - q ? 1 : 0;
    // corner offsets in (x, y, z, w): corner q is the origin minus its
    // integer offset plus q * G4
    float cx[5], cy[5], cz[5], cw[5];
    cx[0] = x0;
    cy[0] = y0;
    cz[0] = z0;
    cw[0] = w0;
    const float gq[3] = { G4, 2.0f * G4, 3.0f * G4 };
    for (int q = 0; q < 3; ++q) {
        cx[q + 1] = x0 - (float)off[q][0] + gq[q];
        cy[q + 1] = y0 - (float)off[q][1] + gq[q];
        cz[q + 1] = z0 - (float)off[q][2] + gq[q];
        cw[q + 1] = w0 - (float)off[q][3] + gq[q];
    }
    cx[4] = x0 - 1.0f + 4.0f * G4;
    cy[4] = y0 - 1.0f + 4.0f * G4;
    cz[4] = z0 - 1.0f + 4.0f * G4;
    cw[4] = w0 - 1.0f + 4.0f * G4;
    const int ci[5][4] = { { i, j, k, l },
                           { i + off[0][0], j + off[0][1], k + off[0][2], l + off[0][3] },
                           { i + off[1][0], j + off[1][1], k + off[1][2], l + off[1][3] },
                           { i + off[2][0], j + off[2][1], k + off[2][2], l + off[2][3] },
                           { i + 1, j + 1, k + 1, l + 1 } };
    float g[5][4], tt[5], t2[5], t4[5], nn[5];
    for (int q = 0; q < 5; ++q) {
        OREF_COUNT(12);  // corner index adds, table decode (4 components)
        uint32_t h = scramble((uint32_t)ci[q][0], (uint32_t)ci[q][1],
                              scramble((uint32_t)ci[q][2], (uint32_t)ci[q][3], (uint32_t)seed));
        for (int a = 0; a < 4; ++a)
            g[q][a] = kGrad4[h & 31][a];
        tt[q] = 0.5f - cx[q] * cx[q] - cy[q] * cy[q] - cz[q] * cz[q] - cw[q] * cw[q];
        t2[q] = tt[q] * tt[q];
        t4[q] = t2[q] * t2[q];
        nn[q] = t4[q] * (g[q][0] * cx[q] + g[q][1] * cy[q] + g[q][2] * cz[q] + g[q][3] * cw[q]);
        if (tt[q] < 0.0f)
            nn[q] = 0.0f;
    }
    const float scale = 54.0f;
    float noise       = scale * (nn[0] + nn[1] + nn[2] + nn[3] + nn[4]);
    if (dnoise) {
        float d[4] = { 0, 0, 0, 0 };
        for (int q = 0; q < 5; ++q) {
            if (tt[q] < 0.0f)
                g[q][0] = g[q][1] = g[q][2] = g[q][3] = 0.0f;
            float temp = t2[q] * tt[q]
                         * (g[q][0] * cx[q] + g[q][1] * cy[q] + g[q][2] * cz[q] + g[q][3] * cw[q]);
            if (q == 0) {
                d[0] = temp * cx[q];
                d[1] = temp * cy[q];
                d[2] = temp * cz[q];
                d[3] = temp * cw[q];
            } else {
                d[0] += temp * cx[q];
                d[1] += temp * cy[q];
                d[2] += temp * cz[q];
                d[3] += temp * cw[q];
            }
        }
        for (int a = 0; a < 4; ++a) {
            d[a] *= -8.0f;
            d[a] += t4[0] * g[0][a] + t4[1] * g[1][a] + t4[2] * g[2][a] + t4[3] * g[3][a]
                    + t4[4] * g[4][a];
            dnoise[a] = d[a] * scale;
        }
    }
    return noise;
}

float simplex_f(int dim, const float* p, int seed, float* dnoise)
{
    switch (dim) {
    case 1: return simplex1(p[0], seed, dnoise);
    case 2: return simplex2(p[0], p[1], seed, dnoise);
    case 3: return simplex3(p[0], p[1], p[2], seed, dnoise);
    default: return simplex4(p[0], p[1], p[2], p[3], seed, dnoise);
    }
}

// Dual2 simplex: SimplexNoiseScalar dual overloads (oslnoise.h:2925-2957):
// derivatives by the chain rule  dn/dP . dP/dx  summed left to right.
D2<float> simplex_df(int dim, const D2<float>* p, int seed)
{
    float v[4] = { 0, 0, 0, 0 }, d[4] = { 0, 0, 0, 0 };
    for (int i = 0; i < dim; ++i)
        v[i] = p[i].v;
    float r = simplex_f(dim, v, seed, d);
    if (dim == 4)
        return D2<float>(r, d[0] * p[0].dx + d[1] * p[1].dx + d[2] * p[2].dx + d[3] * p[3].dx,
                         d[0] * p[0].dy + d[1] * p[1].dy + d[2] * p[2].dy + d[3] * p[3].dy);
    if (dim == 1)
        return D2<float>(r, d[0] * p[0].dx, d[0] * p[0].dy);
    if (dim == 2)
        return D2<float>(r, d[0] * p[0].dx + d[1] * p[1].dx,
                         d[0] * p[0].dy + d[1] * p[1].dy);
    return D2<float>(r, d[0] * p[0].dx + d[1] * p[1].dx + d[2] * p[2].dx,
                     d[0] * p[0].dy + d[1] * p[1].dy + d[2] * p[2].dy);
}

D2<V3> simplex_dv(int dim, const D2<float>* p)
{
    D2<float> a = simplex_df(dim, p, 0);
    D2<float> b = simplex_df(dim, p, 1);
    D2<float> c = simplex_df(dim, p, 2);
    return D2<V3>(V3(a.v, b.v, c.v), V3(a.dx, b.dx, c.dx), V3(a.dy, b.dy, c.dy));
}

// ---------------------------------------------------------------------------
float smoothstep_f(float e0, float e1, float x)
{
    if (x < e0)
        return 0.0f;
    else if (x >= e1)
        return 1.0f;
    float t = (x - e0) / (e1 - e0);
    return (3.0f - 2.0f * t) * (t * t);
}

D2<float> smoothstep_d(D2<float> e0, D2<float> e1, D2<float> x)
{
    if (x.v < e0.v)
        return D2<float>(0.0f);
    else if (x.v >= e1.v)
        return D2<float>(1.0f);
    D2<float> t = (x - e0) / (e1 - e0);
    // (T(3) - T(2)*t)*t*t with T(3), T(2) promoted to duals
    return (D2<float>(3.0f) - D2<float>(2.0f) * t) * t * t;
}

}  // namespace oref
