// CDNA4 device library: integer hashing and lattice noises.
//
// One work-item evaluates one shading point; the batched CPU path's W=16
// lane loop becomes the 64-wide wavefront. Perlin/cell/hash noise use no
// tables at all (pure integer VALU work: the Jenkins hash below); simplex
// uses a 16-entry gradient table held in registers as bit masks.
//
// Reference algorithms (file:line in the reference tree):
//   bjmix/bjfinal               src/include/OSL/oslnoise.h:154-180
//   inthash(k0..k4)             src/include/OSL/oslnoise.h:229-289
//   bits_to_01                  src/include/OSL/oslnoise.h:135-141
//   CellNoise/HashNoise         src/include/OSL/oslnoise.h:324-480
//   floorfrac/fade/grad/scale   src/include/OSL/oslnoise.h:872-1041
//   perlin<CGScalar> 1-4D       src/include/OSL/oslnoise.h:1327-2282
//   simplexnoise1/2/3/4         src/include/OSL/sfm_simplex.h:559-1230
#pragma once
#include "oslmi_types.h"

namespace oslmi {

// ---------------------------------------------------------------- hashing
OSLMI_DEV uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

OSLMI_DEV void bjmix(uint32_t& a, uint32_t& b, uint32_t& c)
{
    a -= c;  a ^= rotl32(c, 4);  c += b;
    b -= a;  b ^= rotl32(a, 6);  a += c;
    c -= b;  c ^= rotl32(b, 8);  b += a;
    a -= c;  a ^= rotl32(c, 16); c += b;
    b -= a;  b ^= rotl32(a, 19); a += c;
    c -= b;  c ^= rotl32(b, 4);  b += a;
}

OSLMI_DEV uint32_t bjfinal(uint32_t a, uint32_t b, uint32_t c)
{
    c ^= b; c -= rotl32(b, 14);
    a ^= c; a -= rotl32(c, 11);
    b ^= a; b -= rotl32(a, 25);
    c ^= b; c -= rotl32(b, 16);
    a ^= c; a -= rotl32(c, 4);
    b ^= a; b -= rotl32(a, 14);
    c ^= b; c -= rotl32(b, 24);
    return c;
}

#define OSLMI_SEED(n) (0xdeadbeefu + ((uint32_t)(n) << 2) + 13u)
OSLMI_DEV uint32_t inthash(uint32_t k0)
{
    return bjfinal(OSLMI_SEED(1) + k0, OSLMI_SEED(1), OSLMI_SEED(1));
}
OSLMI_DEV uint32_t inthash(uint32_t k0, uint32_t k1)
{
    return bjfinal(OSLMI_SEED(2) + k0, OSLMI_SEED(2) + k1, OSLMI_SEED(2));
}
OSLMI_DEV uint32_t inthash(uint32_t k0, uint32_t k1, uint32_t k2)
{
    return bjfinal(OSLMI_SEED(3) + k0, OSLMI_SEED(3) + k1, OSLMI_SEED(3) + k2);
}
OSLMI_DEV uint32_t inthash(uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3)
{
    uint32_t a = OSLMI_SEED(4) + k0, b = OSLMI_SEED(4) + k1, c = OSLMI_SEED(4) + k2;
    bjmix(a, b, c);
    a += k3;
    return bjfinal(a, b, c);
}
OSLMI_DEV uint32_t inthash(uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3, uint32_t k4)
{
    uint32_t a = OSLMI_SEED(5) + k0, b = OSLMI_SEED(5) + k1, c = OSLMI_SEED(5) + k2;
    bjmix(a, b, c);
    b += k4;
    a += k3;
    return bjfinal(a, b, c);
}

OSLMI_DEV float bits_to_01(uint32_t bits)
{
    const float k = (float)(1.0 / 4294967295.0);
    return (float)bits * k;
}

OSLMI_DEV int ifloor(float x) { return (int)floorf(x); }

// ----------------------------------------------------- cell / hash noise
// Keys are ifloor(x) (cell) or the float's bit pattern (hash).
template<bool CELL> OSLMI_DEV uint32_t nkey(float x)
{
    return CELL ? (uint32_t)ifloor(x) : __float_as_uint(x);
}

template<bool CELL> OSLMI_DEV float intnoise1(float x) { return bits_to_01(inthash(nkey<CELL>(x))); }
template<bool CELL> OSLMI_DEV float intnoise2(float x, float y)
{
    return bits_to_01(inthash(nkey<CELL>(x), nkey<CELL>(y)));
}
template<bool CELL> OSLMI_DEV float intnoise3(float x, float y, float z)
{
    return bits_to_01(inthash(nkey<CELL>(x), nkey<CELL>(y), nkey<CELL>(z)));
}
template<bool CELL> OSLMI_DEV float intnoise4(float x, float y, float z, float w)
{
    return bits_to_01(inthash(nkey<CELL>(x), nkey<CELL>(y), nkey<CELL>(z), nkey<CELL>(w)));
}
// Vec3 variants append the words 0,1,2; 3 and 4 keys share one bjmix
// (inthashVec, oslnoise.h:412-450).
template<bool CELL> OSLMI_DEV V3 vintnoise1(float x)
{
    uint32_t k = nkey<CELL>(x);
    return V3(bits_to_01(inthash(k, 0u)), bits_to_01(inthash(k, 1u)), bits_to_01(inthash(k, 2u)));
}
template<bool CELL> OSLMI_DEV V3 vintnoise2(float x, float y)
{
    uint32_t k0 = nkey<CELL>(x), k1 = nkey<CELL>(y);
    return V3(bits_to_01(inthash(k0, k1, 0u)), bits_to_01(inthash(k0, k1, 1u)),
              bits_to_01(inthash(k0, k1, 2u)));
}
template<bool CELL> OSLMI_DEV V3 vintnoise3(float x, float y, float z)
{
    uint32_t a = OSLMI_SEED(4) + nkey<CELL>(x), b = OSLMI_SEED(4) + nkey<CELL>(y),
             c = OSLMI_SEED(4) + nkey<CELL>(z);
    bjmix(a, b, c);
    return V3(bits_to_01(bjfinal(a + 0u, b, c)), bits_to_01(bjfinal(a + 1u, b, c)),
              bits_to_01(bjfinal(a + 2u, b, c)));
}
template<bool CELL> OSLMI_DEV V3 vintnoise4(float x, float y, float z, float w)
{
    uint32_t a = OSLMI_SEED(5) + nkey<CELL>(x), b = OSLMI_SEED(5) + nkey<CELL>(y),
             c = OSLMI_SEED(5) + nkey<CELL>(z);
    bjmix(a, b, c);
    uint32_t a2 = a + nkey<CELL>(w);
    return V3(bits_to_01(bjfinal(a2, b + 0u, c)), bits_to_01(bjfinal(a2, b + 1u, c)),
              bits_to_01(bjfinal(a2, b + 2u, c)));
}

// hash() shadeop (inthashf, oslnoise.h:565-610)
OSLMI_DEV int hash_op1(float x) { return (int)inthash(__float_as_uint(x)); }
OSLMI_DEV int hash_op2(float x, float y) { return (int)inthash(__float_as_uint(x), __float_as_uint(y)); }
OSLMI_DEV int hash_op3(float x, float y, float z)
{
    return (int)inthash(__float_as_uint(x), __float_as_uint(y), __float_as_uint(z));
}
OSLMI_DEV int hash_op4(float x, float y, float z, float w)
{
    return (int)inthash(__float_as_uint(x), __float_as_uint(y), __float_as_uint(z), __float_as_uint(w));
}
OSLMI_DEV int hash_opi(int x) { return (int)inthash((uint32_t)x); }

// ------------------------------------------------------- periodic lattice
// imod (oslnoise.h:847-859): the remainder of a / b in [0, b) for b > 0.
OSLMI_DEV int imod(int a, int b)
{
    int r = a % b;
    if (r < 0)
        r += b;
    return r;
}

// The lattice period of a float period argument (HashScalarPeriodic,
// oslnoise.h:1162-1180): ifloor(p), at least 1. A floor that no int holds
// (NaN, 2^31 and beyond) gives 1, as x86's truncating conversion (INT_MIN)
// followed by the clamp does on the CPU.
OSLMI_DEV int period_of(float p)
{
    const float f = floorf(p);
    return (f >= 1.0f && f < 2147483648.0f) ? (int)f : 1;
}

// PeriodicAdaptionOf::wrap (oslnoise.h:532-539) and gabornoise.h:185-192:
// s reduced into [0, floor(period)), the period at least 1. Periodic cell
// and hash noise hash the wrapped coordinates.
OSLMI_DEV float wrap_period(float s, float period)
{
    period = floorf(period);
    if (period < 1.0f)
        period = 1.0f;
    return s - period * floorf(s / period);
}

// Lattice-coordinate maps applied before hashing: NoWrap for the plain
// noises, Wrap (imod by the per-axis lattice period) for the periodic ones
// (HashScalarPeriodic / HashVectorPeriodic, oslnoise.h:1157-1290).
struct NoWrap {
    OSLMI_DEV uint32_t operator()(int x, int) const { return (uint32_t)x; }
};
struct Wrap {
    int p[4];
    OSLMI_DEV uint32_t operator()(int x, int axis) const { return (uint32_t)imod(x, p[axis]); }
};
OSLMI_DEV Wrap make_wrap(float a, float b = 1.0f, float c = 1.0f, float d = 1.0f)
{
    Wrap w;
    w.p[0] = period_of(a);
    w.p[1] = period_of(b);
    w.p[2] = period_of(c);
    w.p[3] = period_of(d);
    return w;
}

// ----------------------------------------------------------------- perlin
template<class T> struct NOps;
template<> struct NOps<float> {
    static OSLMI_DEV float one() { return 1.0f; }
    static OSLMI_DEV float c(float x) { return x; }
    static OSLMI_DEV float neg_if(float v, bool b) { return b ? -v : v; }
    static OSLMI_DEV float sel(bool b, float t, float f) { return b ? t : f; }
};
template<> struct NOps<Df> {
    static OSLMI_DEV Df one() { return Df(1.0f); }
    static OSLMI_DEV Df c(float x) { return Df(x); }
    static OSLMI_DEV Df neg_if(const Df& v, bool b) { return b ? -v : v; }
    static OSLMI_DEV Df sel(bool b, const Df& t, const Df& f) { return b ? t : f; }
};

OSLMI_DEV float floorfrac(float x, int* i)
{
    float f = floorf(x);
    *i      = (int)f;
    return x - (float)*i;
}
OSLMI_DEV Df floorfrac(const Df& x, int* i) { return Df(floorfrac(x.v, i), x.dx, x.dy); }

template<class T> OSLMI_DEV T fade(const T& t)
{
    return t * t * t * (t * (t * NOps<T>::c(6.0f) - NOps<T>::c(15.0f)) + NOps<T>::c(10.0f));
}
template<class T, class Q> OSLMI_DEV T lerp(const T& a, const T& b, const Q& x)
{
    return a * (NOps<Q>::one() - x) + b * x;
}
template<class T, class Q>
OSLMI_DEV T bilerp(const T& v0, const T& v1, const T& v2, const T& v3, const Q& s, const Q& t)
{
    Q s1 = NOps<Q>::one() - s;
    return (NOps<Q>::one() - t) * (v0 * s1 + v1 * s) + t * (v2 * s1 + v3 * s);
}
template<class T, class Q>
OSLMI_DEV T trilerp(const T& v0, const T& v1, const T& v2, const T& v3, const T& v4,
                    const T& v5, const T& v6, const T& v7, const Q& s, const Q& t, const Q& r)
{
    Q s1 = NOps<Q>::one() - s;
    Q t1 = NOps<Q>::one() - t;
    Q r1 = NOps<Q>::one() - r;
    return r1 * (t1 * (v0 * s1 + v1 * s) + t * (v2 * s1 + v3 * s))
           + r * (t1 * (v4 * s1 + v5 * s) + t * (v6 * s1 + v7 * s));
}

template<class T> OSLMI_DEV T grad1(int hash, const T& x)
{
    int h   = hash & 15;
    float g = (float)(1 + (h & 7));
    if (h & 8)
        g = -g;
    return g * x;
}
template<class T> OSLMI_DEV T grad2(int hash, const T& x, const T& y)
{
    int h = hash & 7;
    T u   = NOps<T>::sel(h < 4, x, y);
    T v   = 2.0f * NOps<T>::sel(h < 4, y, x);
    return NOps<T>::neg_if(u, h & 1) + NOps<T>::neg_if(v, h & 2);
}
template<class T> OSLMI_DEV T grad3(int hash, const T& x, const T& y, const T& z)
{
    int h = hash & 15;
    T u   = NOps<T>::sel(h < 8, x, y);
    T v   = NOps<T>::sel(h < 4, y, NOps<T>::sel((h == 12) || (h == 14), x, z));
    return NOps<T>::neg_if(u, h & 1) + NOps<T>::neg_if(v, h & 2);
}
template<class T> OSLMI_DEV T grad4(int hash, const T& x, const T& y, const T& z, const T& w)
{
    int h = hash & 31;
    T u   = NOps<T>::sel(h < 24, x, y);
    T v   = NOps<T>::sel(h < 16, y, z);
    T s   = NOps<T>::sel(h < 8, z, w);
    return NOps<T>::neg_if(u, h & 1) + NOps<T>::neg_if(v, h & 2) + NOps<T>::neg_if(s, h & 4);
}

// Component selector of the lattice hash: shift 0 is HashScalar, 8 and 16
// are the sliceup() fields of HashVector. grad() reads at most 5 low bits.
OSLMI_DEV int hslice(uint32_t h, int shift) { return (int)(h >> shift); }

template<class T, class W = NoWrap> OSLMI_DEV T perlin1(const T& x, int shift, const W& wr = W())
{
    int X;
    T fx = floorfrac(x, &X);
    T u  = fade(fx);
    T r  = lerp(grad1(hslice(inthash(wr(X, 0)), shift), fx),
                grad1(hslice(inthash(wr(X + 1, 0)), shift), fx - 1.0f), u);
    return 0.25f * r;
}

template<class T, class W = NoWrap>
OSLMI_DEV T perlin2(const T& x, const T& y, int shift, const W& wr = W())
{
    int X, Y;
    T fx = floorfrac(x, &X);
    T fy = floorfrac(y, &Y);
    T u  = fade(fx);
    T v  = fade(fy);
    uint32_t X0 = wr(X, 0), X1 = wr(X + 1, 0), Y0 = wr(Y, 1), Y1 = wr(Y + 1, 1);
    T r = bilerp(grad2(hslice(inthash(X0, Y0), shift), fx, fy),
                 grad2(hslice(inthash(X1, Y0), shift), fx - 1.0f, fy),
                 grad2(hslice(inthash(X0, Y1), shift), fx, fy - 1.0f),
                 grad2(hslice(inthash(X1, Y1), shift), fx - 1.0f, fy - 1.0f), u, v);
    return 0.6616f * r;
}

// 3D: the eight corner hashes are computed once and shared by the three
// components of the vector variant (as the reference's reimplementation at
// oslnoise.h:1866-1946 does).
struct Corners3 {
    uint32_t h[8];
};
template<class W = NoWrap> OSLMI_DEV Corners3 corners3(int X, int Y, int Z, const W& wr = W())
{
    Corners3 c;
    uint32_t X0 = wr(X, 0), X1 = wr(X + 1, 0), Y0 = wr(Y, 1), Y1 = wr(Y + 1, 1), Z0 = wr(Z, 2),
             Z1 = wr(Z + 1, 2);
    c.h[0] = inthash(X0, Y0, Z0);
    c.h[1] = inthash(X1, Y0, Z0);
    c.h[2] = inthash(X0, Y1, Z0);
    c.h[3] = inthash(X1, Y1, Z0);
    c.h[4] = inthash(X0, Y0, Z1);
    c.h[5] = inthash(X1, Y0, Z1);
    c.h[6] = inthash(X0, Y1, Z1);
    c.h[7] = inthash(X1, Y1, Z1);
    return c;
}

template<class T>
OSLMI_DEV T perlin3_eval(const Corners3& c, int shift, const T& fx, const T& fy, const T& fz,
                         const T& u, const T& v, const T& w)
{
    T fx1 = fx - 1.0f, fy1 = fy - 1.0f, fz1 = fz - 1.0f;
    T r = trilerp(grad3(hslice(c.h[0], shift), fx, fy, fz),
                  grad3(hslice(c.h[1], shift), fx1, fy, fz),
                  grad3(hslice(c.h[2], shift), fx, fy1, fz),
                  grad3(hslice(c.h[3], shift), fx1, fy1, fz),
                  grad3(hslice(c.h[4], shift), fx, fy, fz1),
                  grad3(hslice(c.h[5], shift), fx1, fy, fz1),
                  grad3(hslice(c.h[6], shift), fx, fy1, fz1),
                  grad3(hslice(c.h[7], shift), fx1, fy1, fz1), u, v, w);
    return 0.9820f * r;
}

template<class T, class W = NoWrap>
OSLMI_DEV T perlin3(const T& x, const T& y, const T& z, const W& wr = W())
{
    int X, Y, Z;
    T fx = floorfrac(x, &X), fy = floorfrac(y, &Y), fz = floorfrac(z, &Z);
    Corners3 c = corners3(X, Y, Z, wr);
    return perlin3_eval(c, 0, fx, fy, fz, fade(fx), fade(fy), fade(fz));
}

// Vector perlin: the three components as float/Df results.
template<class T, class W = NoWrap>
OSLMI_DEV void vperlin3(const T& x, const T& y, const T& z, T r[3], const W& wr = W())
{
    int X, Y, Z;
    T fx = floorfrac(x, &X), fy = floorfrac(y, &Y), fz = floorfrac(z, &Z);
    T u = fade(fx), v = fade(fy), w = fade(fz);
    Corners3 c = corners3(X, Y, Z, wr);
    r[0]       = perlin3_eval(c, 0, fx, fy, fz, u, v, w);
    r[1]       = perlin3_eval(c, 8, fx, fy, fz, u, v, w);
    r[2]       = perlin3_eval(c, 16, fx, fy, fz, u, v, w);
}

template<class T, class WR = NoWrap>
OSLMI_DEV T perlin4(const T& x, const T& y, const T& z, const T& w, int shift, const WR& wr = WR())
{
    int X, Y, Z, W;
    T fx = floorfrac(x, &X), fy = floorfrac(y, &Y), fz = floorfrac(z, &Z), fw = floorfrac(w, &W);
    T u = fade(fx), v = fade(fy), t = fade(fz), s = fade(fw);
    T fx1 = fx - 1.0f, fy1 = fy - 1.0f, fz1 = fz - 1.0f, fw1 = fw - 1.0f;
    uint32_t X0 = wr(X, 0), X1 = wr(X + 1, 0), Y0 = wr(Y, 1), Y1 = wr(Y + 1, 1), Z0 = wr(Z, 2),
             Z1 = wr(Z + 1, 2), W0 = wr(W, 3), W1 = wr(W + 1, 3);
#define OSLMI_G4(a, b, c, d, px, py, pz, pw) grad4(hslice(inthash(a, b, c, d), shift), px, py, pz, pw)
    T lo = trilerp(OSLMI_G4(X0, Y0, Z0, W0, fx, fy, fz, fw), OSLMI_G4(X1, Y0, Z0, W0, fx1, fy, fz, fw),
                   OSLMI_G4(X0, Y1, Z0, W0, fx, fy1, fz, fw), OSLMI_G4(X1, Y1, Z0, W0, fx1, fy1, fz, fw),
                   OSLMI_G4(X0, Y0, Z1, W0, fx, fy, fz1, fw), OSLMI_G4(X1, Y0, Z1, W0, fx1, fy, fz1, fw),
                   OSLMI_G4(X0, Y1, Z1, W0, fx, fy1, fz1, fw),
                   OSLMI_G4(X1, Y1, Z1, W0, fx1, fy1, fz1, fw), u, v, t);
    T hi = trilerp(OSLMI_G4(X0, Y0, Z0, W1, fx, fy, fz, fw1), OSLMI_G4(X1, Y0, Z0, W1, fx1, fy, fz, fw1),
                   OSLMI_G4(X0, Y1, Z0, W1, fx, fy1, fz, fw1),
                   OSLMI_G4(X1, Y1, Z0, W1, fx1, fy1, fz, fw1),
                   OSLMI_G4(X0, Y0, Z1, W1, fx, fy, fz1, fw1),
                   OSLMI_G4(X1, Y0, Z1, W1, fx1, fy, fz1, fw1),
                   OSLMI_G4(X0, Y1, Z1, W1, fx, fy1, fz1, fw1),
                   OSLMI_G4(X1, Y1, Z1, W1, fx1, fy1, fz1, fw1), u, v, t);
#undef OSLMI_G4
    return 0.8344f * lerp(lo, hi, s);
}

// --------------------------------------------------------------- simplex
OSLMI_DEV uint32_t scramble(uint32_t v0, uint32_t v1, uint32_t v2)
{
    return bjfinal(v0, v1, v2 ^ 0xdeadbeefu);
}
// 3D gradient table (12 cube edges + 4 repeats, sfm_simplex.h:129-183) as
// per-component bit masks: bit i set in P/N means component = +1/-1.
OSLMI_DEV void simplex_grad3(uint32_t h, float g[3])
{
    uint32_t bit = 1u << (h & 15u);
    g[0] = (bit & 0x1311u) ? 1.0f : ((bit & 0x2C44u) ? -1.0f : 0.0f);
    g[1] = (bit & 0x4622u) ? 1.0f : ((bit & 0x8988u) ? -1.0f : 0.0f);
    g[2] = (bit & 0x300Fu) ? 1.0f : ((bit & 0xC0F0u) ? -1.0f : 0.0f);
}
OSLMI_DEV void simplex_grad2(uint32_t h, float g[2])
{
    uint32_t bit = 1u << (h & 7u);
    g[0] = (bit & 0x8Au) ? 1.0f : ((bit & 0x15u) ? -1.0f : 0.0f);
    g[1] = (bit & 0x58u) ? 1.0f : ((bit & 0xA1u) ? -1.0f : 0.0f);
}

OSLMI_DEV float simplex1(float x, uint32_t seed, float* dnoise)
{
    int i0    = ifloor(x);
    int i1    = i0 + 1;
    float x0  = x - (float)i0;
    float x1  = x0 - 1.0f;
    float x20 = x0 * x0;
    float t0  = 1.0f - x20;
    float t20 = t0 * t0;
    float t40 = t20 * t20;
    int h0    = (int)scramble((uint32_t)i0, seed, 0u);
    float gx0 = 1.0f + (float)(h0 & 7);
    if (h0 & 8)
        gx0 = -gx0;
    float n0  = t40 * gx0 * x0;
    float x21 = x1 * x1;
    float t1  = 1.0f - x21;
    float t21 = t1 * t1;
    float t41 = t21 * t21;
    int h1    = (int)scramble((uint32_t)i1, seed, 0u);
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

OSLMI_DEV float simplex2(float x, float y, uint32_t seed, float* dnoise)
{
    const float F2 = 0.366025403f;
    const float G2 = 0.211324865f;
    float s        = (x + y) * F2;
    float xs = x + s, ys = y + s;
    int i = ifloor(xs), j = ifloor(ys);
    float t  = (float)(i + j) * G2;
    float X0 = (float)i - t, Y0 = (float)j - t;
    float x0 = x - X0, y0 = y - Y0;
    int i1 = (x0 > y0);
    int j1 = !i1;
    float x1 = x0 - (float)i1 + G2;
    float y1 = y0 - (float)j1 + G2;
    float x2 = x0 - 1.0f + 2.0f * G2;
    float y2 = y0 - 1.0f + 2.0f * G2;
    float g0[2], g1[2], g2[2];
    float t0  = 0.5f - x0 * x0 - y0 * y0;
    float t20 = t0 * t0, t40 = t20 * t20;
    simplex_grad2(scramble((uint32_t)i, (uint32_t)j, seed), g0);
    float n0 = t40 * (g0[0] * x0 + g0[1] * y0);
    if (t0 < 0.0f)
        n0 = 0.0f;
    float t1  = 0.5f - x1 * x1 - y1 * y1;
    float t21 = t1 * t1, t41 = t21 * t21;
    simplex_grad2(scramble((uint32_t)(i + i1), (uint32_t)(j + j1), seed), g1);
    float n1 = t41 * (g1[0] * x1 + g1[1] * y1);
    if (t1 < 0.0f)
        n1 = 0.0f;
    float t2  = 0.5f - x2 * x2 - y2 * y2;
    float t22 = t2 * t2, t42 = t22 * t22;
    simplex_grad2(scramble((uint32_t)(i + 1), (uint32_t)(j + 1), seed), g2);
    float n2 = t42 * (g2[0] * x2 + g2[1] * y2);
    if (t2 < 0.0f)
        n2 = 0.0f;
    const float scale = 64.0f;
    float noise       = scale * (n0 + n1 + n2);
    if (dnoise) {
        if (t0 < 0.0f)
            g0[0] = g0[1] = 0.0f;
        float temp0 = t20 * t0 * (g0[0] * x0 + g0[1] * y0);
        float dx = temp0 * x0, dy = temp0 * y0;
        if (t1 < 0.0f)
            g1[0] = g1[1] = 0.0f;
        float temp1 = t21 * t1 * (g1[0] * x1 + g1[1] * y1);
        dx += temp1 * x1;
        dy += temp1 * y1;
        if (t2 < 0.0f)
            g2[0] = g2[1] = 0.0f;
        float temp2 = t22 * t2 * (g2[0] * x2 + g2[1] * y2);
        dx += temp2 * x2;
        dy += temp2 * y2;
        dx *= -8.0f;
        dy *= -8.0f;
        dx += t40 * g0[0] + t41 * g1[0] + t42 * g2[0];
        dy += t40 * g0[1] + t41 * g1[1] + t42 * g2[1];
        dnoise[0] = dx * scale;
        dnoise[1] = dy * scale;
    }
    return noise;
}

template<bool DERIVS>
OSLMI_DEV float simplex3(float x, float y, float z, uint32_t seed, float* dnoise)
{
    const float F3 = 0.333333333f;
    const float G3 = 0.166666667f;
    float s        = (x + y + z) * F3;
    float xs = x + s, ys = y + s, zs = z + s;
    int i = ifloor(xs), j = ifloor(ys), k = ifloor(zs);
    float t  = (float)(i + j + k) * G3;
    float X0 = (float)i - t, Y0 = (float)j - t, Z0 = (float)k - t;
    float x0 = x - X0, y0 = y - Y0, z0 = z - Z0;
    int bg0 = (x0 >= y0), bg1 = (y0 >= z0), bg2 = (x0 >= z0);
    int nbg0 = !bg0, nbg1 = !bg1, nbg2 = !bg2;
    int i1 = bg0 & (bg1 | bg2);
    int j1 = nbg0 & bg1;
    int k1 = nbg1 & ((bg0 & nbg2) | nbg0);
    int i2 = bg0 | (bg1 & bg2);
    int j2 = bg1 | nbg0;
    int k2 = (bg0 & nbg1) | (nbg0 & (nbg1 | nbg2));
    float x1 = x0 - (float)i1 + G3, y1 = y0 - (float)j1 + G3, z1 = z0 - (float)k1 + G3;
    float x2 = x0 - (float)i2 + 2.0f * G3, y2 = y0 - (float)j2 + 2.0f * G3,
          z2 = z0 - (float)k2 + 2.0f * G3;
    float x3 = x0 - 1.0f + 3.0f * G3, y3 = y0 - 1.0f + 3.0f * G3, z3 = z0 - 1.0f + 3.0f * G3;
    uint32_t ks = scramble((uint32_t)k, seed, 0u);
    uint32_t ks1 = scramble((uint32_t)(k + k1), seed, 0u);
    uint32_t ks2 = scramble((uint32_t)(k + k2), seed, 0u);
    uint32_t ks3 = scramble((uint32_t)(k + 1), seed, 0u);
    float g0[3], g1[3], g2[3], g3[3];
    simplex_grad3(scramble((uint32_t)i, (uint32_t)j, ks), g0);
    simplex_grad3(scramble((uint32_t)(i + i1), (uint32_t)(j + j1), ks1), g1);
    simplex_grad3(scramble((uint32_t)(i + i2), (uint32_t)(j + j2), ks2), g2);
    simplex_grad3(scramble((uint32_t)(i + 1), (uint32_t)(j + 1), ks3), g3);
    float t0  = 0.5f - x0 * x0 - y0 * y0 - z0 * z0;
    float t20 = t0 * t0, t40 = t20 * t20;
    float n0 = t40 * (g0[0] * x0 + g0[1] * y0 + g0[2] * z0);
    if (t0 < 0.0f)
        n0 = 0.0f;
    float t1  = 0.5f - x1 * x1 - y1 * y1 - z1 * z1;
    float t21 = t1 * t1, t41 = t21 * t21;
    float n1 = t41 * (g1[0] * x1 + g1[1] * y1 + g1[2] * z1);
    if (t1 < 0.0f)
        n1 = 0.0f;
    float t2  = 0.5f - x2 * x2 - y2 * y2 - z2 * z2;
    float t22 = t2 * t2, t42 = t22 * t22;
    float n2 = t42 * (g2[0] * x2 + g2[1] * y2 + g2[2] * z2);
    if (t2 < 0.0f)
        n2 = 0.0f;
    float t3  = 0.5f - x3 * x3 - y3 * y3 - z3 * z3;
    float t23 = t3 * t3, t43 = t23 * t23;
    float n3 = t43 * (g3[0] * x3 + g3[1] * y3 + g3[2] * z3);
    if (t3 < 0.0f)
        n3 = 0.0f;
    const float scale = 68.0f;
    float noise       = scale * (n0 + n1 + n2 + n3);
    if (DERIVS) {
        if (t0 < 0.0f)
            g0[0] = g0[1] = g0[2] = 0.0f;
        float temp0 = t20 * t0 * (g0[0] * x0 + g0[1] * y0 + g0[2] * z0);
        float dx = temp0 * x0, dy = temp0 * y0, dz = temp0 * z0;
        if (t1 < 0.0f)
            g1[0] = g1[1] = g1[2] = 0.0f;
        float temp1 = t21 * t1 * (g1[0] * x1 + g1[1] * y1 + g1[2] * z1);
        dx += temp1 * x1;
        dy += temp1 * y1;
        dz += temp1 * z1;
        if (t2 < 0.0f)
            g2[0] = g2[1] = g2[2] = 0.0f;
        float temp2 = t22 * t2 * (g2[0] * x2 + g2[1] * y2 + g2[2] * z2);
        dx += temp2 * x2;
        dy += temp2 * y2;
        dz += temp2 * z2;
        if (t3 < 0.0f)
            g3[0] = g3[1] = g3[2] = 0.0f;
        float temp3 = t23 * t3 * (g3[0] * x3 + g3[1] * y3 + g3[2] * z3);
        dx += temp3 * x3;
        dy += temp3 * y3;
        dz += temp3 * z3;
        dx *= -8.0f;
        dy *= -8.0f;
        dz *= -8.0f;
        dx += t40 * g0[0] + t41 * g1[0] + t42 * g2[0] + t43 * g3[0];
        dy += t40 * g0[1] + t41 * g1[1] + t42 * g2[1] + t43 * g3[1];
        dz += t40 * g0[2] + t41 * g1[2] + t42 * g2[2] + t43 * g3[2];
        dnoise[0] = dx * scale;
        dnoise[1] = dy * scale;
        dnoise[2] = dz * scale;
    }
    return noise;
}

// 4D gradients: the 32 edges of the tesseract (sfm_simplex.h:184-193) as
// per-component masks over h & 31 (bit set in P / N: component +1 / -1).
OSLMI_DEV void simplex_grad4(uint32_t h, float g[4])
{
    uint32_t bit = 1u << (h & 31u);
    g[0] = (bit & 0x0F0F0F00u) ? 1.0f : ((bit & 0xF0F0F000u) ? -1.0f : 0.0f);
    g[1] = (bit & 0x3333000Fu) ? 1.0f : ((bit & 0xCCCC00F0u) ? -1.0f : 0.0f);
    g[2] = (bit & 0x55003333u) ? 1.0f : ((bit & 0xAA00CCCCu) ? -1.0f : 0.0f);
    g[3] = (bit & 0x00555555u) ? 1.0f : ((bit & 0x00AAAAAAu) ? -1.0f : 0.0f);
}

// 4D simplex noise (simplexnoise4, sfm_simplex.h:970-1230). The simplex's
// corner order comes from each coordinate's rank among x0, y0, z0, w0 under
// the reference's six strict comparisons (a coordinate that loses a tie
// ranks below); corner q (1..3) steps along the coordinates ranked 4 - q
// and above. This equals the reference's 64-entry corner table for every
// comparison pattern real inputs produce.
template<bool DERIVS>
OSLMI_DEV float simplex4(float x, float y, float z, float w, uint32_t seed, float* dnoise)
{
    const float F4 = 0.309016994f;
    const float G4 = 0.138196601f;
    float s        = (x + y + z + w) * F4;
    float xs = x + s, ys = y + s, zs = z + s, ws = w + s;
    int i = ifloor(xs), j = ifloor(ys), k = ifloor(zs), l = ifloor(ws);
    float t  = (float)(i + j + k + l) * G4;
    float x0 = x - ((float)i - t), y0 = y - ((float)j - t);
    float z0 = z - ((float)k - t), w0 = w - ((float)l - t);
    const int xy = x0 > y0, xz = x0 > z0, yz = y0 > z0, xw = x0 > w0, yw = y0 > w0, zw = z0 > w0;
    const int rx = xy + xz + xw;
    const int ry = (1 - xy) + yz + yw;
    const int rz = (1 - xz) + (1 - yz) + zw;
    const int rw = (1 - xw) + (1 - yw) + (1 - zw);
    float cx[5], cy[5], cz[5], cw[5];
    int ci[5][4];
    cx[0] = x0, cy[0] = y0, cz[0] = z0, cw[0] = w0;
    ci[0][0] = i, ci[0][1] = j, ci[0][2] = k, ci[0][3] = l;
#pragma unroll
    for (int q = 1; q <= 3; ++q) {
        const int oi = rx >= 4 - q, oj = ry >= 4 - q, ok = rz >= 4 - q, ol = rw >= 4 - q;
        const float gq = q == 1 ? G4 : (q == 2 ? 2.0f * G4 : 3.0f * G4);
        cx[q] = x0 - (float)oi + gq;
        cy[q] = y0 - (float)oj + gq;
        cz[q] = z0 - (float)ok + gq;
        cw[q] = w0 - (float)ol + gq;
        ci[q][0] = i + oi, ci[q][1] = j + oj, ci[q][2] = k + ok, ci[q][3] = l + ol;
    }
    cx[4] = x0 - 1.0f + 4.0f * G4, cy[4] = y0 - 1.0f + 4.0f * G4;
    cz[4] = z0 - 1.0f + 4.0f * G4, cw[4] = w0 - 1.0f + 4.0f * G4;
    ci[4][0] = i + 1, ci[4][1] = j + 1, ci[4][2] = k + 1, ci[4][3] = l + 1;
    float g[5][4], tt[5], t2[5], t4[5];
    float noise_sum = 0.0f;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const uint32_t kl = scramble((uint32_t)ci[q][2], (uint32_t)ci[q][3], seed);
        simplex_grad4(scramble((uint32_t)ci[q][0], (uint32_t)ci[q][1], kl), g[q]);
        tt[q] = 0.5f - cx[q] * cx[q] - cy[q] * cy[q] - cz[q] * cz[q] - cw[q] * cw[q];
        t2[q] = tt[q] * tt[q];
        t4[q] = t2[q] * t2[q];
        float nq = t4[q] * (g[q][0] * cx[q] + g[q][1] * cy[q] + g[q][2] * cz[q] + g[q][3] * cw[q]);
        if (tt[q] < 0.0f)
            nq = 0.0f;
        noise_sum = q == 0 ? nq : noise_sum + nq;
    }
    const float scale = 54.0f;
    float noise       = scale * noise_sum;
    if (DERIVS) {
        float d[4];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            if (tt[q] < 0.0f)
                g[q][0] = g[q][1] = g[q][2] = g[q][3] = 0.0f;
            float temp = t2[q] * tt[q]
                         * (g[q][0] * cx[q] + g[q][1] * cy[q] + g[q][2] * cz[q] + g[q][3] * cw[q]);
            d[0] = q == 0 ? temp * cx[q] : d[0] + temp * cx[q];
            d[1] = q == 0 ? temp * cy[q] : d[1] + temp * cy[q];
            d[2] = q == 0 ? temp * cz[q] : d[2] + temp * cz[q];
            d[3] = q == 0 ? temp * cw[q] : d[3] + temp * cw[q];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            d[a] *= -8.0f;
            d[a] += t4[0] * g[0][a] + t4[1] * g[1][a] + t4[2] * g[2][a] + t4[3] * g[3][a]
                    + t4[4] * g[4][a];
            dnoise[a] = d[a] * scale;
        }
    }
    return noise;
}

OSLMI_DEV Df simplex4_d(const Df& x, const Df& y, const Df& z, const Df& w, uint32_t seed)
{
    float d[4];
    float r = simplex4<true>(x.v, y.v, z.v, w.v, seed, d);
    return Df(r, d[0] * x.dx + d[1] * y.dx + d[2] * z.dx + d[3] * w.dx,
              d[0] * x.dy + d[1] * y.dy + d[2] * z.dy + d[3] * w.dy);
}

// Dual2 simplex via the chain rule (SimplexNoiseScalar, oslnoise.h:2925-2957)
OSLMI_DEV Df simplex3_d(const Df& x, const Df& y, const Df& z, uint32_t seed)
{
    float d[3];
    float r = simplex3<true>(x.v, y.v, z.v, seed, d);
    return Df(r, d[0] * x.dx + d[1] * y.dx + d[2] * z.dx, d[0] * x.dy + d[1] * y.dy + d[2] * z.dy);
}

// ----------------------------------------------------- smoothstep / mix
OSLMI_DEV float smoothstep(float e0, float e1, float x)
{
    if (x < e0)
        return 0.0f;
    else if (x >= e1)
        return 1.0f;
    float t = (x - e0) / (e1 - e0);
    return (3.0f - 2.0f * t) * (t * t);
}
OSLMI_DEV Df smoothstep(const Df& e0, const Df& e1, const Df& x)
{
    if (x.v < e0.v)
        return Df(0.0f);
    else if (x.v >= e1.v)
        return Df(1.0f);
    Df t = (x - e0) / (e1 - e0);
    return (Df(3.0f) - Df(2.0f) * t) * t * t;
}
// mix of llvm_gen_mix (batched_llvm_gen.cpp:1868-1960): a*(1-x) + b*x
OSLMI_DEV float mix(float a, float b, float x) { return a * (1.0f - x) + b * x; }

}  // namespace oslmi
