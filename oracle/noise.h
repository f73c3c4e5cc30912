// CPU oracle: integer hashing and lattice noises (perlin, cell, hash).
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
#pragma once
#include "oracle_core.h"

namespace oref {

// --- Jenkins lookup3 mix/final -------------------------------------------
// src/include/OSL/oslnoise.h:154-180 restates OIIO::bjhash::bjmix/bjfinal
// (Bob Jenkins, lookup3.c, "mix" and "final").
#ifdef OREF_COUNTING
// a rotate is one instruction (v_alignbit_b32 on CDNA)
inline uint32_t rotl32(uint32_t x, int k)
{
    OREF_COUNT(1);
    unsigned v = x.v;
    return uint32_t((v << k) | (v >> (32 - k)));
}
#else
inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
#endif

inline void bjmix(uint32_t& a, uint32_t& b, uint32_t& c)
{
    a -= c;  a ^= rotl32(c, 4);  c += b;
    b -= a;  b ^= rotl32(a, 6);  a += c;
    c -= b;  c ^= rotl32(b, 8);  b += a;
    a -= c;  a ^= rotl32(c, 16); c += b;
    b -= a;  b ^= rotl32(a, 19); a += c;
    c -= b;  c ^= rotl32(b, 4);  b += a;
}

inline uint32_t bjfinal(uint32_t a, uint32_t b, uint32_t c)
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

// reference_inthash<N> (oslnoise.h:192-215): lookup3 over N words, seed
// 0xdeadbeef + (N<<2) + 13.  The unrolled inthash(k0..k4) overloads
// (oslnoise.h:229-289) compute the same values.
uint32_t inthash_n(const uint32_t* k, int n);
inline uint32_t inthash(uint32_t k0) { return inthash_n(&k0, 1); }
inline uint32_t inthash(uint32_t k0, uint32_t k1)
{
    uint32_t k[2] = { k0, k1 };
    return inthash_n(k, 2);
}
inline uint32_t inthash(uint32_t k0, uint32_t k1, uint32_t k2)
{
    uint32_t k[3] = { k0, k1, k2 };
    return inthash_n(k, 3);
}
inline uint32_t inthash(uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3)
{
    uint32_t k[4] = { k0, k1, k2, k3 };
    return inthash_n(k, 4);
}
inline uint32_t inthash(uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3,
                        uint32_t k4)
{
    uint32_t k[5] = { k0, k1, k2, k3, k4 };
    return inthash_n(k, 5);
}

// oslnoise.h:135-141: bits * float(1/(2^32-1)) with the u32->f32 rounding
// of the multiplication's implicit conversion.
inline float bits_to_01(uint32_t bits)
{
    const float k = (float)(1.0 / 4294967295.0);
    return (float)bits * k;
}

// OIIO::ifloor
inline int ifloor(float x) { return (int)floorf(x); }

// imod (oslnoise.h:847-859): the remainder of a / b in [0, b) for b > 0.
inline int imod(int a, int b)
{
    int r = a % b;
    if (r < 0)
        r += b;
    return r;
}

// The lattice period of a float period argument (HashScalarPeriodic,
// oslnoise.h:1162-1180): ifloor(p), at least 1. A floor that no int holds
// (NaN, beyond 2^31) gives 1, as the x86 truncating conversion (INT_MIN)
// followed by the clamp does.
inline int period_of(float p)
{
    float f = floorf(p);
    return (f >= 1.0f && f < 2147483648.0f) ? (int)f : 1;
}

// PeriodicAdaptionOf::wrap (oslnoise.h:532-539; gabornoise.h:185-192):
// s reduced into [0, floor(period)) with the period at least 1.
inline float wrap_period(float s, float period)
{
    period = floorf(period);
    if (period < 1.0f)
        period = 1.0f;
    return s - period * floorf(s / period);
}

// --- cell / hash noise (oslnoise.h:337-480) --------------------------------
// dim in 1..4 (the 4th coordinate is the 't' argument). vec selects the
// Vec3-valued variant, which hashes an extra trailing word 0,1,2.
float cellnoise_f(int dim, const float* p);
V3 cellnoise_v(int dim, const float* p);
float hashnoise_f(int dim, const float* p);
V3 hashnoise_v(int dim, const float* p);

// hash() shadeop: inthashf of the float bit patterns, returned as int
// (oslnoise.h:555-610; src/liboslexec/wide/wide_ophash.cpp:28-80).
int hash_op_f(int dim, const float* p);
int hash_op_i(int x);

// --- perlin ---------------------------------------------------------------
// Signed perlin noise ("perlin"/"snoise"), CGScalar branches of
// oslnoise.h:1327-2282. dim in 1..4.
float snoise_f(int dim, const float* p);
V3 snoise_v(int dim, const float* p);
D2<float> snoise_df(int dim, const D2<float>* p);
D2<V3> snoise_dv(int dim, const D2<float>* p);
// The same noise in the interpolation order of the CGDefault (SIMD)
// implementation the batched path uses for sparsely active batches;
// component comp (0..2) of a vector result.
float snoise_f_cgdefault(int dim, const float* p, int comp);
D2<float> snoise_df_cgdefault(int dim, const D2<float>* p, int comp);
// Unsigned ("uperlin"/"noise") = 0.5*(snoise+1) (oslnoise.h:2284-2400).
float unoise_f(int dim, const float* p);
V3 unoise_v(int dim, const float* p);
D2<float> unoise_df(int dim, const D2<float>* p);
D2<V3> unoise_dv(int dim, const D2<float>* p);

// Periodic perlin ("psnoise" signed, "pnoise" unsigned) and periodic
// cell/hash noise; period holds one float per input dimension.
float psnoise_f(int dim, const float* p, const float* period);
V3 psnoise_v(int dim, const float* p, const float* period);
D2<float> psnoise_df(int dim, const D2<float>* p, const float* period);
D2<V3> psnoise_dv(int dim, const D2<float>* p, const float* period);
float pnoise_f(int dim, const float* p, const float* period);
V3 pnoise_v(int dim, const float* p, const float* period);
D2<float> pnoise_df(int dim, const D2<float>* p, const float* period);
D2<V3> pnoise_dv(int dim, const D2<float>* p, const float* period);
float pcellnoise_f(int dim, const float* p, const float* period);
V3 pcellnoise_v(int dim, const float* p, const float* period);
float phashnoise_f(int dim, const float* p, const float* period);
V3 phashnoise_v(int dim, const float* p, const float* period);

// --- simplex (sfm_simplex.h, SimplexNoiseScalar oslnoise.h:2878-2990) -----
float simplex_f(int dim, const float* p, int seed, float* dnoise);
D2<float> simplex_df(int dim, const D2<float>* p, int seed);
D2<V3> simplex_dv(int dim, const D2<float>* p);

// --- smoothstep / mix -----------------------------------------------------
float smoothstep_f(float e0, float e1, float x);               // dual.h:1314
D2<float> smoothstep_d(D2<float> e0, D2<float> e1, D2<float> x); // dual.h:1324

}  // namespace oref
