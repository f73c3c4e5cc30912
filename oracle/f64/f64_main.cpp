// FP64 evaluation build of the oracle (test infrastructure).
//
// The oracle sources recompiled with `float` replaced by binary64 and the
// transcendental approximations of omath.h replaced by the exact functions
// they stand for (OREF_F64). The result is the shading path's mathematics
// evaluated in double precision: the reference against which the binary32
// results (the GPU's and the binary32 oracle's alike) are given an error.
// Integer hashing, tables and control flow are the same code; coordinates
// are hashed by their binary32 value (oracle_core.h f2u).
//
// Exports oref64_* wrappers over double planes; everything else in this
// library is hidden (built with -fvisibility=hidden).
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

typedef float oref_f32;

#define OREF_F64 1
#define float double
#define fmaf fma
#define sqrtf sqrt
#define fabsf fabs
#define copysignf copysign
#define floorf floor
#define rintf rint
#define fminf fmin
#define fmaxf fmax
#define powf pow
#define expf exp
#define exp2f exp2
#define expm1f expm1
#define logf log
#define sinf sin
#define cosf cos
#include "../noise.cpp"
#include "../gabor.cpp"
#include "../color.cpp"
#include "../closure.cpp"
#include "../shade.cpp"
#include "../mtx.cpp"
#include "../mtx_api.cpp"
#include "../oracle_api.cpp"
#undef float

#define OREF64_API extern "C" __attribute__((visibility("default")))

// fbm_cell: P 3 planes -> Cout 3 planes
OREF64_API void oref64_shade_c1(const double* P, double* Cout, int64_t n, int nthreads)
{
    oref_shade_c1(P, Cout, n, nthreads);
}

// layered3: P, N, wo, wi 3 planes each, dPdx/dPdy uniform -> 4 planes
OREF64_API void oref64_shade_c2(const double* P, const double* N, const double* wo,
                                const double* wi, const double* dPdx, const double* dPdy,
                                double* out, int64_t n, int nthreads)
{
    oref_shade_c2(P, N, wo, wi, dPdx, dPdy, out, nullptr, n, nthreads);
}

// layered MaterialX closure, planes as oref_mtx_layered
OREF64_API void oref64_mtx_layered(const double* mat, const double* N, const double* V,
                                   const double* L, const double* prm, const double* rnd,
                                   double* eval, double* samp, int64_t n, int nthreads)
{
    oref_mtx_layered(mat, N, V, L, prm, rnd, eval, samp, n, nthreads);
}

// math building blocks (fn as oref_math)
OREF64_API void oref64_math(int fn, const double* in, double* out, int64_t n)
{
    oref_math(fn, in, out, n);
}
