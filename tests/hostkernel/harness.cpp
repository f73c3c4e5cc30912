// Host build of one emitted group kernel, for CPU tests of the emitter's
// lowering: the kernel source (KERNEL, a file name) is compiled as plain
// C++ with the device qualifiers defined away, and its grid-stride loop
// runs on one host thread over every point (blockIdx = threadIdx = 0,
// blockDim = gridDim = 1). Reads n, then 3-plane P, N, dPdu, dPdv, wo, wi
// (float32), and the libbsdl table file; writes the 4 eval planes, or with
// a fourth argument "export" the closure export (1 + 8 records of
// OSLMI_CLOSURE_RECORD planes) of a kernel built with it, or with "out<k>"
// the group's first k output planes.
//
// usage: harness <input> <output> <bsdl_luts.bin> [export | out<k>]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace __hip_internal {
typedef ::uint32_t uint32_t;
typedef ::int32_t int32_t;
typedef ::int64_t int64_t;
typedef ::uint64_t uint64_t;
}  // namespace __hip_internal
using std::isfinite;
using std::isinf;
using std::isnan;
#define __HIPCC_RTC__ 1
#define OSLMI_DEV inline
#define OSLMI_HD inline
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
#define amdgpu_waves_per_eu(x) unused
static inline float __fmaf_rn(float a, float b, float c) { return fmaf(a, b, c); }
static inline float __uint_as_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t __float_as_uint(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline int __float_as_int(float f)
{
    int u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float __int_as_float(int u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
struct HostDim {
    unsigned x;
};
static HostDim blockIdx { 0 }, threadIdx { 0 }, blockDim { 1 }, gridDim { 1 };
static inline void __syncthreads() {}

#include KERNEL

int main(int argc, char** argv)
{
    if (argc != 4 && argc != 5)
        return 1;
    const bool xp   = argc == 5 && strcmp(argv[4], "export") == 0;
    const int nouts = (argc == 5 && strncmp(argv[4], "out", 3) == 0) ? atoi(argv[4] + 3) : 0;
    if (argc == 5 && !xp && (nouts < 1 || nouts > OSLMI_MAX_OUTPUT_PLANES))
        return 1;
    FILE* f = fopen(argv[1], "rb");
    int64_t n = 0;
    if (!f || fread(&n, 8, 1, f) != 1 || n < 1)
        return 2;
    std::vector<float> P(3 * n), N(3 * n), du(3 * n), dv(3 * n), wo(3 * n), wi(3 * n), ev(4 * n);
    for (std::vector<float>* v : { &P, &N, &du, &dv, &wo, &wi })
        if (fread(v->data(), 4, 3 * n, f) != (size_t)(3 * n))
            return 3;
    fclose(f);
    std::vector<float> lut(8448);
    FILE* t = fopen(argv[3], "rb");
    char magic[8];
    if (!t || fread(magic, 1, 8, t) != 8 || fread(lut.data(), 4, 8448, t) != 8448)
        return 4;
    fclose(t);
    oslmi::GroupArgs A;
    memset(&A, 0, sizeof A);
    A.n                 = n;
    A.g[oslmi::kG_P]    = P.data();
    A.g[oslmi::kG_N]    = N.data();
    A.g[oslmi::kG_dPdu] = du.data();
    A.g[oslmi::kG_dPdv] = dv.data();
    A.wo                = wo.data();
    A.bsdl              = lut.data();
    std::vector<float> clo, outs;
    if (nouts) {
        outs.assign((size_t)nouts * n, 0.0f);
        for (int j = 0; j < nouts; ++j)
            A.out[j] = outs.data() + (size_t)j * n;
    } else if (xp) {
        clo.assign((size_t)(1 + 8 * OSLMI_CLOSURE_RECORD) * n, 0.0f);
        A.clo     = clo.data();
        A.clo_max = 8;
    } else {
        A.wi   = wi.data();
        A.eval = ev.data();
    }
    oslmi_group_main(A);
    FILE* o = fopen(argv[2], "wb");
    const std::vector<float>& res = nouts ? outs : xp ? clo : ev;
    if (!o || fwrite(res.data(), 4, res.size(), o) != res.size())
        return 5;
    fclose(o);
    return 0;
}
