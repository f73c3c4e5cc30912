// Wide shadeop kernels (noise, hash) behind the C ABI in include/oslmi.h.
//
// Each kernel maps one work-item to one shading point and reads/writes
// SoA planes with unit stride, so every wavefront load is one coalesced
// 256-byte transaction per plane.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "oslmi.h"
#include "oslmi_internal.h"
#include "oslmi_noise.h"

namespace oslmi {

enum NoiseKind { kPerlin = 0, kUPerlin, kCell, kHash, kSimplex, kUSimplex };

// Perlin of INDIM coordinates, OUTDIM components (8-bit slices of one hash
// per corner for the vector form), over the lattice map W.
template<class T, int OUTDIM, int INDIM, class W>
__device__ __forceinline__ void perlin_n(const T* p, T* r, const W& w)
{
    if (INDIM == 3 && OUTDIM == 3) {
        vperlin3<T, W>(p[0], p[1], p[2], r, w);
    } else {
#pragma unroll
        for (int c = 0; c < OUTDIM; ++c) {
            int sh = 8 * c;
            r[c]   = INDIM == 1   ? perlin1<T, W>(p[0], sh, w)
                     : INDIM == 2 ? perlin2<T, W>(p[0], p[1], sh, w)
                     : INDIM == 3 ? perlin3<T, W>(p[0], p[1], p[2], w)
                                  : perlin4<T, W>(p[0], p[1], p[2], p[3], sh, w);
        }
    }
}

// the lattice periods of point i (PER) or the identity map
template<int INDIM>
__device__ __forceinline__ Wrap point_wrap(const float* __restrict__ per, int64_t n, int64_t i)
{
    float q[4] = { 1.f, 1.f, 1.f, 1.f };
#pragma unroll
    for (int c = 0; c < INDIM; ++c)
        q[c] = per[c * n + i];
    return make_wrap(q[0], q[1], q[2], q[3]);
}

// PER: periodic form, per holds INDIM period planes (no derivatives).
template<int KIND, int OUTDIM, int INDIM, bool PER>
__global__ __launch_bounds__(256) void k_noise(const float* __restrict__ in,
                                               const float* __restrict__ per,
                                               float* __restrict__ out, int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float p[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int c = 0; c < INDIM; ++c)
            p[c] = in[c * n + i];
        float r[3] = { 0.f, 0.f, 0.f };
        if (KIND == kCell || KIND == kHash) {
            constexpr bool CELL = (KIND == kCell);
            if (PER)
#pragma unroll
                for (int c = 0; c < INDIM; ++c)
                    p[c] = wrap_period(p[c], per[c * n + i]);
            if (OUTDIM == 1) {
                r[0] = INDIM == 1   ? intnoise1<CELL>(p[0])
                       : INDIM == 2 ? intnoise2<CELL>(p[0], p[1])
                       : INDIM == 3 ? intnoise3<CELL>(p[0], p[1], p[2])
                                    : intnoise4<CELL>(p[0], p[1], p[2], p[3]);
            } else {
                V3 v = INDIM == 1   ? vintnoise1<CELL>(p[0])
                       : INDIM == 2 ? vintnoise2<CELL>(p[0], p[1])
                       : INDIM == 3 ? vintnoise3<CELL>(p[0], p[1], p[2])
                                    : vintnoise4<CELL>(p[0], p[1], p[2], p[3]);
                r[0] = v.x;
                r[1] = v.y;
                r[2] = v.z;
            }
        } else if (KIND == kPerlin || KIND == kUPerlin) {
            if (PER)
                perlin_n<float, OUTDIM, INDIM>(p, r, point_wrap<INDIM>(per, n, i));
            else
                perlin_n<float, OUTDIM, INDIM>(p, r, NoWrap());
            if (KIND == kUPerlin)
#pragma unroll
                for (int c = 0; c < OUTDIM; ++c)
                    r[c] = 0.5f * (r[c] + 1.0f);
        } else {  // simplex
#pragma unroll
            for (int c = 0; c < OUTDIM; ++c) {
                r[c] = INDIM == 1   ? simplex1(p[0], (uint32_t)c, nullptr)
                       : INDIM == 2 ? simplex2(p[0], p[1], (uint32_t)c, nullptr)
                       : INDIM == 3 ? simplex3<false>(p[0], p[1], p[2], (uint32_t)c, nullptr)
                                    : simplex4<false>(p[0], p[1], p[2], p[3], (uint32_t)c, nullptr);
                if (KIND == kUSimplex)
                    r[c] = 0.5f * (r[c] + 1.0f);
            }
        }
#pragma unroll
        for (int c = 0; c < OUTDIM; ++c)
            out[c * n + i] = r[c];
    }
}

// Dual2 variants (value + dx + dy planes in and out).
template<int KIND, int OUTDIM, int INDIM, bool PER>
__global__ __launch_bounds__(256) void k_noise_d(const float* __restrict__ in,
                                                 const float* __restrict__ per,
                                                 float* __restrict__ out, int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Df p[4];
#pragma unroll
        for (int c = 0; c < INDIM; ++c)
            p[c] = Df(in[c * n + i], in[(INDIM + c) * n + i], in[(2 * INDIM + c) * n + i]);
        Df r[3];
        if (KIND == kPerlin || KIND == kUPerlin) {
            if (PER)
                perlin_n<Df, OUTDIM, INDIM>(p, r, point_wrap<INDIM>(per, n, i));
            else
                perlin_n<Df, OUTDIM, INDIM>(p, r, NoWrap());
            if (KIND == kUPerlin)
#pragma unroll
                for (int c = 0; c < OUTDIM; ++c)
                    r[c] = 0.5f * (r[c] + 1.0f);
        } else {  // simplex, derivatives by the chain rule
#pragma unroll
            for (int c = 0; c < OUTDIM; ++c) {
                float d[3] = { 0.f, 0.f, 0.f };
                float v;
                if (INDIM == 1) {
                    v    = simplex1(p[0].v, (uint32_t)c, d);
                    r[c] = Df(v, d[0] * p[0].dx, d[0] * p[0].dy);
                } else if (INDIM == 2) {
                    v    = simplex2(p[0].v, p[1].v, (uint32_t)c, d);
                    r[c] = Df(v, d[0] * p[0].dx + d[1] * p[1].dx, d[0] * p[0].dy + d[1] * p[1].dy);
                } else if (INDIM == 3) {
                    r[c] = simplex3_d(p[0], p[1], p[2], (uint32_t)c);
                } else {
                    r[c] = simplex4_d(p[0], p[1], p[2], p[3], (uint32_t)c);
                }
                if (KIND == kUSimplex)
                    r[c] = Df(0.5f * (r[c].v + 1.0f), 0.5f * r[c].dx, 0.5f * r[c].dy);
            }
        }
#pragma unroll
        for (int c = 0; c < OUTDIM; ++c) {
            out[c * n + i]                = r[c].v;
            out[(OUTDIM + c) * n + i]     = r[c].dx;
            out[(2 * OUTDIM + c) * n + i] = r[c].dy;
        }
    }
}

__global__ __launch_bounds__(256) void k_zero(float* __restrict__ out, int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = 0.0f;
}

template<int INDIM>
__global__ __launch_bounds__(256) void k_hash(const float* __restrict__ in, int32_t* __restrict__ out,
                                              int64_t n)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float p[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int c = 0; c < INDIM; ++c)
            p[c] = in[c * n + i];
        out[i] = INDIM == 1   ? hash_op1(p[0])
                 : INDIM == 2 ? hash_op2(p[0], p[1])
                 : INDIM == 3 ? hash_op3(p[0], p[1], p[2])
                              : hash_op4(p[0], p[1], p[2], p[3]);
    }
}

typedef void (*NoiseKernel)(const float*, const float*, float*, int64_t);

template<int KIND, int OUTDIM, bool PER> static NoiseKernel pick_plain(int indim)
{
    switch (indim) {
    case 1: return k_noise<KIND, OUTDIM, 1, PER>;
    case 2: return k_noise<KIND, OUTDIM, 2, PER>;
    case 3: return k_noise<KIND, OUTDIM, 3, PER>;
    case 4: return k_noise<KIND, OUTDIM, 4, PER>;
    }
    return nullptr;
}
template<int KIND, int OUTDIM, bool PER> static NoiseKernel pick_deriv(int indim)
{
    switch (indim) {
    case 1: return k_noise_d<KIND, OUTDIM, 1, PER>;
    case 2: return k_noise_d<KIND, OUTDIM, 2, PER>;
    case 3: return k_noise_d<KIND, OUTDIM, 3, PER>;
    case 4: return k_noise_d<KIND, OUTDIM, 4, PER>;
    }
    return nullptr;
}
template<int KIND, bool PER = false> static NoiseKernel pick(int outdim, int indim, bool derivs)
{
    if (derivs)
        return outdim == 1 ? pick_deriv<KIND, 1, PER>(indim) : pick_deriv<KIND, 3, PER>(indim);
    return outdim == 1 ? pick_plain<KIND, 1, PER>(indim) : pick_plain<KIND, 3, PER>(indim);
}

}  // namespace oslmi

using namespace oslmi;

extern "C" int oslmi_noise(const char* name, int outdim, int indim, int derivs, const float* d_in,
                           float* d_out, int64_t n, void* stream)
{
    if (!name || n < 0 || (outdim != 1 && outdim != 3) || indim < 1 || indim > 4)
        return set_error(OSLMI_ERR_INVALID, "oslmi_noise: bad arguments");
    if (n == 0)
        return OSLMI_OK;
    if (!d_in || !d_out)
        return set_error(OSLMI_ERR_INVALID, "oslmi_noise: null buffer");
    std::string s(name);
    int kind;
    if (s == "perlin" || s == "snoise")
        kind = kPerlin;
    else if (s == "uperlin" || s == "noise")
        kind = kUPerlin;
    else if (s == "cell" || s == "cellnoise")
        kind = kCell;
    else if (s == "hash" || s == "hashnoise")
        kind = kHash;
    else if (s == "simplex")
        kind = kSimplex;
    else if (s == "usimplex")
        kind = kUSimplex;
    else
        return set_error(OSLMI_ERR_UNSUPPORTED, "oslmi_noise: unknown noise '" + s + "'");
    // cellnoise/hashnoise have no derivatives: the value planes are computed
    // by the plain kernel and the derivative planes are cleared.
    bool dv = derivs && kind != kCell && kind != kHash;
    NoiseKernel k = nullptr;
    switch (kind) {
    case kPerlin: k = pick<kPerlin>(outdim, indim, dv); break;
    case kUPerlin: k = pick<kUPerlin>(outdim, indim, dv); break;
    case kCell: k = pick<kCell>(outdim, indim, false); break;
    case kHash: k = pick<kHash>(outdim, indim, false); break;
    case kSimplex: k = pick<kSimplex>(outdim, indim, dv); break;
    case kUSimplex: k = pick<kUSimplex>(outdim, indim, dv); break;
    }
    if (!k)
        return set_error(OSLMI_ERR_UNSUPPORTED, "oslmi_noise: unsupported dimension for " + s);
    hipStream_t st = (hipStream_t)stream;
    // the value planes of a Dual2 input are its first indim planes
    dim3 grid(grid_for(n)), block(256);
    hipLaunchKernelGGL(k, grid, block, 0, st, d_in, nullptr, d_out, n);
    if (derivs && !dv)
        hipLaunchKernelGGL(k_zero, dim3(grid_for(2 * outdim * n)), block, 0, st,
                           d_out + outdim * n, 2 * outdim * n);
    return check_launch("oslmi_noise");
}

extern "C" int oslmi_pnoise(const char* name, int outdim, int indim, int derivs, const float* d_in,
                            const float* d_period, float* d_out, int64_t n, void* stream)
{
    if (!name || n < 0 || (outdim != 1 && outdim != 3) || indim < 1 || indim > 4)
        return set_error(OSLMI_ERR_INVALID, "oslmi_pnoise: bad arguments");
    if (n == 0)
        return OSLMI_OK;
    if (!d_in || !d_period || !d_out)
        return set_error(OSLMI_ERR_INVALID, "oslmi_pnoise: null buffer");
    std::string s(name);
    int kind;
    if (s == "perlin" || s == "snoise" || s == "psnoise")
        kind = kPerlin;
    else if (s == "uperlin" || s == "noise" || s == "pnoise")
        kind = kUPerlin;
    else if (s == "cell" || s == "cellnoise")
        kind = kCell;
    else if (s == "hash" || s == "hashnoise")
        kind = kHash;
    else
        return set_error(OSLMI_ERR_UNSUPPORTED, "oslmi_pnoise: periodic noise type '" + s
                                                    + "' is unknown");
    bool dv = derivs && kind != kCell && kind != kHash;
    NoiseKernel k = nullptr;
    switch (kind) {
    case kPerlin: k = pick<kPerlin, true>(outdim, indim, dv); break;
    case kUPerlin: k = pick<kUPerlin, true>(outdim, indim, dv); break;
    case kCell: k = pick<kCell, true>(outdim, indim, false); break;
    case kHash: k = pick<kHash, true>(outdim, indim, false); break;
    }
    if (!k)
        return set_error(OSLMI_ERR_UNSUPPORTED, "oslmi_pnoise: unsupported dimension for " + s);
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(grid_for(n)), block(256);
    hipLaunchKernelGGL(k, grid, block, 0, st, d_in, d_period, d_out, n);
    if (derivs && !dv)
        hipLaunchKernelGGL(k_zero, dim3(grid_for(2 * outdim * n)), block, 0, st,
                           d_out + outdim * n, 2 * outdim * n);
    return check_launch("oslmi_pnoise");
}

extern "C" int oslmi_hash(int indim, const float* d_in, int32_t* d_out, int64_t n, void* stream)
{
    if (indim < 1 || indim > 4 || n < 0)
        return set_error(OSLMI_ERR_INVALID, "oslmi_hash: bad arguments");
    if (n == 0)
        return OSLMI_OK;
    if (!d_in || !d_out)
        return set_error(OSLMI_ERR_INVALID, "oslmi_hash: null buffer");
    void (*k)(const float*, int32_t*, int64_t) = indim == 1   ? k_hash<1>
                                                 : indim == 2 ? k_hash<2>
                                                 : indim == 3 ? k_hash<3>
                                                              : k_hash<4>;
    hipLaunchKernelGGL(k, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, d_in, d_out, n);
    return check_launch("oslmi_hash");
}
