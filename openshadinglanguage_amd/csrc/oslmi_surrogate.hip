// Learned surrogate of the layered closure: a 14 -> 4x64 SiLU -> 4 MLP
// trained and evaluated with BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) on gfx950.
//
// Inputs (14): L xyz and V x, z in the shading frame (N = +z, V.y = 0),
// coat ior, coat roughness, sheen roughness, F0 rgb, F90 rgb, normalised
// affinely (x - offset) * scale in FP32 before rounding to bf16.
// Outputs (4): log of the eval weight rgb and log pdf, normalised by
// (y - mean) / std.
//
// Training (one step = one minibatch): row-major bf16 activations feed a
// generic C = A B^T GEMM kernel (one wave per 16x16 tile, K in steps of 32,
// split-K partials for the weight gradients); elementwise kernels add
// biases, apply SiLU and its derivative, form the loss gradient and
// transpose through LDS; Adam updates FP32 master weights and refreshes
// the bf16 copies. All reductions run in a fixed order, so training is
// deterministic.
//
// Inference: one fused kernel. Each wave keeps every weight fragment in the
// accumulation registers (the MFMA reads its A operand there), reads the
// bias initialisers from LDS, and walks pairs of 16-query column tiles: a
// layer's MFMA accumulators (lane: 4 consecutive features of one query, in
// VGPRs) become the next layer's B operand in place, because the packed
// weight fragments use the matching order of the 64 hidden features along
// K. That fits a wave in 240 registers, two waves per SIMD. The SiLU costs
// two transcendentals (exp2, rcp) and two VALU operations per unit, with
// the -log2(e) scaling folded into the packed weights, and the output
// epilogue is one exp2 (output normalisation folded into W5 and b5).
// The same kernel routes the queries as it reads them (one pass over the
// 14 input planes): exact zeros are written at once, queries outside the
// training domain or in a routing cell flagged by calibration are appended
// to a fallback list that the exact closure kernel then evaluates, and the
// rest are queued per wave in LDS until two 16-query tiles are ready.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "oslmi.h"
#include "oslmi_internal.h"

namespace oslmi {
namespace surrogate {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 14, NIN_PAD = 32, H = 64, NOUT = 4, NOUT_PAD = 32;
// parameter layout (floats): W1 [64][32], b1, W2..W4 [64][64], b2..b4, W5 [32][64], b5 [32]
constexpr int OFF_W[6] = { 0, 0, 2112, 6272, 10432, 14592 };
constexpr int OFF_B[6] = { 0, 2048, 6208, 10368, 14528, 16640 };
constexpr int LAYER_OUT[6] = { 0, H, H, H, H, NOUT_PAD };
constexpr int LAYER_IN[6]  = { 0, NIN_PAD, H, H, H, H };
constexpr int NPARAM       = 16672;
constexpr int EW_ROWS      = 64;    // samples per elementwise block
constexpr int KCHUNK       = 2048;  // samples per split-K partial

struct Norm {
    float in_off[NIN], in_scale[NIN];
    float out_mu[NOUT], out_sd[NOUT];
};

// Training-domain bounds (the closure workload's parameter ranges).
struct Domain {
    float lo[9], hi[9];  // coat ior, coat roughness, sheen roughness, F0 rgb, F90 rgb
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ---- training kernels --------------------------------------------------
// C[m][n] = sum_k A[m][k] B[n][k]; A, B bf16 with K contiguous; one wave
// per 16x16 tile; blockIdx.z selects a K chunk whose partial sum goes to
// C + z * cstride. M, N multiples of 16, K and kchunk multiples of 32, lda
// and ldb multiples of 8.
__global__ __launch_bounds__(64) void k_gemm_nt(int K, const bf16* __restrict__ A, int lda,
                                                const bf16* __restrict__ B, int ldb,
                                                float* __restrict__ C, int ldc, int kchunk,
                                                int64_t cstride)
{
    const int lane = threadIdx.x;
    const int r = lane & 15, g = lane >> 4;
    const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    const int kb = blockIdx.z * kchunk;
    const int ke = kb + kchunk < K ? kb + kchunk : K;
    const bf16* a = A + (int64_t)(m0 + r) * lda + 8 * g;
    const bf16* b = B + (int64_t)(n0 + r) * ldb + 8 * g;
    f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int k = kb; k < ke; k += 32) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(a + k);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(b + k);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    float* c = C + blockIdx.z * cstride;
#pragma unroll
    for (int j = 0; j < 4; j++)
        c[(int64_t)(m0 + 4 * g + j) * ldc + n0 + r] = acc[j];
}

// Normalised inputs X [B][32] and X^T [32][B] (bf16, zero padded) and
// normalised log targets Y [B][4] from query planes q [14][ldq] and exact
// outputs t [4][ldq] (weight rgb, pdf).
__global__ __launch_bounds__(256) void k_prep(int B, Norm nm, const float* __restrict__ q,
                                              const float* __restrict__ t, int64_t ldq,
                                              bf16* __restrict__ X, bf16* __restrict__ XT,
                                              float* __restrict__ Y)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B)
        return;
#pragma unroll
    for (int k = 0; k < NIN_PAD; k++) {
        float v = 0.0f;
        if (k < NIN)
            v = (q[k * ldq + s] - nm.in_off[k]) * nm.in_scale[k];
        const bf16 h = (bf16)v;
        X[(int64_t)s * NIN_PAD + k] = h;
        XT[(int64_t)k * B + s]      = h;
    }
    if (t) {
#pragma unroll
        for (int o = 0; o < NOUT; o++) {
            const float raw = t[o * ldq + s];
            const float y   = o < 3 ? logf(fmaxf(raw, 1e-6f)) : logf(fmaxf(raw, 1e-12f));
            Y[(int64_t)s * NOUT + o] = (y - nm.out_mu[o]) / nm.out_sd[o];
        }
    }
}

// Hidden layer epilogue over a block of 64 samples: Z += b, A = bf16(silu(Z)),
// A^T through LDS.
__global__ __launch_bounds__(256) void k_bias_silu(int B, float* __restrict__ Z,
                                                   const float* __restrict__ bias,
                                                   bf16* __restrict__ A, bf16* __restrict__ AT)
{
    __shared__ bf16 tile[EW_ROWS][H + 2];
    const int s0 = blockIdx.x * EW_ROWS;
    for (int e = threadIdx.x; e < EW_ROWS * H; e += blockDim.x) {
        const int s = e / H, o = e % H;
        const int64_t idx = (int64_t)(s0 + s) * H + o;
        const float z = Z[idx] + bias[o];
        Z[idx]        = z;
        const bf16 a  = (bf16)silu(z);
        A[idx]        = a;
        tile[s][o]    = a;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < EW_ROWS * H; e += blockDim.x) {
        const int o = e / EW_ROWS, s = e % EW_ROWS;
        AT[(int64_t)o * B + s0 + s] = tile[s][o];
    }
}

// Output layer and loss: pred = Z5 + b5 (4 outputs), dZ5 = 2 (pred - Y) / (4B)
// as bf16 [B][32] and [32][B]; bias-gradient and loss partials per block.
__global__ __launch_bounds__(256) void k_loss(int B, const float* __restrict__ Z5,
                                              const float* __restrict__ b5,
                                              const float* __restrict__ Y,
                                              bf16* __restrict__ dZ, bf16* __restrict__ dZT,
                                              float* __restrict__ bpart,
                                              float* __restrict__ lpart)
{
    __shared__ float red[NOUT_PAD][EW_ROWS + 1];
    __shared__ float lred[EW_ROWS];
    const int s0 = blockIdx.x * EW_ROWS;
    for (int e = threadIdx.x; e < EW_ROWS * NOUT_PAD; e += blockDim.x) {
        const int s = e / NOUT_PAD, o = e % NOUT_PAD;
        float d = 0.0f;
        if (o < NOUT) {
            const float pred = Z5[(int64_t)(s0 + s) * NOUT_PAD + o] + b5[o];
            d = pred - Y[(int64_t)(s0 + s) * NOUT + o];
        }
        red[o][s] = d;
    }
    __syncthreads();
    if (threadIdx.x < EW_ROWS) {
        const int s = threadIdx.x;
        float l = 0.0f;
        for (int o = 0; o < NOUT; o++)
            l += red[o][s] * red[o][s];
        lred[s] = l;
    }
    __syncthreads();
    const float scale = 2.0f / (4.0f * (float)B);
    for (int e = threadIdx.x; e < EW_ROWS * NOUT_PAD; e += blockDim.x) {
        const int s = e / NOUT_PAD, o = e % NOUT_PAD;
        const float g = red[o][s] * scale;
        red[o][s]     = g;
        dZ[(int64_t)(s0 + s) * NOUT_PAD + o] = (bf16)g;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < EW_ROWS * NOUT_PAD; e += blockDim.x) {
        const int o = e / EW_ROWS, s = e % EW_ROWS;
        dZT[(int64_t)o * B + s0 + s] = (bf16)red[o][s];
    }
    if (threadIdx.x < NOUT_PAD) {
        const int o = threadIdx.x;
        float acc = 0.0f;
        for (int s = 0; s < EW_ROWS; s++)
            acc += red[o][s];
        bpart[(int64_t)blockIdx.x * NOUT_PAD + o] = acc;
    }
    if (threadIdx.x == 0) {
        float acc = 0.0f;
        for (int s = 0; s < EW_ROWS; s++)
            acc += lred[s];
        lpart[blockIdx.x] = acc;
    }
}

// Hidden layer backward epilogue: dZ = dA * silu'(Z) as bf16 [B][64] and
// [64][B]; bias-gradient partials per block.
__global__ __launch_bounds__(256) void k_silu_bwd(int B, const float* __restrict__ dA,
                                                  const float* __restrict__ Z,
                                                  bf16* __restrict__ dZ, bf16* __restrict__ dZT,
                                                  float* __restrict__ bpart)
{
    __shared__ float tile[EW_ROWS][H + 1];
    const int s0 = blockIdx.x * EW_ROWS;
    for (int e = threadIdx.x; e < EW_ROWS * H; e += blockDim.x) {
        const int s = e / H, o = e % H;
        const int64_t idx = (int64_t)(s0 + s) * H + o;
        const float g     = dA[idx] * dsilu(Z[idx]);
        tile[s][o]        = g;
        dZ[idx]           = (bf16)g;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < EW_ROWS * H; e += blockDim.x) {
        const int o = e / EW_ROWS, s = e % EW_ROWS;
        dZT[(int64_t)o * B + s0 + s] = (bf16)tile[s][o];
    }
    if (threadIdx.x < H) {
        const int o = threadIdx.x;
        float acc = 0.0f;
        for (int s = 0; s < EW_ROWS; s++)
            acc += tile[s][o];
        bpart[(int64_t)blockIdx.x * H + o] = acc;
    }
}

// Reduces the gradient of every parameter (weight partials over split-K
// chunks, bias partials over blocks, fixed order) into G. Weight partials
// of layer l: [nsplit][out][in] at wpart + nsplit * OFF_W[l]; bias
// partials: [nblk][out] at bpart + nblk * (l - 1) * H.
__global__ __launch_bounds__(256) void k_reduce_grads(const float* __restrict__ wpart,
                                                      const float* __restrict__ bpart,
                                                      int nsplit, int nblk,
                                                      float* __restrict__ G)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM)
        return;
    int l = 1;
    while (l < 5 && p >= OFF_W[l + 1])
        l++;
    const int nw = LAYER_OUT[l] * LAYER_IN[l];
    if (p >= OFF_B[l])
        return;  // biases: k_reduce_bias
    float g = 0.0f;
    const int e = p - OFF_W[l];
    const float* slab = wpart + (int64_t)nsplit * OFF_W[l];
    for (int k = 0; k < nsplit; k++)
        g += slab[(int64_t)k * nw + e];
    G[p] = g;
}

// Bias gradients: one wave per bias, each lane summing every 64th
// per-block partial, then a fixed-order shuffle reduction (deterministic).
__global__ __launch_bounds__(64) void k_reduce_bias(const float* __restrict__ bpart, int nblk,
                                                    float* __restrict__ G)
{
    int l = 1, o = blockIdx.x;
    while (l < 5 && o >= LAYER_OUT[l]) {
        o -= LAYER_OUT[l];
        l++;
    }
    const int nout    = LAYER_OUT[l];
    const float* slab = bpart + (int64_t)nblk * (l - 1) * H;
    float g = 0.0f;
    for (int k = threadIdx.x; k < nblk; k += 64)
        g += slab[(int64_t)k * nout + o];
    for (int off = 32; off > 0; off >>= 1)
        g += __shfl_down(g, off);
    if (threadIdx.x == 0)
        G[OFF_B[l] + o] = g;
}

// Adam on FP32 master weights; refreshes the bf16 W and W^T copies.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, float* __restrict__ Mm,
                                              float* __restrict__ Vv,
                                              const float* __restrict__ G, float lr, float b1,
                                              float b2, float eps, float bc1, float bc2,
                                              bf16* __restrict__ Wb, bf16* __restrict__ WTb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM)
        return;
    const float g = G[p];
    const float m = b1 * Mm[p] + (1.0f - b1) * g;
    const float v = b2 * Vv[p] + (1.0f - b2) * g * g;
    Mm[p]         = m;
    Vv[p]         = v;
    const float w = P[p] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    P[p]          = w;
    int l = 1;
    while (l < 5 && p >= OFF_W[l + 1])
        l++;
    if (p < OFF_B[l]) {
        const int e = p - OFF_W[l];
        const int o = e / LAYER_IN[l], i = e % LAYER_IN[l];
        Wb[OFF_W[l] + e]                        = (bf16)w;
        WTb[OFF_W[l] + i * LAYER_OUT[l] + o] = (bf16)w;
    }
}

__global__ void k_refresh_bf16(const float* __restrict__ P, bf16* __restrict__ Wb,
                               bf16* __restrict__ WTb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM)
        return;
    int l = 1;
    while (l < 5 && p >= OFF_W[l + 1])
        l++;
    if (p < OFF_B[l]) {
        const int e = p - OFF_W[l];
        const int o = e / LAYER_IN[l], i = e % LAYER_IN[l];
        Wb[OFF_W[l] + e]                     = (bf16)P[p];
        WTb[OFF_W[l] + i * LAYER_OUT[l] + o] = (bf16)P[p];
    }
}

}  // namespace surrogate
}  // namespace oslmi

namespace oslmi {
namespace surrogate {

// ---- inference ---------------------------------------------------------
// Packed fragments: 30 per lane (W1 4 tiles; W2..W4 4 tiles x 2 K steps;
// W5 1 tile x 2 K steps), 8 bf16 each, stored [fragment][lane]; biases as
// accumulator initialisers: 17 float4 per lane, stored [slot][lane].
constexpr int NFRAG = 30, NBIAS = 17;
constexpr int NCELL = 64;

struct InferConsts {
    Norm nm;
    Domain dom;
};

// Hidden feature order along K of the packed W2..W5 fragments: K step s,
// lane group g, slot j holds feature 16 (2s + j/4) + 4g + j%4, which is
// where the previous layer's accumulators left it.
__host__ __device__ inline int hidden_k(int s, int g, int j)
{
    return 16 * (2 * s + j / 4) + 4 * g + (j % 4);
}

// SiLU for inference with the hardware exp2 and reciprocal (v_exp_f32,
// v_rcp_f32, about 1 ulp each), on pre-activations scaled by -log2(e):
// the packed weights and biases of layers 1-4 carry that factor and W5 its
// inverse, so with y = -z log2(e) the unit computes
//   y * rcp(1 + 2^y) = -log2(e) * z / (1 + e^-z) = -log2(e) * SiLU(z)
// (three VALU instructions and two transcendentals, no scaling multiply),
// and every next layer sees its input scaled by the same -log2(e), which
// W2..W4 (scaled by -log2(e) for their own output and by its inverse for
// their input) and W5 absorb. The result is rounded to bf16 next, far
// coarser than these errors.
constexpr float kLog2e = 1.44269504088896341f;
__device__ __forceinline__ float silu_fast(float y)
{
    return y * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y));
}

__device__ __forceinline__ bf16x8 act_pair(const f32x4& lo, const f32x4& hi)
{
    bf16x8 r;
    r[0] = (bf16)silu_fast(lo[0]);
    r[1] = (bf16)silu_fast(lo[1]);
    r[2] = (bf16)silu_fast(lo[2]);
    r[3] = (bf16)silu_fast(lo[3]);
    r[4] = (bf16)silu_fast(hi[0]);
    r[5] = (bf16)silu_fast(hi[1]);
    r[6] = (bf16)silu_fast(hi[2]);
    r[7] = (bf16)silu_fast(hi[3]);
    return r;
}

// Routing cell of a query: coat roughness (8 bins over [0.05, 1]) x
// cosine between L and V's mirror direction (8 bins over [-1, 1]).
__device__ __forceinline__ int route_cell(float lx, float lz, float vx, float vz, float cr)
{
    const float d = -lx * vx + lz * vz;
    int a = (int)((cr - 0.05f) * (8.0f / 0.95f));
    int b = (int)((d + 1.0f) * 4.0f);
    a = a < 0 ? 0 : (a > 7 ? 7 : a);
    b = b < 0 ? 0 : (b > 7 ? 7 : b);
    return a * 8 + b;
}

// Class of a query: 0 exact zero (V.z > 0 and L.z <= 0), 1 network
// (inside the training domain and, in mode 1, not in a routed cell), 2
// fallback (mode 1: the exact closure) or unanswered (mode 0: NaN).
__device__ __forceinline__ int route_class(const Domain& dom, unsigned long long cell_mask, int mode,
                                           const float v[NIN])
{
    const float lz = v[2], vz = v[4];
    bool inside = vz > 0.0f;
#pragma unroll
    for (int k = 0; k < 9; k++)
        inside = inside && v[5 + k] >= dom.lo[k] && v[5 + k] <= dom.hi[k];
    const int cell = route_cell(v[0], lz, v[3], vz, v[6]);
    if (inside && mode == 1 && ((cell_mask >> cell) & 1ull))
        inside = false;
    if (vz > 0.0f && lz <= 0.0f)
        return 0;
    return inside ? 1 : 2;
}

// Per-wave queue of network queries in LDS: a ring of QCAP slots of 16
// floats (14 raw inputs, the query index as bits, padding).
constexpr int QCAP = 96, QSLOT = 16, WAVES_PER_BLOCK = 4;
static_assert(QSLOT == 16, "slot reads index with k & (QSLOT - 1)");
// Per-wave fallback list in LDS (flushed to the global list when full).
constexpr int FBCAP = 64;

// The network over TILES 16-query tiles whose inputs sit in queue slots
// head + 16 u + r (u < TILES, r = lane % 16) of this wave's ring; `count`
// of them are valid. Writes the outputs (or appends to the fallback list).
template<int TILES>
__device__ __forceinline__ void run_tiles(const float* __restrict__ ring, int head, int count,
                                          const float noff[8], const float nscl[8],
                                          const bf16x8 w1[4], const bf16x8 w2[4][2],
                                          const bf16x8 w3[4][2], const bf16x8 w4[4][2],
                                          const bf16x8 w5[2], const f32x4* __restrict__ s_b,
                                          int64_t n, float* __restrict__ out, bool spot_check,
                                          float min_weight, bool spot[TILES], int qi_out[TILES])
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    bf16x8 x[TILES];
    int qi[TILES];
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        const int e     = u * 16 + r;
        const float* sl = ring + ((head + e) % QCAP) * QSLOT;
        // the whole slot is read unconditionally (always inside the ring)
        // and masked afterwards: conditional reads made the compiler wait
        // for every outstanding global load, the next chunk's included
        const int qv = __float_as_int(sl[14]);
        qi[u]        = e < count ? qv : -1;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            // layer-1 operand: this lane group's 8 inputs (normalised)
            const int k   = 8 * g + j;
            const float v = sl[k & (QSLOT - 1)];  // inside the slot; k >= NIN masked
            x[u][j] = (bf16)(k < NIN && e < count ? (v - noff[j]) * nscl[j] : 0.0f);
        }
    }
    f32x4 a[TILES][4];
    bf16x8 h0[TILES], h1[TILES];
    f32x4 y2[TILES];
#pragma unroll
    for (int u = 0; u < TILES; u++)
#pragma unroll
        for (int t = 0; t < 4; t++)
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1[t], x[u], s_b[t * 64 + lane], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        h0[u] = act_pair(a[u][0], a[u][1]);
        h1[u] = act_pair(a[u][2], a[u][3]);
    }
#pragma unroll
    for (int u = 0; u < TILES; u++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2[t][0], h0[u], s_b[(4 + t) * 64 + lane], 0, 0, 0);
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2[t][1], h1[u], a[u][t], 0, 0, 0);
        }
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        h0[u] = act_pair(a[u][0], a[u][1]);
        h1[u] = act_pair(a[u][2], a[u][3]);
    }
#pragma unroll
    for (int u = 0; u < TILES; u++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3[t][0], h0[u], s_b[(8 + t) * 64 + lane], 0, 0, 0);
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3[t][1], h1[u], a[u][t], 0, 0, 0);
        }
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        h0[u] = act_pair(a[u][0], a[u][1]);
        h1[u] = act_pair(a[u][2], a[u][3]);
    }
#pragma unroll
    for (int u = 0; u < TILES; u++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4[t][0], h0[u], s_b[(12 + t) * 64 + lane], 0, 0, 0);
            a[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4[t][1], h1[u], a[u][t], 0, 0, 0);
        }
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        h0[u] = act_pair(a[u][0], a[u][1]);
        h1[u] = act_pair(a[u][2], a[u][3]);
    }
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        y2[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w5[0], h0[u], s_b[16 * 64 + lane], 0, 0, 0);
        y2[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w5[1], h1[u], y2[u], 0, 0, 0);
    }
    // lanes of group 0 hold outputs 0..3 of their query. Spot check (when
    // spot_check): a predicted weight below the threshold has a large
    // relative error (the log-space fit is absolute); such a query is not
    // written but flagged in spot[u] / qi_out[u] for the exact closure.
#pragma unroll
    for (int u = 0; u < TILES; u++) {
        float v[4];
#pragma unroll
        for (int o = 0; o < 4; o++)
            v[o] = __builtin_amdgcn_exp2f(y2[u][o]);
        const bool mine = g == 0 && qi[u] >= 0;
        spot[u]   = spot_check && mine && fminf(fminf(v[0], v[1]), v[2]) < min_weight;
        qi_out[u] = qi[u];
        if (mine && !spot[u]) {
#pragma unroll
            for (int o = 0; o < 4; o++)
                out[o * n + qi[u]] = v[o];
        }
    }
}

// Routing and network in one pass over the queries. Each wave walks chunks
// of 64 consecutive queries (lane = query): it reads the 14 input planes
// once (coalesced), answers the exact zeros at once, appends fallback
// queries to the fallback list (one atomic per wave and chunk, wave ballot
// for the positions), and queues network queries in its LDS ring (ballot
// compaction, query order kept). Whenever 32 are queued, the network runs
// on two 16-query tiles; the rest is flushed after the last chunk. The next
// chunk's inputs are loaded before the network runs on the current queue.
// The weights stay in the accumulation registers for the whole kernel (the
// MFMA reads its A operand there), the bias initialisers in LDS.
template<int TILES>
__global__ __launch_bounds__(256) void k_infer_fused(InferConsts kc, unsigned long long cell_mask,
                                                     int mode, const bf16x8* __restrict__ frag,
                                                     const f32x4* __restrict__ biasp,
                                                     const float* __restrict__ q, int64_t n,
                                                     float* __restrict__ out, float min_weight,
                                                     int* __restrict__ fbidx, int* __restrict__ fbcount)
{
    __shared__ f32x4 s_b[NBIAS * 64];
    __shared__ float s_ring[WAVES_PER_BLOCK][QCAP * QSLOT];
    __shared__ int s_fb[WAVES_PER_BLOCK][FBCAP];
    for (int j = threadIdx.x; j < NBIAS * 64; j += blockDim.x)
        s_b[j] = biasp[j];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g    = lane >> 4;
    float* ring    = s_ring[wv];
    int* fbl       = s_fb[wv];
    const unsigned long long lt = (1ull << lane) - 1ull;
    // this wave's fallback list: fbn (wave-uniform) indices in fbl, moved
    // to the global list with one atomic when full and at the end
    int fbn         = 0;
    auto fb_flush   = [&]() {
        if (fbn == 0)
            return;
        int base = 0;
        if (lane == 0)
            base = atomicAdd(fbcount, fbn);
        base = __shfl(base, 0);
        if (lane < fbn)
            fbidx[base + lane] = fbl[lane];
        fbn = 0;
        __builtin_amdgcn_wave_barrier();
    };
    auto fb_append = [&](bool flag, int idx) {
        const unsigned long long m = __ballot(flag);
        if (m == 0)
            return;
        const int cnt = __popcll(m);
        if (fbn + cnt > FBCAP)
            fb_flush();
        if (flag)
            fbl[fbn + __popcll(m & lt)] = idx;
        fbn += cnt;
        __builtin_amdgcn_wave_barrier();
    };
    bf16x8 w1[4], w2[4][2], w3[4][2], w4[4][2], w5[2];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        w1[t] = frag[t * 64 + lane];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            w2[t][s] = frag[(4 + t * 2 + s) * 64 + lane];
            w3[t][s] = frag[(12 + t * 2 + s) * 64 + lane];
            w4[t][s] = frag[(20 + t * 2 + s) * 64 + lane];
        }
    }
    w5[0] = frag[28 * 64 + lane];
    w5[1] = frag[29 * 64 + lane];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        asm volatile("" : "+a"(w1[t]));
#pragma unroll
        for (int s = 0; s < 2; s++) {
            asm volatile("" : "+a"(w2[t][s]));
            asm volatile("" : "+a"(w3[t][s]));
            asm volatile("" : "+a"(w4[t][s]));
        }
    }
    asm volatile("" : "+a"(w5[0]));
    asm volatile("" : "+a"(w5[1]));
    // this lane group's normalisation constants (inputs 8g .. 8g+7)
    float noff[8], nscl[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k = 8 * g + j;
        noff[j]     = k < NIN ? kc.nm.in_off[k < NIN ? k : 0] : 0.0f;
        nscl[j]     = k < NIN ? kc.nm.in_scale[k < NIN ? k : 0] : 0.0f;
    }
    const int64_t nchunks = (n + 63) / 64;
    const int64_t wave    = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wv;
    const int64_t nwaves  = (int64_t)gridDim.x * WAVES_PER_BLOCK;
    // unconditional loads (clamped in-bounds address, value unused past n)
    auto load = [&](int64_t chunk, float v[NIN]) {
        const int64_t i  = chunk * 64 + lane;
        const int64_t ic = i < n ? i : n - 1;
#pragma unroll
        for (int k = 0; k < NIN; k++)
            v[k] = q[k * n + ic];
    };
    int head = 0, pending = 0;
    float cur[NIN];
    if (wave < nchunks)
        load(wave, cur);
    for (int64_t chunk = wave; chunk < nchunks; chunk += nwaves) {
        const int64_t i = chunk * 64 + lane;
        const bool live = i < n;
        int c = live ? route_class(kc.dom, cell_mask, mode, cur) : -1;
        if (c == 0) {
            out[i] = out[n + i] = out[2 * n + i] = out[3 * n + i] = 0.0f;
        } else if (c == 2 && mode == 0) {
            out[i] = out[n + i] = out[2 * n + i] = out[3 * n + i] = __builtin_nanf("");
        }
        if (mode == 1)
            fb_append(c == 2, (int)i);
        const unsigned long long nb = __ballot(c == 1);
        if (c == 1) {
            float* sl = ring + ((head + pending + __popcll(nb & lt)) % QCAP) * QSLOT;
#pragma unroll
            for (int k = 0; k < NIN; k++)
                sl[k] = cur[k];
            sl[14] = __int_as_float((int)i);
        }
        pending += __popcll(nb);
        __builtin_amdgcn_wave_barrier();
        // the next chunk's inputs are in flight while the network runs
        if (chunk + nwaves < nchunks)
            load(chunk + nwaves, cur);
        while (pending >= 16 * TILES) {
            bool spot[TILES];
            int qs[TILES];
            run_tiles<TILES>(ring, head, 16 * TILES, noff, nscl, w1, w2, w3, w4, w5, s_b, n, out,
                             mode == 1, min_weight, spot, qs);
            head = (head + 16 * TILES) % QCAP;
            pending -= 16 * TILES;
            __builtin_amdgcn_wave_barrier();
            if (mode == 1)
#pragma unroll
                for (int u = 0; u < TILES; u++)
                    fb_append(spot[u], qs[u]);
        }
    }
    if (pending > 0) {
        bool spot[TILES];
        int qs[TILES];
        run_tiles<TILES>(ring, head, pending, noff, nscl, w1, w2, w3, w4, w5, s_b, n, out,
                         mode == 1, min_weight, spot, qs);
        if (mode == 1)
#pragma unroll
            for (int u = 0; u < TILES; u++)
                fb_append(spot[u], qs[u]);
    }
    if (mode == 1)
        fb_flush();
}

__global__ void k_zero_int(int* p) { *p = 0; }

// ---- host context ----------------------------------------------------------
struct Context {
    int64_t max_batch = 0;
    Norm nm {};
    Domain dom {};
    float* P = nullptr;  // fp32 master parameters
    float* Mm = nullptr;
    float* Vv = nullptr;
    float* G = nullptr;
    bf16* Wb = nullptr;   // bf16 W (layer layouts as P)
    bf16* WTb = nullptr;  // bf16 W^T
    // activations / workspaces for max_batch samples
    bf16 *X = nullptr, *XT = nullptr;
    float* Y = nullptr;
    float* Z[6] = {};
    bf16 *A[5] = {}, *AT[5] = {};
    bf16 *dZ = nullptr, *dZT = nullptr;
    float* dA = nullptr;
    float* wpart = nullptr;
    float* bpart = nullptr;
    float* lpart = nullptr;
    // inference
    bf16x8* frag = nullptr;
    f32x4* biasp = nullptr;
    unsigned long long cell_mask = 0;  // routed cells (bit c: cell c)
    float min_weight = 0.0f;  // spot-check threshold on predicted weights
    bool packed = false;
    std::vector<float> host_params;
};

int ctx_alloc(Context* c)
{
    const int64_t B = c->max_batch;
    const int64_t nsplit = (B + KCHUNK - 1) / KCHUNK;
    const int64_t nblk   = B / EW_ROWS;
    struct Item {
        void** p;
        size_t bytes;
    } items[] = {
        { (void**)&c->P, NPARAM * 4 }, { (void**)&c->Mm, NPARAM * 4 }, { (void**)&c->Vv, NPARAM * 4 },
        { (void**)&c->G, NPARAM * 4 }, { (void**)&c->Wb, NPARAM * 2 }, { (void**)&c->WTb, NPARAM * 2 },
        { (void**)&c->X, (size_t)B * NIN_PAD * 2 }, { (void**)&c->XT, (size_t)B * NIN_PAD * 2 },
        { (void**)&c->Y, (size_t)B * NOUT * 4 },
        { (void**)&c->Z[1], (size_t)B * H * 4 }, { (void**)&c->Z[2], (size_t)B * H * 4 },
        { (void**)&c->Z[3], (size_t)B * H * 4 }, { (void**)&c->Z[4], (size_t)B * H * 4 },
        { (void**)&c->Z[5], (size_t)B * NOUT_PAD * 4 },
        { (void**)&c->A[1], (size_t)B * H * 2 }, { (void**)&c->A[2], (size_t)B * H * 2 },
        { (void**)&c->A[3], (size_t)B * H * 2 }, { (void**)&c->A[4], (size_t)B * H * 2 },
        { (void**)&c->AT[1], (size_t)B * H * 2 }, { (void**)&c->AT[2], (size_t)B * H * 2 },
        { (void**)&c->AT[3], (size_t)B * H * 2 }, { (void**)&c->AT[4], (size_t)B * H * 2 },
        { (void**)&c->dZ, (size_t)B * H * 2 }, { (void**)&c->dZT, (size_t)B * H * 2 },
        { (void**)&c->dA, (size_t)B * H * 4 },
        { (void**)&c->wpart, (size_t)nsplit * NPARAM * 4 },
        { (void**)&c->bpart, (size_t)nblk * 5 * H * 4 },
        { (void**)&c->lpart, (size_t)nblk * 4 },
        { (void**)&c->frag, (size_t)NFRAG * 64 * 16 }, { (void**)&c->biasp, (size_t)NBIAS * 64 * 16 },
    };
    for (auto& it : items) {
        if (int rc = check_hip(hipMalloc(it.p, it.bytes), "hipMalloc(surrogate)"))
            return rc;
    }
    if (int rc = check_hip(hipMemset(c->Mm, 0, NPARAM * 4), "hipMemset"))
        return rc;
    if (int rc = check_hip(hipMemset(c->Vv, 0, NPARAM * 4), "hipMemset"))
        return rc;
    return OSLMI_OK;
}

void ctx_free(Context* c)
{
    void* ptrs[] = { c->P, c->Mm, c->Vv, c->G, c->Wb, c->WTb, c->X, c->XT, c->Y, c->Z[1], c->Z[2],
                     c->Z[3], c->Z[4], c->Z[5], c->A[1], c->A[2], c->A[3], c->A[4], c->AT[1],
                     c->AT[2], c->AT[3], c->AT[4], c->dZ, c->dZT, c->dA, c->wpart, c->bpart,
                     c->lpart, c->frag, c->biasp };
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
}

static void gemm(hipStream_t st, int M, int N, int K, const bf16* A, int lda, const bf16* B,
                 int ldb, float* C, int ldc, int kchunk, int64_t cstride)
{
    const int splits = (K + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(k_gemm_nt, dim3(N / 16, M / 16, splits), dim3(64), 0, st, K, A, lda, B, ldb,
                       C, ldc, kchunk, cstride);
}

// Forward + backward over a batch of B samples (B a multiple of 64, at most
// max_batch); leaves the reduced gradients in G and per-block loss partials.
int fwd_bwd(Context* c, int64_t B64, const float* q, const float* t, int64_t ldq, hipStream_t st)
{
    const int B    = (int)B64;
    const int nblk = B / EW_ROWS;
    const int nsplit = (B + KCHUNK - 1) / KCHUNK;
    hipLaunchKernelGGL(k_prep, dim3((B + 255) / 256), dim3(256), 0, st, B, c->nm, q, t, ldq, c->X,
                       c->XT, c->Y);
    // forward
    const bf16* in = c->X;
    int kin = NIN_PAD;
    for (int l = 1; l <= 4; l++) {
        gemm(st, B, H, kin, in, kin, c->Wb + OFF_W[l], kin, c->Z[l], H, kin, 0);
        hipLaunchKernelGGL(k_bias_silu, dim3(nblk), dim3(256), 0, st, B, c->Z[l], c->P + OFF_B[l],
                           c->A[l], c->AT[l]);
        in  = c->A[l];
        kin = H;
    }
    gemm(st, B, NOUT_PAD, H, c->A[4], H, c->Wb + OFF_W[5], H, c->Z[5], NOUT_PAD, H, 0);
    hipLaunchKernelGGL(k_loss, dim3(nblk), dim3(256), 0, st, B, c->Z[5], c->P + OFF_B[5], c->Y,
                       c->dZ, c->dZT, c->bpart + (int64_t)nblk * 4 * H, c->lpart);
    // backward
    for (int l = 5; l >= 1; l--) {
        const int nout = LAYER_OUT[l], nin = LAYER_IN[l];
        const bf16* inT = l == 1 ? c->XT : c->AT[l - 1];
        gemm(st, nout, nin, B, c->dZT, B, inT, B, c->wpart + (int64_t)nsplit * OFF_W[l], nin,
             KCHUNK, (int64_t)nout * nin);
        if (l > 1) {
            gemm(st, B, H, nout, c->dZ, nout, c->WTb + OFF_W[l], nout, c->dA, H, nout, 0);
            hipLaunchKernelGGL(k_silu_bwd, dim3(nblk), dim3(256), 0, st, B, c->dA, c->Z[l - 1],
                               c->dZ, c->dZT, c->bpart + (int64_t)nblk * (l - 2) * H);
        }
    }
    hipLaunchKernelGGL(k_reduce_grads, dim3((NPARAM + 255) / 256), dim3(256), 0, st, c->wpart,
                       c->bpart, nsplit, nblk, c->G);
    hipLaunchKernelGGL(k_reduce_bias, dim3(4 * H + NOUT_PAD), dim3(64), 0, st, c->bpart, nblk,
                       c->G);
    return check_launch("surrogate fwd/bwd");
}

// Packs the inference fragments and bias initialisers from host parameters.
// Scale of W5's output row o in the packed weights: the network's output
// y_o (normalised log target) becomes log2 of the answer,
//   log2(exp(y_o sd_o + mu_o)) = (y_o sd_o + mu_o) log2(e),
// and the layer's input arrives scaled by -log2(e) (see silu_fast), so the
// row carries -sd_o and the bias (b_o sd_o + mu_o) log2(e); the kernel's
// epilogue is then a single exp2.
float out_row_scale(const Norm& nm, int o) { return o < NOUT ? -nm.out_sd[o] : 0.0f; }

int pack_inference(Context* c)
{
    const std::vector<float>& P = c->host_params;
    std::vector<bf16> frag((size_t)NFRAG * 64 * 8);
    std::vector<float> bias((size_t)NBIAS * 64 * 4);
    auto W = [&](int l, int o, int i) { return P[OFF_W[l] + o * LAYER_IN[l] + i]; };
    int f = 0;
    for (int t = 0; t < 4; t++, f++)
        for (int lane = 0; lane < 64; lane++)
            for (int j = 0; j < 8; j++)
                frag[((size_t)f * 64 + lane) * 8 + j] =
                    (bf16)(-kLog2e * W(1, 16 * t + (lane & 15), 8 * (lane >> 4) + j));
    for (int l = 2; l <= 4; l++)
        for (int t = 0; t < 4; t++)
            for (int s = 0; s < 2; s++, f++)
                for (int lane = 0; lane < 64; lane++)
                    for (int j = 0; j < 8; j++)
                        frag[((size_t)f * 64 + lane) * 8 + j] =
                            (bf16)W(l, 16 * t + (lane & 15), hidden_k(s, lane >> 4, j));
    for (int s = 0; s < 2; s++, f++)
        for (int lane = 0; lane < 64; lane++)
            for (int j = 0; j < 8; j++)
                frag[((size_t)f * 64 + lane) * 8 + j] =
                    (bf16)(W(5, lane & 15, hidden_k(s, lane >> 4, j)) * out_row_scale(c->nm, lane & 15));
    int b = 0;
    for (int l = 1; l <= 4; l++)
        for (int t = 0; t < 4; t++, b++)
            for (int lane = 0; lane < 64; lane++)
                for (int j = 0; j < 4; j++)
                    bias[((size_t)b * 64 + lane) * 4 + j] =
                        -kLog2e * P[OFF_B[l] + 16 * t + 4 * (lane >> 4) + j];
    for (int lane = 0; lane < 64; lane++)
        for (int j = 0; j < 4; j++) {
            const int o = 4 * (lane >> 4) + j;
            float v     = P[OFF_B[5] + o];
            if (o < NOUT)
                v = (v * c->nm.out_sd[o] + c->nm.out_mu[o]) * kLog2e;
            bias[((size_t)b * 64 + lane) * 4 + j] = v;
        }
    if (int rc = check_hip(hipMemcpy(c->frag, frag.data(), frag.size() * sizeof(bf16),
                                     hipMemcpyHostToDevice), "hipMemcpy(fragments)"))
        return rc;
    if (int rc = check_hip(hipMemcpy(c->biasp, bias.data(), bias.size() * sizeof(float),
                                     hipMemcpyHostToDevice), "hipMemcpy(biases)"))
        return rc;
    c->packed = true;
    return OSLMI_OK;
}

int sync_host_params(Context* c)
{
    c->host_params.resize(NPARAM);
    if (int rc = check_hip(hipMemcpy(c->host_params.data(), c->P, NPARAM * 4,
                                     hipMemcpyDeviceToHost), "hipMemcpy(params)"))
        return rc;
    return pack_inference(c);
}

}  // namespace surrogate

// exact fallback (oslmi_layered.hip)
int layered_eval_local_indexed(const oslmi_layered_material* m, int64_t n, const float* d_query,
                               const int* d_index, const int* d_count, float* d_out,
                               hipStream_t stream);

}  // namespace oslmi

using namespace oslmi;
using namespace oslmi::surrogate;

struct oslmi_surrogate {
    Context c;
};

extern "C" oslmi_surrogate* oslmi_surrogate_create(int64_t max_batch)
{
    if (max_batch < 64 || max_batch % KCHUNK != 0) {
        set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_create: max_batch must be a positive "
                                     "multiple of 2048");
        return nullptr;
    }
    oslmi_surrogate* s = new oslmi_surrogate();
    s->c.max_batch     = max_batch;
    for (int k = 0; k < NIN; k++)
        s->c.nm.in_scale[k] = 1.0f;
    for (int o = 0; o < NOUT; o++)
        s->c.nm.out_sd[o] = 1.0f;
    if (ctx_alloc(&s->c) != OSLMI_OK) {
        ctx_free(&s->c);
        delete s;
        return nullptr;
    }
    return s;
}

extern "C" void oslmi_surrogate_destroy(oslmi_surrogate* s)
{
    if (!s)
        return;
    (void)hipDeviceSynchronize();
    ctx_free(&s->c);
    delete s;
}

extern "C" int oslmi_surrogate_set_normalization(oslmi_surrogate* s, const float* in_offset,
                                                 const float* in_scale, const float* out_mean,
                                                 const float* out_std, const float* domain_lo,
                                                 const float* domain_hi)
{
    if (!s || !in_offset || !in_scale || !out_mean || !out_std || !domain_lo || !domain_hi)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_set_normalization: null argument");
    memcpy(s->c.nm.in_off, in_offset, sizeof(float) * NIN);
    memcpy(s->c.nm.in_scale, in_scale, sizeof(float) * NIN);
    memcpy(s->c.nm.out_mu, out_mean, sizeof(float) * NOUT);
    memcpy(s->c.nm.out_sd, out_std, sizeof(float) * NOUT);
    memcpy(s->c.dom.lo, domain_lo, sizeof(float) * 9);
    memcpy(s->c.dom.hi, domain_hi, sizeof(float) * 9);
    s->c.packed = false;  // the packed output layer carries the output scaling
    return OSLMI_OK;
}

extern "C" int oslmi_surrogate_param_count(void) { return NPARAM; }

extern "C" int oslmi_surrogate_set_params(oslmi_surrogate* s, const float* params, int64_t count)
{
    if (!s || !params || count != NPARAM)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_set_params: need 16672 floats");
    if (int rc = check_hip(hipMemcpy(s->c.P, params, NPARAM * 4, hipMemcpyHostToDevice),
                           "hipMemcpy(params)"))
        return rc;
    if (int rc = check_hip(hipMemset(s->c.Mm, 0, NPARAM * 4), "hipMemset"))
        return rc;
    if (int rc = check_hip(hipMemset(s->c.Vv, 0, NPARAM * 4), "hipMemset"))
        return rc;
    hipLaunchKernelGGL(k_refresh_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, s->c.P,
                       s->c.Wb, s->c.WTb);
    if (int rc = check_launch("oslmi_surrogate_set_params"))
        return rc;
    s->c.host_params.assign(params, params + NPARAM);
    return pack_inference(&s->c);
}

extern "C" int oslmi_surrogate_get_params(oslmi_surrogate* s, float* params, int64_t count)
{
    if (!s || !params || count != NPARAM)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_get_params: need 16672 floats");
    if (int rc = sync_host_params(&s->c))
        return rc;
    memcpy(params, s->c.host_params.data(), NPARAM * 4);
    return OSLMI_OK;
}

static int batch_ok(oslmi_surrogate* s, int64_t batch, const float* q, const char* who)
{
    if (!s || !q)
        return set_error(OSLMI_ERR_INVALID, std::string(who) + ": null argument");
    if (batch < 64 || batch % 64 != 0 || batch > s->c.max_batch)
        return set_error(OSLMI_ERR_INVALID, std::string(who) +
                                                ": batch must be a multiple of 64 within "
                                                "max_batch");
    return OSLMI_OK;
}

extern "C" int oslmi_surrogate_grads(oslmi_surrogate* s, int64_t batch, const float* d_query,
                                     const float* d_target, float* grads, float* loss)
{
    if (int rc = batch_ok(s, batch, d_query, "oslmi_surrogate_grads"))
        return rc;
    if (!d_target || !grads)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_grads: null argument");
    if (int rc = fwd_bwd(&s->c, batch, d_query, d_target, batch, 0))
        return rc;
    if (int rc = check_hip(hipMemcpy(grads, s->c.G, NPARAM * 4, hipMemcpyDeviceToHost),
                           "hipMemcpy(grads)"))
        return rc;
    if (loss) {
        std::vector<float> lp(batch / EW_ROWS);
        if (int rc = check_hip(hipMemcpy(lp.data(), s->c.lpart, lp.size() * 4,
                                         hipMemcpyDeviceToHost), "hipMemcpy(loss)"))
            return rc;
        double acc = 0.0;
        for (float v : lp)
            acc += v;
        *loss = (float)(acc / (4.0 * (double)batch));
    }
    return OSLMI_OK;
}

extern "C" int oslmi_surrogate_train_step(oslmi_surrogate* s, int64_t batch, const float* d_query,
                                          const float* d_target, float lr, float beta1,
                                          float beta2, float eps, int step, void* stream)
{
    if (int rc = batch_ok(s, batch, d_query, "oslmi_surrogate_train_step"))
        return rc;
    if (!d_target || step < 1)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_train_step: bad argument");
    hipStream_t st = (hipStream_t)stream;
    if (int rc = fwd_bwd(&s->c, batch, d_query, d_target, batch, st))
        return rc;
    const float bc1 = 1.0f - powf(beta1, (float)step);
    const float bc2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, s->c.P, s->c.Mm,
                       s->c.Vv, s->c.G, lr, beta1, beta2, eps, bc1, bc2, s->c.Wb, s->c.WTb);
    s->c.packed = false;
    return check_launch("oslmi_surrogate_train_step");
}

extern "C" int oslmi_surrogate_set_routing(oslmi_surrogate* s, const unsigned char* cells,
                                           int count)
{
    if (!s || !cells || count != NCELL)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_set_routing: need 64 cells");
    unsigned long long m = 0;
    for (int k = 0; k < NCELL; k++)
        if (cells[k])
            m |= 1ull << k;
    s->c.cell_mask = m;
    return OSLMI_OK;
}

extern "C" int oslmi_surrogate_set_spot_check(oslmi_surrogate* s, float min_weight)
{
    if (!s || !(min_weight >= 0.0f))
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_set_spot_check: need min_weight >= 0");
    s->c.min_weight = min_weight;
    return OSLMI_OK;
}

// Persistent grid of the fused kernel: the workgroups that fit at once (two
// waves per SIMD: 228 of the 512 registers per lane, so 2 workgroups of 4
// waves per CU), fewer when there are fewer 64-query chunks than waves.
static unsigned infer_grid(int64_t n)
{
    const int64_t chunks = (n + 63) / 64;
    int64_t blocks       = (chunks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    const int64_t cap    = 256 * 2;
    if (blocks > cap)
        blocks = cap;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

extern "C" int oslmi_surrogate_infer(oslmi_surrogate* s, const oslmi_layered_material* m,
                                     int64_t n, const float* d_query, float* d_out,
                                     int* d_fallback_index, int* d_fallback_count, void* stream)
{
    if (!s || (n > 0 && (!d_query || !d_out)))
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_infer: null argument");
    if (n < 0 || n > 0x7fffffff)
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_infer: bad n");
    if (n == 0)
        return OSLMI_OK;
    if (!s->c.packed)
        if (int rc = sync_host_params(&s->c))
            return rc;
    hipStream_t st = (hipStream_t)stream;
    InferConsts kc { s->c.nm, s->c.dom };
    const int mode = m ? 1 : 0;
    if (mode == 1 && (!d_fallback_index || !d_fallback_count))
        return set_error(OSLMI_ERR_INVALID, "oslmi_surrogate_infer: fallback buffers needed");
    if (mode == 1)
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, d_fallback_count);
    hipLaunchKernelGGL(k_infer_fused<2>, dim3(infer_grid(n)), dim3(256), 0, st, kc, s->c.cell_mask, mode,
                       s->c.frag, s->c.biasp, d_query, n, d_out, s->c.min_weight,
                       mode == 1 ? d_fallback_index : nullptr,
                       mode == 1 ? d_fallback_count : nullptr);
    if (int rc = check_launch("oslmi_surrogate_infer"))
        return rc;
    if (mode == 1)
        return layered_eval_local_indexed(m, n, d_query, d_fallback_index, d_fallback_count,
                                          d_out, st);
    return OSLMI_OK;
}
