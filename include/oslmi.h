/*
 * oslmi.h -- C ABI of the MI355X batched shading executor.
 *
 * This is the drop-in boundary below the reference's batched execution API.
 * Every pointer argument named d_* is a device (HBM) pointer; arrays are
 * structure-of-arrays planes of n elements ("SoA"), the GPU form of the
 * reference's Block<T,W> / Wide<T,W> layout (src/include/OSL/wide.h:439).
 * A Dual2 operand of k components occupies 3k planes in the order
 * [val_0..val_{k-1}, dx_0..dx_{k-1}, dy_0..dy_{k-1}].
 *
 * Calls return OSLMI_OK (0) or a negative error code; oslmi_last_error()
 * gives a message for the calling thread. stream may be NULL (the null
 * stream) or a hipStream_t. No call synchronizes unless it says so.
 */
#ifndef OSLMI_H
#define OSLMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSLMI_OK 0
#define OSLMI_ERR_INVALID -1     /* bad argument */
#define OSLMI_ERR_HIP -2         /* a HIP runtime call failed */
#define OSLMI_ERR_UNSUPPORTED -3 /* request outside what the library implements */
#define OSLMI_ERR_COMPILE -4     /* group code generation / compilation failed */

/* Library version and a message describing the last failure on this thread. */
const char* oslmi_version(void);
const char* oslmi_last_error(void);

/* Select and probe the current device; returns OSLMI_ERR_HIP if there is
 * no usable gfx950 GPU. */
int oslmi_init(void);

/* ---- Wide shadeops ---------------------------------------------------
 * Batched noise over n points. Replaces the per-width shadeop entry points
 * osl_b16_<ISA>_{noise,snoise,cellnoise,hashnoise,simplexnoise,...}_<args>
 * (src/liboslexec/wide/wide_opnoise_impl_xmacro.h:57-127, names from
 * define_opname_macros.h:17-35) and the generic noise signature
 * NoiseGenericFunc(outdim, out, derivs, indim, in, period, params)
 * (src/include/OSL/oslnoise.h:108-111).
 * name: "perlin"|"snoise", "uperlin"|"noise", "cell"|"cellnoise",
 *       "hash"|"hashnoise", "simplex", "usimplex" (batched_llvm_gen.cpp:6096-6134).
 * outdim 1 or 3; indim 1..4; derivs selects Dual2 in/out planes
 * (cell/hash noise output zero derivatives, as the reference forces). */
int oslmi_noise(const char* name, int outdim, int indim, int derivs, const float* d_in,
                float* d_out, int64_t n, void* stream);

/* hash() shadeop over 1..4 float planes (src/liboslexec/wide/wide_ophash.cpp:28-80). */
int oslmi_hash(int indim, const float* d_in, int32_t* d_out, int64_t n, void* stream);

/* ---- Prebuilt group lowerings -----------------------------------------
 * The two benchmark groups compiled ahead of time (OSL sources in
 * openshadinglanguage_amd/shaders/). Each replaces one
 * BatchedExecutor<W>::execute over the whole point set
 * (src/include/OSL/oslexec.h:1000-1004) plus, for layered3, the renderer's
 * closure evaluation (src/testrender/shading.h:364-372).
 *
 * fbm_cell: d_P 3 planes in; d_Cout 3 planes out. */
int oslmi_shade_fbm_cell(int64_t n, const float* d_P, float* d_Cout, void* stream);

/* layered3: d_P, d_N, d_wo, d_wi 3 planes each; dPdx/dPdy host float[3]
 * (uniform screen derivatives of P); d_out 4 planes (R, G, B, pdf);
 * d_aov optional (NULL or 20 planes: fbm v/dx/dy, vn v xyz/dx xyz/dy xyz,
 * gabor v/dx/dy, Cd xyz, alpha, 0). */
int oslmi_shade_layered3(int64_t n, const float* d_P, const float* d_N, const float* d_wo,
                         const float* d_wi, const float* dPdx, const float* dPdy, float* d_out,
                         float* d_aov, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OSLMI_H */
