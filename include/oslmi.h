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

/* Periodic noise (pnoise/psnoise with or without a name): the same planes
 * plus d_period, indim planes of n floats holding each point's period per
 * input dimension (floats without derivatives, as in the reference).
 * Replaces the osl_b16_<ISA>_{pnoise,psnoise,pcellnoise,phashnoise}_<args>
 * shadeops (wide_opnoise_periodic_impl_xmacro.h; PeriodicNoise,
 * PeriodicSNoise, PeriodicAdaptionOf at src/include/OSL/oslnoise.h:486-548,
 * :2531-2760). name: "perlin"|"snoise"|"psnoise", "uperlin"|"noise"|"pnoise",
 * "cell"|"cellnoise", "hash"|"hashnoise" (the periodic branch of
 * batched_llvm_gen.cpp:6096-6134; simplex has no periodic form). */
int oslmi_pnoise(const char* name, int outdim, int indim, int derivs, const float* d_in,
                 const float* d_period, float* d_out, int64_t n, void* stream);

/* hash() shadeop over 1..4 float planes (src/liboslexec/wide/wide_ophash.cpp:28-80). */
int oslmi_hash(int indim, const float* d_in, int32_t* d_out, int64_t n, void* stream);

/* ---- Shading system, shader groups, batched execution -----------------
 * The drop-in path: shader masters are loaded from OSO text, assembled into
 * groups of layers with parameter values and connections, lowered by the
 * library's emitter to one fused gfx950 kernel (hiprtc) and executed over
 * any number of points. */
typedef struct oslmi_shadingsys oslmi_shadingsys;
typedef struct oslmi_group oslmi_group;

/* Shader globals of a batch (BatchedShaderGlobals<W>, src/include/OSL/
 * batched_shaderglobals.h:21-140) as SoA planes over n points. Index order:
 * P, dPdx, dPdy, N, Ng, I, u, v, dudx, dudy, dvdx, dvdy, time, dPdu, dPdv, Ps.
 * planes[k] is a device pointer (3 planes for triples, 1 for floats) or NULL,
 * in which case uniform[k] holds the value shared by all points (uniform
 * shader globals, batched_shaderglobals.h:21-75). */
#define OSLMI_NUM_GLOBALS 16
typedef struct oslmi_batched_globals {
    int64_t n;
    const float* planes[OSLMI_NUM_GLOBALS];
    float uniform[OSLMI_NUM_GLOBALS][3];
} oslmi_batched_globals;

/* ShadingSystem construction (src/include/OSL/oslexec.h:170). include_dir:
 * directory of the device library headers (the package's csrc/) handed to
 * the JIT; cache_dir: code-object cache directory or NULL. */
oslmi_shadingsys* oslmi_shadingsys_create(const char* include_dir, const char* cache_dir);
void oslmi_shadingsys_destroy(oslmi_shadingsys* ss);
/* ShadingSystem::attribute (oslexec.h:177-324): "colorspace", "jit_cache",
 * "include_dir", "commonspace" (the synonym of "common", default "world"),
 * "closure_sample" (0/1: surface kernels compiled afterwards also carry
 * CompositeBSDF::sample, for oslmi_execute_batch_sample),
 * "closure_export" (0/1: they also export Ci's lobes, for
 * oslmi_execute_batch_closures), and
 * "gpu_waves_per_eu" (0..8: the occupancy, in waves per SIMD, the emitted
 * kernels are register-allocated for; default 6, 0 leaves it to the
 * compiler). */
int oslmi_shadingsys_attribute(oslmi_shadingsys* ss, const char* name, const char* value);
/* Named coordinate systems, the renderer services behind transform(),
 * matrix(space) and getmatrix(): RendererServices::get_matrix /
 * get_inverse_matrix for a space name (batched_rendererservices.h:62-108,
 * testshade's SimpleRenderer::name_transform, src/testshade/simplerend.cpp:
 * 464-468), and the uniform shader2common / object2common matrices of a
 * batch under the names "shader" and "object" (testshade.cpp:973-979).
 * m: the space-to-common matrix, 16 floats in Imath's row-vector layout
 * (translation in m[12..14]); inv: its inverse, or NULL to have the library
 * invert m as OSL's default get_inverse_matrix does. "common" and the
 * commonspace synonym ("world") are the identity and cannot be named.
 * Takes effect at the next oslmi_execute_batch, oslmi_batch_bind,
 * oslmi_batch_launch or oslmi_batch_capture (also for a batch bound or a
 * graph captured before the change), which then waits for the device once
 * to update the table. */
int oslmi_renderer_name_transform(oslmi_shadingsys* ss, const char* name, const float* m,
                                  const float* inv);
/* Renderer attributes behind getattribute(name, value)
 * (BatchedRendererServices::get_attribute / get_attribute_uniform,
 * batched_rendererservices.h:128-150): type "int", "float", "color"/"point"/"vector"/"normal" or
 * "matrix"; value points at 1, 1, 3 or 16 values of that type. A kernel's
 * getattribute succeeds (returns 1 and writes its value) only when the
 * attribute exists with the value's base type; otherwise it returns 0 and
 * leaves the value unchanged. Takes effect like
 * oslmi_renderer_name_transform. */
int oslmi_renderer_set_attribute(oslmi_shadingsys* ss, const char* name, const char* type,
                                 const void* value);
/* LoadMemoryCompiledShader: register a shader master from OSO text. */
int oslmi_load_memory_compiled_shader(oslmi_shadingsys* ss, const char* shadername,
                                      const char* oso_text);
/* ShaderGroupBegin / Parameter / Shader / ConnectShaders / ShaderGroupEnd
 * (oslexec.h:634-742). Parameter values apply to the next oslmi_shader
 * call. type: "int", "float", "color", "point", "vector", "normal",
 * "matrix", "string" (value: const char* const*). */
oslmi_group* oslmi_shader_group_begin(oslmi_shadingsys* ss, const char* groupname);
int oslmi_parameter(oslmi_shadingsys* ss, oslmi_group* g, const char* name, const char* type,
                    const void* value, int count);
int oslmi_shader(oslmi_shadingsys* ss, oslmi_group* g, const char* usage, const char* shadername,
                 const char* layername);
int oslmi_connect_shaders(oslmi_shadingsys* ss, oslmi_group* g, const char* srclayer,
                          const char* srcparam, const char* dstlayer, const char* dstparam);
int oslmi_shader_group_end(oslmi_shadingsys* ss, oslmi_group* g);
void oslmi_group_release(oslmi_group* g);
/* Renderer outputs ("renderer_outputs"): symbols written to the output
 * planes in order ("layer.symbol" or "symbol" = last layer that has it);
 * a triple takes 3 planes, an int/float 1. */
int oslmi_group_outputs(oslmi_shadingsys* ss, oslmi_group* g, int n, const char* const* names);
/* jit_group (oslexec.h:993-998): emit + compile + load. */
int oslmi_jit_group(oslmi_shadingsys* ss, oslmi_group* g);
/* Emit and compile to a gfx950 code object without loading it (no device
 * needed); fills the code-object cache so a later jit_group only loads. */
int oslmi_compile_group(oslmi_shadingsys* ss, oslmi_group* g);
/* Emit only (no device needed); copies the HIP source into buf and returns
 * its length (debug_output of the reference). */
int oslmi_group_emit(oslmi_shadingsys* ss, oslmi_group* g, char* buf, int64_t bufsize);
int oslmi_group_output_planes(oslmi_shadingsys* ss, oslmi_group* g);
double oslmi_group_jit_seconds(oslmi_group* g);
/* BatchedExecutor<W>::execute over bg->n points (oslexec.h:1000-1004).
 * d_outputs: output_planes planes of n floats. If d_eval is non-NULL the
 * surface closure Ci is evaluated at (d_wo, d_wi) (3 planes each) like the
 * renderer's CompositeBSDF::eval and written as 4 planes (R, G, B, pdf). */
int oslmi_execute_batch(oslmi_shadingsys* ss, oslmi_group* g, const oslmi_batched_globals* bg,
                        float* d_outputs, const float* d_wo, const float* d_wi, float* d_eval,
                        void* stream);
/* The same execution with the surface closure Ci sampled instead of
 * evaluated, as the renderer's CompositeBSDF::sample(wo, rx, ry, rz)
 * (src/testrender/shading.h:376-402, called from simpleraytracer.cpp's
 * path loop): d_wo 3 planes, d_rnd 3 planes (rx, ry, rz in [0,1)),
 * d_samples 7 planes (wi xyz, weight rgb, pdf; all zero when no lobe
 * was drawn). The group must have been compiled after setting the
 * shading-system attribute "closure_sample" to 1. */
int oslmi_execute_batch_sample(oslmi_shadingsys* ss, oslmi_group* g,
                               const oslmi_batched_globals* bg, float* d_outputs,
                               const float* d_wo, const float* d_rnd, float* d_samples,
                               void* stream);
/* The same execution with the surface closure Ci handed out as testrender
 * gets it from OSL: the list of components process_bsdf_closure passes to
 * CompositeBSDF::add_bsdf (src/testrender/shading.cpp:1448-1667), each as
 * (closure id, weight rgb, parameters) with testrender's ids and parameter
 * order (register_closures, shading.cpp:185-296; ClosureIDs, shading.h:
 * 25-61; oslexec.h:1041 register_closure). Layers are resolved into the
 * weights, so MX_LAYER_ID does not appear. d_closures: 1 + max_lobes *
 * OSLMI_CLOSURE_RECORD planes of n floats: plane 0 the count (at most
 * max_lobes, 1..8); lobe k's record at planes 1 + k * OSLMI_CLOSURE_RECORD:
 * id, weight r, g, b, then OSLMI_CLOSURE_PARAMS parameter floats
 * (DIFFUSE_ID 3: N; MICROFACET_ID 8: distribution (0 ggx, 1 beckmann), N,
 * U, xalpha, yalpha, eta, refract; MX_DIELECTRIC_ID 17: N, U, reflection
 * tint, transmission tint, roughness_x, roughness_y, ior;
 * MX_GENERALIZED_SCHLICK_ID 19: N, U, reflection tint, transmission tint,
 * roughness_x, roughness_y, f0, f90, exponent; MX_SHEEN_ID 23: N, albedo,
 * roughness, mode), zero-padded. d_wo (NULL: -I) is the direction
 * MaterialX lobes are built at (it does not change what is exported). The
 * group must have been compiled after setting the shading-system
 * attribute "closure_export" to 1. */
#define OSLMI_CLOSURE_PARAMS 21
#define OSLMI_CLOSURE_RECORD (4 + OSLMI_CLOSURE_PARAMS)
int oslmi_execute_batch_closures(oslmi_shadingsys* ss, oslmi_group* g,
                                 const oslmi_batched_globals* bg, float* d_outputs,
                                 const float* d_wo, int max_lobes, float* d_closures,
                                 void* stream);

/* ---- The reference's own batch layout ----------------------------------
 * BatchedShaderGlobals<16> byte for byte (src/include/OSL/
 * batched_shaderglobals.h:21-140): the 64-byte UniformShaderGlobals block,
 * then VaryingShaderGlobals<16>, whose Block<Vec3,16> fields are x[16],
 * y[16], z[16] (src/include/OSL/wide.h:439-445) and Block<float,16> /
 * Block<int,16> / Block<pointer,16> 16 values each. 4032 bytes, 64-byte
 * aligned. A renderer that fills these for the batched CPU path hands the
 * same blocks (copied to device memory) to oslmi_execute_wide16. The
 * uniform pointers and object2common / shader2common / Ci are opaque
 * renderer data here and are not read. */
#if defined(__cplusplus)
#define OSLMI_ALIGN64 alignas(64)
#else
#define OSLMI_ALIGN64 _Alignas(64)
#endif
typedef struct oslmi_bsg16 {
    /* UniformShaderGlobals (batched_shaderglobals.h:21-75) */
    OSLMI_ALIGN64 void* renderstate;
    void* tracedata;
    void* objdata;
    void* context;
    void* renderer;
    int raytype;
    int pad[5];
    /* VaryingShaderGlobals<16> (batched_shaderglobals.h:77-140) */
    OSLMI_ALIGN64 float P[3][16];
    float dPdx[3][16], dPdy[3][16], dPdz[3][16];
    float I[3][16], dIdx[3][16], dIdy[3][16];
    float N[3][16];
    float Ng[3][16];
    float u[16], dudx[16], dudy[16];
    float v[16], dvdx[16], dvdy[16];
    float dPdu[3][16], dPdv[3][16];
    float time[16];
    float dtime[16];
    float dPdtime[3][16];
    float Ps[3][16], dPsdx[3][16], dPsdy[3][16];
    const void* object2common[16];
    const void* shader2common[16];
    void* Ci[16];
    float surfacearea[16];
    int flipHandedness[16];
    int backfacing[16];
} oslmi_bsg16;

/* Where a symbol's value goes on the renderer side (SymLocationDesc,
 * src/include/OSL/oslexec.h:69-104; ShadingSystem::add_symlocs, :1072-1076).
 * Only the output arena is supported: point i's value is written to
 * output_base + offset + shadeindex[i] * stride. type: "float", "int",
 * "color", "point", "vector" or "normal"; derivs must be 0. name:
 * "layer.symbol" or "symbol" (the last layer that has it). */
#define OSLMI_ARENA_HEAP 0
#define OSLMI_ARENA_OUTPUTS 1
#define OSLMI_ARENA_USERDATA 2
typedef struct oslmi_symloc {
    const char* name;
    const char* type;
    int arena;
    int derivs;
    int64_t offset;
    int64_t stride;
} oslmi_symloc;
/* add_symlocs(group, symlocs) / clear_symlocs(group). Symbols placed in
 * the output arena become the group's renderer outputs: before the group
 * is compiled they set its output list; afterwards each must already be one
 * of its outputs. At most 16 per group. */
int oslmi_add_symlocs(oslmi_shadingsys* ss, oslmi_group* g, int n, const oslmi_symloc* locs);
int oslmi_clear_symlocs(oslmi_shadingsys* ss, oslmi_group* g);

/* BatchedExecutor<16>::execute (src/include/OSL/oslexec.h:1000-1004) for
 * nbatch batches at once: d_bsg nbatch BatchedShaderGlobals<16> blocks in
 * device memory; batch b holds 16 points except the last, which holds
 * last_batch_size (1..16, the reference's batch_size); d_shadeindex the
 * batches' Wide<const int,16> shade indices (nbatch * 16 ints, device);
 * outputs placed by the group's symlocs relative to d_output_base (device).
 * Allocates a per-group device workspace on first use (the globals and
 * outputs in structure-of-arrays form, reused by later calls). */
int oslmi_execute_wide16(oslmi_shadingsys* ss, oslmi_group* g, int64_t nbatch,
                         int last_batch_size, const oslmi_bsg16* d_bsg,
                         const int32_t* d_shadeindex, void* d_output_base, void* stream);

/* ---- Multi-group batches -----------------------------------------------
 * Many groups shaded in one persistent launch. The reference executes one
 * group per BatchedExecutor<W>::execute call (oslexec.h:1000-1004); a
 * renderer with many materials loops over them (testshade.cpp:1886-1909
 * does one group). A batch replaces that loop with one kernel: group k's
 * points become tiles sorted by group, and every workgroup pulls tiles from
 * a device counter until the list is empty. Item k of a bind is group k and
 * takes the arguments oslmi_execute_batch takes. */
typedef struct oslmi_batch oslmi_batch;
typedef struct oslmi_batch_item {
    oslmi_batched_globals globals;
    float* d_outputs;
    const float* d_wo;
    const float* d_wi;
    float* d_eval;
} oslmi_batch_item;
/* Launch shape: block (64, 128 or 256 work-items), points per work-item
 * per tile (1..8; a tile is block * ppt points), workgroups per CU of the
 * persistent grid (1..32), blackbody table in LDS (1) or read through the
 * caches (0), tile order (0: groups in item order, 1: groups by measured
 * cost per tile, most expensive first; both keep each group's tiles
 * together), streams the batch's kernels are spread over (1..4; more than
 * one lets the kernels overlap, forked from and joined back into the
 * caller's stream; eager launches only: a captured graph records the
 * kernels on the capturing stream). */
typedef struct oslmi_batch_config {
    int block, ppt, blocks_per_cu, bb_lds, order, streams;
} oslmi_batch_config;
/* Emits and compiles (hiprtc, code-object cache) the batch's persistent
 * kernels over ngroups ended groups: groups_per_kernel groups per kernel
 * (0: 8), the kernels launched one after another. load = 0 only fills the
 * code-object cache (no device needed). */
oslmi_batch* oslmi_batch_create(oslmi_shadingsys* ss, int ngroups, oslmi_group* const* groups,
                                int groups_per_kernel, int load);
void oslmi_batch_destroy(oslmi_batch* b);
/* Longest compile time among the batch's kernels (compiled concurrently). */
double oslmi_batch_jit_seconds(oslmi_batch* b);
/* Number of persistent kernels a launch runs. */
int oslmi_batch_kernels(oslmi_batch* b);
/* Binds one item per group (nitems == ngroups) and uploads the argument
 * blocks and tile list. Drops a captured graph. */
int oslmi_batch_bind(oslmi_batch* b, int nitems, const oslmi_batch_item* items);
int oslmi_batch_set_config(oslmi_batch* b, const oslmi_batch_config* cfg);
int oslmi_batch_get_config(oslmi_batch* b, oslmi_batch_config* cfg);
/* One persistent launch over every bound item (the captured graph when
 * oslmi_batch_capture has run on this stream). */
int oslmi_batch_launch(oslmi_batch* b, void* stream);
/* The same kernel launched once per item, item after item (the reference's
 * one-execute-per-group loop), for comparison. */
int oslmi_batch_launch_per_group(oslmi_batch* b, void* stream);
/* Captures oslmi_batch_launch into a hipGraph for replay on `stream`. */
int oslmi_batch_capture(oslmi_batch* b, void* stream);
/* Times `reps` launches of candidate shapes (block x ppt x blocks_per_cu,
 * then bb_lds x order x streams) and keeps the fastest; writes it to *best
 * and its mean time in ms to *best_ms. With graph != 0 every candidate is
 * timed as a captured graph replayed on `stream` (a created stream), and
 * the winner stays captured for oslmi_batch_launch on that stream. Runs the
 * kernels, so the bound outputs are overwritten with the (identical)
 * results. */
int oslmi_batch_autotune(oslmi_batch* b, void* stream, int reps, int graph,
                         oslmi_batch_config* best, double* best_ms);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): oslmi_batch_bind + oslmi_batch_launch in one call that
 * allocates nothing. The argument blocks, tile list and tile counters go to
 * a fixed workspace of the library in each GPU's memory (1 MiB: at most 256
 * items and 32768 tiles of block * ppt points under the batch's current
 * launch shape; larger requests fail with OSLMI_ERR_INVALID) instead of
 * buffers the batch allocates. Needs oslmi_init on the device first (it
 * sets the workspace up), and the shading system's transform table already
 * in device memory if a group uses named spaces (one earlier execute or
 * bind). Runs on `stream` only (no side streams) and returns once its
 * argument copies have reached the device, which waits for the stream's
 * earlier work. Every call on a GPU shares that GPU's workspace: calls that
 * overlap on one GPU (on different streams or threads) race. The batch's
 * own bound items and captured graph are left untouched. */
int oslmi_batch_launch_noalloc(oslmi_batch* b, int nitems, const oslmi_batch_item* items,
                               void* stream);

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

/* ---- Layered MaterialX closure -----------------------------------------
 * The closure layer(sheen_bsdf(N, sheen_albedo, sheen_roughness),
 *   layer(dielectric_bsdf(N, 0, coat_refl_tint, coat_refr_tint, r, r, ior, "ggx"),
 *         generalized_schlick_bsdf(N, 0, base_refl_tint, base_refr_tint, rb, rb,
 *                                  F0, F90, base_exponent, "ggx")))
 * (stdosl.h:522, :566, :606, :707) turned into lobes and evaluated/sampled
 * like the renderer does for a camera-ray hit: process_bsdf_closure with
 * evaluate_layer_opacity, CompositeBSDF::prepare/eval/sample
 * (src/testrender/shading.cpp:1199-1272, :1448-1667, shading.h:319-393) over
 * libbsdl's MTX lobes (libbsdl/include/BSDL/MTX/bsdf_{dielectric,schlick,
 * sheen}_impl.h). Transmission tints must be zero (OSLMI_ERR_UNSUPPORTED
 * otherwise). */
typedef struct oslmi_layered_material {
    float coat_refl_tint[3];
    float coat_refr_tint[3];
    float base_refl_tint[3];
    float base_refr_tint[3];
    float base_exponent;
    float base_roughness; /* < 0: each point's coat roughness */
    float sheen_albedo[3];
} oslmi_layered_material;

/* Per point: d_N, d_V (towards the viewer, -I), d_L (eval direction) 3
 * planes each; d_params 9 planes (coat ior, coat roughness, sheen
 * roughness, F0 rgb, F90 rgb); d_rnd 8 planes ((rx, ry) of samples 0..3).
 * d_eval (NULL or 4 planes): CompositeBSDF::eval(V, L) as weight rgb, pdf.
 * d_samples (NULL or 28 planes): CompositeBSDF::sample(V, rx, ry, 0) for
 * k = 0..3 in planes 7k..7k+6 = wi xyz, weight rgb, pdf. */
int oslmi_layered_closure(const oslmi_layered_material* m, int64_t n, const float* d_N,
                          const float* d_V, const float* d_L, const float* d_params,
                          const float* d_rnd, float* d_eval, float* d_samples, void* stream);

/* CompositeBSDF::eval of the same material for queries in the shading
 * frame (N = +z, V = (x, 0, z)): d_query 14 planes = L xyz, V x, V z, coat
 * ior, coat roughness, sheen roughness, F0 rgb, F90 rgb; d_out 4 planes =
 * weight rgb, pdf. */
int oslmi_layered_eval_local(const oslmi_layered_material* m, int64_t n, const float* d_query,
                             float* d_out, void* stream);

/* Host copy of the energy-compensation tables (MiniMicrofacetGGX 16x16 then
 * DielectricReflFront 32x16x16 = 8448 floats), baked on first use like
 * libbsdl/src/genluts.cpp:128-192 and cached beside the library. No device
 * needed. */
int oslmi_bsdl_tables(float* out, int64_t count);

/* ---- Learned surrogate of the layered closure --------------------------
 * No counterpart in the reference: a 14 -> 4 x 64 SiLU -> 4 MLP that
 * approximates oslmi_layered_eval_local (inputs: its 14 query planes;
 * outputs: log eval weight rgb and log pdf), trained and evaluated with
 * BF16 MFMA / FP32 accumulation kernels. Parameters: 16672 floats, row-major
 * W1[64][32] (input columns 14..31 zero), b1[64], W2..W4[64][64], b2..b4,
 * W5[32][64] (output rows 4..31 zero), b5[32]. */
typedef struct oslmi_surrogate oslmi_surrogate;
/* max_batch: the largest training batch, a multiple of 2048. Allocates
 * device workspaces on the current device. */
oslmi_surrogate* oslmi_surrogate_create(int64_t max_batch);
void oslmi_surrogate_destroy(oslmi_surrogate* s);
int oslmi_surrogate_param_count(void);
/* Input normalisation (x - offset) * scale (14 each), output normalisation
 * (y - mean) / std (4 each), and the training domain of the 9 material
 * parameters (queries outside it are answered by the exact closure). */
int oslmi_surrogate_set_normalization(oslmi_surrogate* s, const float* in_offset,
                                      const float* in_scale, const float* out_mean,
                                      const float* out_std, const float* domain_lo,
                                      const float* domain_hi);
int oslmi_surrogate_set_params(oslmi_surrogate* s, const float* params, int64_t count);
int oslmi_surrogate_get_params(oslmi_surrogate* s, float* params, int64_t count);
/* Gradient of the mean squared error of the normalised outputs over one
 * batch (d_query 14 planes, d_target 4 planes = exact weight rgb, pdf; both
 * with leading dimension batch, batch a multiple of 64) into host grads
 * (16672 floats); *loss (optional) receives the loss. Synchronizes. */
int oslmi_surrogate_grads(oslmi_surrogate* s, int64_t batch, const float* d_query,
                          const float* d_target, float* grads, float* loss);
/* One Adam step (step counts from 1) on one batch, asynchronous on stream. */
int oslmi_surrogate_train_step(oslmi_surrogate* s, int64_t batch, const float* d_query,
                               const float* d_target, float lr, float beta1, float beta2,
                               float eps, int step, void* stream);
/* Routing table: 64 cells (8 coat-roughness bins x 8 bins of the cosine
 * between L and the mirror direction of V); a non-zero cell sends its
 * queries to the exact closure. */
int oslmi_surrogate_set_routing(oslmi_surrogate* s, const unsigned char* cells, int count);
/* Spot check (with fallback only): a query whose predicted weight is below
 * min_weight in any channel is answered by the exact closure instead (the
 * network fits log values, so its relative error grows as values vanish).
 * 0 disables it (the default). */
int oslmi_surrogate_set_spot_check(oslmi_surrogate* s, float min_weight);
/* Answers n queries (d_query 14 planes) into d_out (4 planes: weight rgb,
 * pdf). With m non-NULL, queries outside the training domain or in routed
 * cells are listed in d_fallback_index / *d_fallback_count (device, n
 * ints / 1 int) and evaluated exactly with material m; with m NULL the
 * network answers every in-domain query and the others get NaN. Queries
 * with V.z > 0 and L.z <= 0 get zeros (exact). */
int oslmi_surrogate_infer(oslmi_surrogate* s, const oslmi_layered_material* m, int64_t n,
                          const float* d_query, float* d_out, int* d_fallback_index,
                          int* d_fallback_count, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OSLMI_H */
