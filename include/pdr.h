/*
 * pdr.h — C ABI of libpdr, the MI355X (gfx950) implementation of
 * photara_drender's differentiable rasterize -> interpolate -> texture chain and
 * its atlas refiner.
 *
 * Memory: every array argument of one call is either host memory (the call then
 * stages through device memory and returns when the result is in the caller's
 * buffer, like the reference's synchronous Vulkan submits) or device memory
 * (the call enqueues its kernels on `stream` and returns; results are valid
 * once the stream reaches them).  Mixing host and device arrays in one call is
 * PDR_INVALID_ARGUMENT.  `stream` is a hipStream_t (NULL = the HIP default
 * stream, as everywhere in HIP); host-memory calls run on the context's own
 * stream.  Layouts are the reference's: row-major, float32 / uint32,
 * images [height, width, channels], lower-left origin.
 *
 * Errors: a non-zero pdr_status; pdr_last_error() returns the message of the
 * last failure on the calling thread.  Status codes map onto the reference's
 * exception classes (std::invalid_argument, std::out_of_range,
 * std::overflow_error, std::runtime_error) and messages are the reference's.
 *
 * Threading: calls on one context are serialised by a per-context mutex, as the
 * reference serialises dispatches (src/context_internal.hpp:101).
 */
#ifndef PDR_H
#define PDR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdr_context_s* pdr_context;

typedef enum pdr_status {
    PDR_OK = 0,
    PDR_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    PDR_OUT_OF_RANGE = 2,     /* std::out_of_range */
    PDR_OVERFLOW = 3,         /* std::overflow_error */
    PDR_RUNTIME = 4           /* std::runtime_error */
} pdr_status;

enum { PDR_CULL_NONE = 0, PDR_CULL_BACK = 1, PDR_CULL_FRONT = 2 };       /* types.hpp:11-15 */
enum { PDR_ADDRESS_CLAMP = 0, PDR_ADDRESS_WRAP = 1, PDR_ADDRESS_MIRROR = 2 }; /* types.hpp:17-21 */

/* Context flags (pdr_context_set_flags). */
enum {
    /* Check index / uv ranges of device-memory inputs with a GPU pass before the
     * work (costs one stream synchronisation).  Host inputs are always checked,
     * as the reference does.  Default: on. */
    PDR_FLAG_VALIDATE_DEVICE_INPUTS = 1u << 0,
    /* The attribute gradient of pdr_interpolate_backward(_batched) and the
     * texel gradient of pdr_texture_backward as ordered gathers instead of
     * atomics: per vertex / texel over its pixels in ascending order (taps in
     * order within a pixel), the reference's summation order.  Bit-reproducible;
     * slower (a radix sort of 3 or 4 entries per pixel, at most 2^30 pixels per
     * call).  The raster backward is deterministic with or without it.
     * Default: off. */
    PDR_FLAG_DETERMINISTIC = 1u << 1,
    /* Autotune the raster's launch parameters (wave / queue thresholds,
     * triangles per wave, triangle traversal order, tiled or triangle-parallel
     * forward): the first call of each workload class (view count,
     * log2 of the triangle and pixel counts) times every candidate with HIP
     * events and caches the fastest in the context; later calls use it.
     * Results do not depend on the choice.  Not applied while the stream is
     * being captured.  Default: off.  No counterpart in the reference. */
    PDR_FLAG_AUTOTUNE = 1u << 2
};

typedef struct pdr_device_info {
    char name[256];
    uint32_t vendor_id;
    uint32_t device_id;
    uint32_t api_version;       /* HIP runtime version */
    uint32_t supports_ray_query; /* always 0: ray queries stay on the host */
} pdr_device_info;

const char* pdr_last_error(void);

/* Context(ContextOptions) — include/photara_drender/context.hpp:14; src/context.cpp:248-382.
 * enable_validation is accepted for signature parity (types.hpp:23-26). */
pdr_status pdr_context_create(uint32_t device_index, int enable_validation, pdr_context* out);
void pdr_context_destroy(pdr_context ctx);
pdr_status pdr_context_set_flags(pdr_context ctx, uint32_t flags);
uint32_t pdr_context_flags(pdr_context ctx);
/* Context::device_info — context.hpp:22 */
pdr_status pdr_context_device_info(pdr_context ctx, pdr_device_info* info);
/* Context::enumerate_devices — context.hpp:23 */
pdr_status pdr_enumerate_devices(pdr_device_info* infos, uint32_t capacity, uint32_t* count);
/* Per-kernel timing: with timing on, every kernel the context launches is
 * bracketed by HIP events on its stream; pdr_context_kernel_time returns the
 * summed milliseconds and launch count of one kernel name since timing was
 * (re)enabled ("raster_fwd", "raster_bwd", "bin_count", "bin_fill",
 * "interp_fwd", "interp_bwd", "texture_fwd", "texture_bwd", "photometric",
 * "adam", ...).  No counterpart in the reference (which records only the
 * refiner's wall-clock phases, src/texture_refiner.cpp:420-435). */
pdr_status pdr_context_set_timing(pdr_context ctx, int enable);
pdr_status pdr_context_kernel_time(pdr_context ctx, const char* name, double* total_ms, uint64_t* launches);

/* The autotuner's cache as text, one class per line:
 * "<kind> <views> <log2 triangles> <log2 pixels> <wave_work> <fwd_waves> <flat_work> <bwd_waves>
 *  <tiled> <bin depth> <order>" (order: 0 = triangle index order, 1 = each view's nearer end of the
 * index range first; import also accepts lines without it, as order 0).
 * export: writes up to capacity bytes (NUL-terminated) and the needed size
 * (with the NUL) to *length; import: adds or replaces the listed classes;
 * clear: forgets every class. */
pdr_status pdr_autotune_export(pdr_context ctx, char* buffer, size_t capacity, size_t* length);
pdr_status pdr_autotune_import(pdr_context ctx, const char* text);
pdr_status pdr_autotune_clear(pdr_context ctx);

/* Blocks until `stream` (NULL = default stream) and the context's own stream are idle. */
pdr_status pdr_synchronize(pdr_context ctx, void* stream);

/* Rasterizer::forward — include/photara_drender/rasterizer.hpp:22-25,
 * src/rasterizer.cpp:373-442, shaders/rasterize_forward.hlsl:45-175.
 * pos [V,4] clip positions, idx [T,3]; viewport4 (x,y,w,h) or NULL for the full
 * image; raster_out [H,W,4] = (b0, b1, z/w, id+1); deriv_out [H,W,4] =
 * (db0/dx, db0/dy, db1/dx, db1/dy) or NULL; viewport_out4 (host) receives the
 * resolved viewport or may be NULL. */
pdr_status pdr_rasterize_forward(pdr_context ctx, const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                                 uint32_t width, uint32_t height, uint32_t cull_mode, int want_derivatives,
                                 const float* viewport4, float* raster_out, float* deriv_out, float* viewport_out4,
                                 void* stream);

/* pdr_rasterize_forward without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as pdr_rasterize_forward, device memory only.  Its scratch (8 B per
 * pixel of the size rounded up to multiples of 4, and to 64 pixels, + 16 B per
 * triangle, each part 256-byte aligned) comes from a fixed
 * 32 MiB array per GPU set up by pdr_context_create; larger calls fail with
 * PDR_OVERFLOW.  Indices are not validated by a readback (out-of-range
 * triangles are skipped, as by the shader) and the launch parameters are the
 * defaults.  All calls on one GPU share the array: calls that overlap on one
 * GPU (other streams or contexts) race. */
pdr_status pdr_rasterize_forward_noalloc(pdr_context ctx, const float* pos, uint32_t V, const uint32_t* idx,
                                         uint32_t T, uint32_t width, uint32_t height, uint32_t cull_mode,
                                         int want_derivatives, const float* viewport4, float* raster_out,
                                         float* deriv_out, float* viewport_out4, void* stream);

/* One view of a batched raster: its own clip positions, size and outputs. */
typedef struct pdr_raster_view {
    const float* positions; /* [V,4] */
    uint32_t width, height;
    float viewport[4];      /* (x,y,w,h); viewport[2] <= 0 selects the full image */
    float* raster;          /* [H,W,4] out (forward) or in (backward) */
    float* derivatives;     /* [H,W,4] out or NULL (forward only) */
    const float* grad_raster; /* [H,W,4] (backward only) */
    float* grad_positions;  /* [V,4] out (backward only) */
} pdr_raster_view;

/* Many views of uneven size in one binning pass and one raster launch
 * (project-specified batching of Rasterizer::forward; device memory only).
 * Each view's output equals pdr_rasterize_forward on that view. */
pdr_status pdr_rasterize_forward_batched(pdr_context ctx, const pdr_raster_view* views, uint32_t nviews, uint32_t V,
                                         const uint32_t* idx, uint32_t T, uint32_t cull_mode, int want_derivatives,
                                         void* stream);

/* Rasterizer::backward — rasterizer.hpp:27-32, src/rasterizer.cpp:444-537,
 * shaders/rasterize_backward.hlsl:36-158 + reduce_vertex_gradients.hlsl:12-29.
 * grad_pos_out [V,4] (z component is always 0, as in the reference). */
pdr_status pdr_rasterize_backward(pdr_context ctx, const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                                  const float* raster, uint32_t width, uint32_t height, const float* grad_raster,
                                  const float* viewport4, float* grad_pos_out, void* stream);

/* Batched backward (device memory only): view k writes views[k].grad_positions. */
pdr_status pdr_rasterize_backward_batched(pdr_context ctx, const pdr_raster_view* views, uint32_t nviews, uint32_t V,
                                          const uint32_t* idx, uint32_t T, void* stream);

/* Rasterizer::interpolate_forward — rasterizer.hpp:34-38, src/rasterizer.cpp:539-601,
 * shaders/interpolate_forward.hlsl:15-54.  attr [V,A], raster [H,W,4] -> out [H,W,A]. */
pdr_status pdr_interpolate_forward(pdr_context ctx, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                   uint32_t T, const float* raster, uint32_t width, uint32_t height, float* out,
                                   void* stream);

/* Rasterizer::interpolate_backward — rasterizer.hpp:40-45, src/rasterizer.cpp:603-683,
 * shaders/interpolate_grad_attributes.hlsl:15-58 + interpolate_grad_raster.hlsl:16-54.
 * grad [H,W,A] -> grad_attr [V,A], grad_raster [H,W,4] (zw = 0). */
pdr_status pdr_interpolate_backward(pdr_context ctx, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                    uint32_t T, const float* raster, uint32_t width, uint32_t height,
                                    const float* grad, float* grad_attr, float* grad_raster, void* stream);

/* One view of a batched interpolation (device memory only). */
typedef struct pdr_interp_view {
    uint32_t width, height;
    const float* raster;  /* [H,W,4] */
    float* out;           /* [H,W,A] forward output */
    const float* grad;    /* [H,W,A] backward input */
    float* grad_raster;   /* [H,W,4] backward output */
} pdr_interp_view;

/* Batched interpolation; the backward sums all views' attribute gradients into grad_attr. */
pdr_status pdr_interpolate_forward_batched(pdr_context ctx, const pdr_interp_view* views, uint32_t nviews,
                                           const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                           uint32_t T, void* stream);
pdr_status pdr_interpolate_backward_batched(pdr_context ctx, const pdr_interp_view* views, uint32_t nviews,
                                            const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                            uint32_t T, float* grad_attr, void* stream);

/* Rasterizer::texture_forward — rasterizer.hpp:47-51, src/rasterizer.cpp:685-735,
 * shaders/texture_forward.hlsl:16-79.  tex [th,tw,C], uv [H,W,2] -> out [H,W,C]. */
pdr_status pdr_texture_forward(pdr_context ctx, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                               uint32_t address_mode, const float* uv, const float* raster, uint32_t width,
                               uint32_t height, float* out, void* stream);

/* Rasterizer::texture_backward — rasterizer.hpp:53-59, src/rasterizer.cpp:737-829,
 * build_texture_sample_adjacency :187-258, shaders/texture_grad_texture.hlsl:15-38,
 * texture_grad_uv.hlsl:38-80.  grad_uv may be NULL when compute_uv_gradient is 0. */
pdr_status pdr_texture_backward(pdr_context ctx, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                                uint32_t address_mode, const float* uv, const float* raster, uint32_t width,
                                uint32_t height, const float* grad, int compute_uv_gradient, float* grad_tex,
                                float* grad_uv, void* stream);

/* ProjectionView subset used by TextureRefiner::refine — types.hpp:201-218. */
typedef struct pdr_refine_view {
    uint32_t width, height, channels; /* channels 3 or 4 */
    const float* image;               /* [H,W,channels] */
    const float* visibility_mask;     /* [H,W] or NULL */
    float world_to_clip[16];          /* row-major, clip = M [p,1] */
    int has_viewport;
    float viewport[4];
} pdr_refine_view;

/* TextureRefineOptions — types.hpp:267-283, plus the project's extended
 * objective (chart total variation and a joint seam term, the loss formulas of
 * python/photara_drender/atlas.py:84-144) which the reference refiner lacks. */
typedef struct pdr_refine_options {
    uint32_t width, height, steps, batch_size;
    float learning_rate, minimum_learning_rate, photometric_epsilon;
    float adam_beta1, adam_beta2, adam_epsilon, clamp_min, clamp_max;
    uint32_t seam_polish_steps;
    float seam_learning_rate, seam_epsilon;
    /* extensions; all zero reproduces TextureRefiner::refine exactly */
    float tv_weight;           /* chart TV weight during the photometric phase */
    float tv_epsilon;          /* default 1e-4 when 0 */
    const int32_t* chart_ids;  /* [height,width] chart per texel, <0 = invalid; NULL = one chart */
    float joint_seam_weight;   /* seam term added at every photometric step */
    uint32_t use_graph;        /* capture the step loop as a hipGraph (default behaviour of pdr_refine_options_default) */
    uint32_t checkpoint_every; /* >0: copy the atlas to checkpoints every k steps */
    float* checkpoints;        /* [steps/checkpoint_every, height, width, 3] host or device, or NULL */
    uint32_t stop_after;       /* >0: run only the first stop_after steps of the `steps`-step schedule (its
                                  learning rates and view order unchanged), then the seam polish; loss_out
                                  entries past it stay 0 (a run that stops at a target checkpoint) */
} pdr_refine_options;

void pdr_refine_options_default(pdr_refine_options* options);

/* TextureRefiner::refine — include/photara_drender/texture_refiner.hpp:21-28,
 * src/texture_refiner.cpp:174-454.  Synchronous.  positions [V,3], uv [V,2],
 * seam_pairs [npairs,2,2] or NULL; color_out [height,width,3];
 * loss_out [steps]; seam_loss_out [seam steps run] or NULL; times_out2 =
 * (precompute_seconds, optimization_seconds) or NULL. */
pdr_status pdr_refine(pdr_context ctx, const float* initial_texture, const float* positions, uint32_t V,
                      const float* uv, const uint32_t* idx, uint32_t T, const pdr_refine_view* views, uint32_t nviews,
                      const float* seam_pairs, uint32_t npairs, const pdr_refine_options* options, float* color_out,
                      float* loss_out, float* seam_loss_out, double* times_out2);

/* ------------------------------------------------------------------ bake */

/* ProjectionView — include/photara_drender/types.hpp:135-152. */
typedef struct pdr_bake_view {
    uint32_t width, height, channels; /* channels 3 or 4 */
    const float* image;               /* [H,W,channels] in [0,1] */
    const float* visibility_mask;     /* [H,W] or NULL */
    float world_to_clip[16];          /* row-major, clip = M [p,1] */
    float camera_position[3];
    int has_viewport;
    float viewport[4];
    float weight;          /* default 1 */
    float depth_bias;      /* default 1e-3 */
    float min_view_cosine; /* default 0.05 */
} pdr_bake_view;

/* TextureBakeOptions — types.hpp:154-179.  blend_mode: 0 best_view,
 * 1 weighted_average, 2 softmax.  visibility_mode: 0 shadow_map,
 * 1 hybrid_ray_query, 2 ray_query.  This backend has no ray queries: like the
 * reference on such a device, ray-query modes fall back to the shadow map when
 * allow_visibility_fallback is set and fail otherwise. */
typedef struct pdr_bake_options {
    uint32_t width, height;
    uint32_t blend_mode;
    uint32_t visibility_mode;
    uint32_t pcf_radius;
    float ray_origin_bias;
    int allow_visibility_fallback;
    float mask_floor;
    float softmax_scale; /* <= 0: scene-relative scale (pdr_suggest_softmax_scale) */
} pdr_bake_options;

void pdr_bake_options_default(pdr_bake_options* options);
void pdr_bake_view_default(pdr_bake_view* view);

/* suggest_softmax_scale — src/texture_padding.cpp:187-222.  Host arrays. */
float pdr_suggest_softmax_scale(const float* positions3, uint32_t vertex_count, const float* camera_positions3,
                                uint32_t camera_count);

/* TextureBaker::bake — src/texture_baker.cpp:156-361 + shaders/project_texture.hlsl.
 * Mesh arrays host or device; images and masks host or device per view.
 * Outputs (host or device, all [height,width]): color_out [.,.,4] (RGB, valid),
 * confidence_out, source_view_out (u32, 0xFFFFFFFF none), valid_out,
 * coverage_out, face_ids_out (u32, 0xFFFFFFFF outside the mesh; NULL to skip);
 * used_ray_query_out (always 0 here; NULL to skip). */
pdr_status pdr_bake(pdr_context ctx, const float* positions, const float* normals, const float* uv,
                    uint32_t vertex_count, const uint32_t* indices, uint32_t triangle_count,
                    const pdr_bake_view* views, uint32_t view_count, const pdr_bake_options* options,
                    float* color_out, float* confidence_out, uint32_t* source_view_out, float* valid_out,
                    float* coverage_out, uint32_t* face_ids_out, int* used_ray_query_out);

/* ------------------------------------------------------------- padding */

/* TexturePaddingOptions — include/photara_drender/texture_padding.hpp:33-44. */
typedef struct pdr_padding_options {
    uint32_t margin;                /* guard band in texels (default 4) */
    int fill_unobserved;            /* fill covered texels without samples (default 1) */
    uint32_t maximum_fill_distance; /* cap of that fill in texels, 0 = unlimited */
    float valid_threshold;          /* a texel is a colour source when valid >= threshold (default 0.5) */
} pdr_padding_options;

/* TexturePaddingStats — texture_padding.hpp:46-57. */
typedef struct pdr_padding_stats {
    uint64_t valid_texels, gutter_texels, unobserved_texels, remaining_unobserved, filled_texels;
    double seconds;
} pdr_padding_stats;

void pdr_padding_options_default(pdr_padding_options* options);

/* pad_texture_atlas — texture_padding.hpp:59-64, src/texture_padding.cpp:226-351
 * (8SSEDT nearest-source field, then the guard band and the unobserved fill).
 * color [height,width,channels] (3 or 4) is updated in place; valid_mask
 * [height,width]; coverage_mask [height,width] or NULL (no coverage: every
 * texel without a sample is a gutter texel); filled_mask [height,width] in/out
 * or NULL (no earlier fills, none recorded); stats_out may be NULL. */
pdr_status pdr_pad_texture_atlas(pdr_context ctx, float* color, uint32_t width, uint32_t height, uint32_t channels,
                                 const float* valid_mask, const float* coverage_mask, float* filled_mask,
                                 const pdr_padding_options* options, pdr_padding_stats* stats_out, void* stream);

/* ------------------------------------------------------- warm-start MLP */
/* The learned warm start has no counterpart in the reference: a per-texel MLP
 * 40 -> 64 -> 64 -> 3 (GELU, tanh form) that predicts the refined atlas colour
 * from the texel's projected samples, added to their confidence-weighted mean.
 * Parameters: PDR_MLP_PARAMS float32 values,
 *   W1 [40,64] | b1 [64] | W2 [64,64] | b2 [64] | W3 [64,3] | b3 [3].
 * Matrix products take bfloat16 operands (round to nearest even) and
 * accumulate in float32.  Arrays are host or device memory as above; features
 * must be 16-byte aligned. */
#define PDR_MLP_FEATURES 40
#define PDR_MLP_PARAMS 6979

/* Features of the warm start, computed like pdr_bake (same arguments,
 * validation and shadow-map visibility; blend_mode is ignored): for every
 * covered texel, the 8 samples of highest weighted-average confidence
 * (visibility * alpha * cos * weight, project_texture.hlsl:208-246), in
 * descending order (ties keep view order), each [R, G, B, confidence, cos];
 * unused slots and uncovered texels are zero.  features_out [height, width, 40]. */
pdr_status pdr_bake_features(pdr_context ctx, const float* positions, const float* normals, const float* uv,
                             uint32_t vertex_count, const uint32_t* indices, uint32_t triangle_count,
                             const pdr_bake_view* views, uint32_t view_count, const pdr_bake_options* options,
                             float* features_out);

/* out_rgb [count, 3]: network output plus the samples' confidence-weighted mean
 * colour; 0 for texels without samples (the bake's colour there). */
pdr_status pdr_mlp_infer(pdr_context ctx, const float* features, uint64_t count, const float* params, float* out_rgb,
                         void* stream);

/* Gradient of loss = sum_t valid_t |y_t - target_t|^2 / (3 max(sum_t valid_t, 1))
 * with respect to params: grad_out [PDR_MLP_PARAMS]; loss_count_out [2] =
 * (loss, sum of valid) or NULL; valid [count] (0/1) or NULL for all texels. */
pdr_status pdr_mlp_gradient(pdr_context ctx, const float* features, const float* target, const float* valid,
                            uint64_t count, const float* params, float* grad_out, float* loss_count_out,
                            void* stream);

/* The gradient above followed by one Adam step (step counts from 1) that
 * updates params, adam_m and adam_v in place. */
pdr_status pdr_mlp_adam_step(pdr_context ctx, const float* features, const float* target, const float* valid,
                             uint64_t count, float* params, float* adam_m, float* adam_v, uint32_t step,
                             float learning_rate, float beta1, float beta2, float epsilon, float* loss_count_out,
                             void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PDR_H */
