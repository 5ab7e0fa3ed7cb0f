/*
 * pdr_oracle — CPU restatement of photara_drender's differentiable raster /
 * interpolate / texture / atlas-refinement path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in photara_drender_amd/ and the "reference algorithm on CPU"
 * baseline timed by bench.py.  Nothing in the product links, loads or calls it.
 *
 * Every function follows the reference operation by operation (same FP32
 * operation order, no FMA contraction: the build uses -ffp-contract=off) and
 * cites the reference file:line it restates.  All pointers are host pointers.
 * Return codes: 0 ok, 1 invalid_argument, 2 out_of_range, 3 overflow,
 * 4 runtime; orc_last_error() holds the message.
 */
#ifndef PDR_ORACLE_H
#define PDR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* orc_last_error(void);
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* src/rasterizer.cpp:61-138 (build_tile_bins) — CSR over 16x16 tiles.
 * offsets must hold tiles+1 entries; pass list=NULL to only size (offsets[tiles]
 * is the entry count). */
int orc_build_tile_bins(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                        uint32_t W, uint32_t H, const float* viewport4,
                        uint32_t* offsets, uint32_t* list);

/* src/rasterizer.cpp:373-442 + shaders/rasterize_forward.hlsl:45-175 */
int orc_rasterize_forward(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                          uint32_t W, uint32_t H, uint32_t cull_mode, int want_deriv,
                          const float* viewport4 /* nullable */,
                          float* raster, float* deriv /* nullable */, float* viewport_out4);

/* src/rasterizer.cpp:444-537 + shaders/rasterize_backward.hlsl:36-158 +
 * shaders/reduce_vertex_gradients.hlsl:12-29 (per-triangle bbox loop, then
 * CSR vertex reduce in ascending corner order). */
int orc_rasterize_backward(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                           const float* raster, uint32_t W, uint32_t H,
                           const float* grad_raster, const float* viewport4 /* nullable */,
                           float* grad_pos);

/* src/rasterizer.cpp:539-601 + shaders/interpolate_forward.hlsl:15-54 */
int orc_interpolate_forward(const float* attr, uint32_t V, uint32_t A,
                            const uint32_t* idx, uint32_t T,
                            const float* raster, uint32_t P, float* out);

/* src/rasterizer.cpp:603-683 + shaders/interpolate_grad_attributes.hlsl:15-58 +
 * shaders/interpolate_grad_raster.hlsl:16-54.
 * reference_mode=1: the O(V*A*P) loop exactly as dispatched by the reference.
 * reference_mode=0: same sums in the same order, visiting only pixels whose
 * triangle contains the vertex (bit-identical for finite gradients). */
int orc_interpolate_backward(const float* attr, uint32_t V, uint32_t A,
                             const uint32_t* idx, uint32_t T,
                             const float* raster, uint32_t P, const float* grad,
                             float* grad_attr, float* grad_raster, int reference_mode);

/* shaders/interpolate_grad_attributes.hlsl:15-58 for outputs o = vertex * A +
 * attribute in [o_begin, o_end) only, written to out[o - o_begin]. */
void orc_interpolate_grad_attributes_range(uint32_t V, uint32_t A, const uint32_t* idx, uint32_t T,
                                           const float* raster, uint32_t P, const float* grad,
                                           uint64_t o_begin, uint64_t o_end, float* out);

/* src/rasterizer.cpp:685-735 + shaders/texture_forward.hlsl:16-79 */
int orc_texture_forward(const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                        uint32_t address_mode, const float* uv, const float* raster,
                        uint32_t P, float* out);

/* src/rasterizer.cpp:737-829 + build_texture_sample_adjacency :187-258 +
 * shaders/texture_grad_texture.hlsl:15-38 + shaders/texture_grad_uv.hlsl:38-80 */
int orc_texture_backward(const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                         uint32_t address_mode, const float* uv, const float* raster,
                         uint32_t P, const float* grad, int compute_uv,
                         float* grad_tex, float* grad_uv /* nullable */);

/* src/texture_refiner.cpp:36-50 */
void orc_transform_positions(const float* pos3, uint32_t V, const float* m16, float* out4);

/* shaders/atlas_photometric_gradient.hlsl:45-108 (sequential pixel order;
 * the loss is reduced per 64-pixel group with the shader's tree, groups summed
 * in order).  grad and *loss accumulate. */
void orc_photometric_gradient(const float* tex, uint32_t tw, uint32_t th,
                              const float* uv, const float* target, const float* mask,
                              uint32_t P, float eps, float norm, float* grad, float* loss);

/* shaders/atlas_seam_gradient.hlsl:50-100 */
void orc_seam_gradient(const float* tex, uint32_t tw, uint32_t th,
                       const float* pairs, uint32_t N, float eps, float norm,
                       float* grad, float* loss);

/* shaders/atlas_adam_update.hlsl:20-43 */
void orc_adam_update(uint32_t n, float lr, float b1, float b2, float eps, float bc1,
                     float bc2, float cmin, float cmax, float* tex, float* grad,
                     float* m, float* v);

/* src/texture_refiner.cpp:315-370,377-402: per-step scalars.  Arrays sized
 * steps (+ seam_steps for the seam arrays). */
void orc_refine_schedule(uint32_t steps, uint32_t batch, uint32_t nviews,
                         const double* view_valid_weight, float lr, float min_lr,
                         float b1, float b2, uint32_t seam_steps,
                         float* lr_out, float* bc1_out, float* bc2_out, float* norm_out,
                         uint32_t* view_idx_out /* steps*batch */,
                         float* seam_bc1_out, float* seam_bc2_out);

/* src/texture_refiner.cpp:206-254: target/mask/valid weight of one view from
 * its raster (covered = raster id != 0). */
void orc_refine_view_inputs(const float* image, uint32_t channels, const float* vis_mask /* nullable */,
                            const float* raster, uint32_t P, float* target3, float* mask,
                            double* valid_weight);

/* FP64 twins with frozen visibility (triangle ids taken from an FP32 raster),
 * for central finite-difference checks of the analytic backward. */
void orc_raster_fixed_f64(const double* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                          const float* raster_ids, uint32_t W, uint32_t H, const double* viewport4,
                          double* raster_out);
void orc_interpolate_forward_f64(const double* attr, uint32_t V, uint32_t A,
                                 const uint32_t* idx, uint32_t T, const double* raster,
                                 uint32_t P, double* out);
void orc_texture_forward_f64(const double* tex, uint32_t tw, uint32_t th, uint32_t C,
                             uint32_t address_mode, const double* uv, const double* raster,
                             uint32_t P, double* out);

/* TextureBaker::bake (src/texture_baker.cpp:156-361) with shadow-map
 * visibility (the reference's path on devices without ray queries) and the
 * projection of shaders/project_texture.hlsl:117-247; views are projected in
 * order, each texel's state carried from view to view.  viewport is the
 * resolved (x, y, w, h); weight is the effective per-view weight (already
 * multiplied by the softmax scale in softmax mode). */
typedef struct orc_bake_view {
    uint32_t width, height, channels;
    const float* image;  /* [h, w, channels] */
    const float* mask;   /* [h, w] (floored by the caller when a mask floor applies) or NULL */
    float world_to_clip[16];
    float camera[3];
    float viewport[4];
    float weight, depth_bias, min_view_cosine;
} orc_bake_view;

int orc_bake(const float* pos3, const float* normal3, const float* uv2, uint32_t V, const uint32_t* idx,
             uint32_t T, const orc_bake_view* views, uint32_t nviews, uint32_t width, uint32_t height,
             uint32_t blend_mode, uint32_t pcf_radius, float* color4, float* confidence, uint32_t* source_view,
             float* valid, float* coverage, uint32_t* face_ids);

/* Warm-start features (project-specified): per covered atlas texel the 8
 * samples of highest weighted-average confidence vis * alpha * cos * weight
 * (project_texture.hlsl:208), descending, ties in view order, each
 * [R, G, B, confidence, cos]; zero elsewhere.  features40 [height, width, 40]. */
int orc_bake_features(const float* pos3, const float* normal3, const float* uv2, uint32_t V, const uint32_t* idx,
                      uint32_t T, const orc_bake_view* views, uint32_t nviews, uint32_t width, uint32_t height,
                      uint32_t pcf_radius, float* features40);

/* pad_texture_atlas (src/texture_padding.cpp:226-351): Danielsson 8SSEDT
 * nearest-source field (:37-97; forward sweep relaxing left, up, up-left,
 * up-right; backward sweep right, down, down-right, down-left; a candidate
 * replaces the current offset only when strictly closer), then gutter texels
 * within `margin` and, with fill_unobserved, covered texels within
 * maximum_fill_distance (0 = unlimited) copy their source's colour.
 * color [h,w,c] and filled_mask [h,w] are updated in place (filled_mask may be
 * NULL: no earlier fills); coverage NULL = no coverage information.
 * stats5 = (valid, gutter, unobserved, remaining_unobserved, filled). */
int orc_pad_texture_atlas(float* color, uint32_t width, uint32_t height, uint32_t channels, const float* valid_mask,
                          const float* coverage, float* filled_mask, uint32_t margin, int fill_unobserved,
                          uint32_t maximum_fill_distance, float valid_threshold, uint64_t* stats5);

/* src/texture_padding.cpp:187-222 */
float orc_suggest_softmax_scale(const float* pos3, uint32_t V, const float* cameras3, uint32_t ncameras);

#ifdef __cplusplus
}
#endif
#endif
