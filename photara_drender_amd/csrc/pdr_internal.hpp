// Host-side interface between the C ABI (pdr_api.cpp) and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pdr_device.hpp"

namespace pdr {

// Per-view table passed by value as a kernel argument (<= kMaxViews views).
struct ViewDesc {
    const float* pos;         // [V,4] clip positions
    float* raster;            // [H,W,4]
    float* deriv;             // [H,W,4] or null
    const float* grad_raster; // [H,W,4]
    float* grad_pos;          // [V,4]
    uint32_t width, height;
    uint32_t tiles_x, tiles_y;
    float vx, vy, vw, vh;     // resolved viewport
    uint32_t tile_base;       // first global tile of this view
    uint32_t pixel_base;      // first global pixel of this view
    uint32_t rtiles_x, rtiles_y;  // kRasterTile x kRasterTile tiles of the tiled forward
    uint32_t rtile_base;          // first global raster tile of this view
    uint32_t key_base;            // first global (depth, id) key of this view (key_index)
    uint32_t key_tx;              // 4 x 4 key tiles per row
};

struct ViewTable {
    ViewDesc v[kMaxViews];
    uint32_t n;
    uint32_t total_tiles;
    uint32_t total_pixels;
    uint32_t total_rtiles;
    uint32_t total_keys;
};

// Layout of the triangle-parallel raster's 64-bit (depth, id) keys: each view's
// keys are stored in 4 x 4-pixel tiles of 128 bytes (one cache line), the tiles
// in row-major order, and each view's keys start at a multiple of 64 keys.  A
// triangle's candidates - a few pixels on each of a few rows - then fall into
// few lines, so one wave's memory-side atomicMin instruction touches fewer
// lines than with row-major keys (the candidate pass 0.75 -> 0.48 ms on the
// 16-view workload), and a 64-lane pass over the keys stays in one view.
constexpr uint32_t kKeyTile = 4;
__host__ __device__ __forceinline__ uint32_t key_tiles(uint32_t n) { return (n + kKeyTile - 1) / kKeyTile; }
__host__ __device__ __forceinline__ uint64_t view_key_count(uint32_t w, uint32_t h) {
    return (static_cast<uint64_t>(key_tiles(w)) * key_tiles(h) * (kKeyTile * kKeyTile) + 63u) & ~uint64_t{63};
}
// Key of pixel (x, y) of a view with key_tx key tiles per row, from its key_base.
__host__ __device__ __forceinline__ uint32_t key_index(uint32_t x, uint32_t y, uint32_t key_tx) {
    return ((y >> 2) * key_tx + (x >> 2)) * (kKeyTile * kKeyTile) + ((y & 3u) << 2) + (x & 3u);
}

// Generic per-view pixel table for interpolation / texture kernels.
struct PixelView {
    const float* raster;   // [H,W,4]
    const float* in0;      // kernel-specific input  [H,W,*]
    const float* in1;      // kernel-specific input  [H,W,*]
    float* out0;           // kernel-specific output [H,W,*]
    float* out1;           // kernel-specific output [H,W,*]
    uint32_t width, height;
    uint32_t tiles_x;     // kGradTile-wide tiles per row (interpolation backward)
    uint32_t tile_base;   // first such tile of this view
    uint32_t pixel_base;
};

struct PixelTable {
    PixelView v[kMaxViews];
    uint32_t n;
    uint32_t total_tiles;
    uint32_t total_pixels;
    uint32_t band_rows;  // interpolation backward: rows per wave band (tiles are kGradTile x 4 bands; set by the launcher)
};

// Scratch of the triangle-parallel raster (owned by the context).
struct RasterScratch {
    unsigned long long* keys = nullptr;  // [total_keys] (depth, id) keys (key_index layout); all ~0 between calls
    // tiled forward (raster tiles of kRasterTile^2 pixels)
    uint32_t* bin_counts = nullptr;      // [total_rtiles + 1] triangles binned per tile (0 between calls)
    uint32_t* bin_offsets = nullptr;     // [total_rtiles + 1] exclusive scan of bin_counts
    uint32_t* bin_list = nullptr;        // [bin_capacity] triangle ids, grouped by tile
    uint8_t* bin_seeded = nullptr;       // [total_rtiles] 1 where an unbinned triangle left keys in `keys`
    uint32_t* medium_list = nullptr;     // [views * T] per view: triangles too large to bin
    uint32_t* medium_count = nullptr;    // [kMaxViews]
    void* tri_records = nullptr;         // [views * T] tri_record_bytes() each: binned triangles' setup
    unsigned long long* tri_spans = nullptr;  // [views * T] binned triangles' tile ranges (0: not binned)
    void* bin_temp = nullptr;            // scan temporary
    size_t bin_temp_bytes = 0;
    uint4* queue = nullptr;              // [views * T] (view, triangle, slabs, 0) of large triangles
    uint32_t* counters = nullptr;        // [4] queue length, next queue entry (0 between calls)
    uint32_t* cid = nullptr;             // [total_keys] backward: compared id per pixel (at key_base + key_index)
    float2* cgrad = nullptr;             // [total_keys] backward: grad_raster.xy where cid != 0 (same layout)
    uint32_t* view_flags = nullptr;      // [kMaxViews] backward: view walked over whole boxes (0 between calls)
    uint8_t* won = nullptr;              // [views * T, padded to words] backward: 1 where a pixel's compared id names the triangle (0 between calls)
    const uint32_t* slot_of = nullptr;   // [3T] backward: where each corner's gradient is written (VertexCsr)
};

// Vertex -> corner adjacency (corners ascending per vertex), kept across calls
// with the indices it was built from (rebuilt on the device when they differ).
struct VertexCsr {
    uint32_t* counts = nullptr;   // [V+1] all zeros between calls
    uint32_t* offsets = nullptr;  // [V+1]
    uint32_t* corners = nullptr;  // [3T] corner ids, ascending per vertex
    uint32_t* slot_of = nullptr;  // [3T] position of each corner in `corners`
    uint32_t* saved = nullptr;    // [3T] the indices the adjacency was built from
    uint32_t* totals = nullptr;   // [csr_total_words(V)] scan tile totals
    uint32_t* state = nullptr;    // [8] stale, built, T, V, blocks finished
};

// ---- raster.hip
hipError_t fill_keys(unsigned long long* keys, size_t n, hipStream_t st);
// k_raster_small + k_raster_large: (depth, id) keys of every covered candidate pixel
// Launch parameters of the raster kernels, chosen per workload class by the
// autotuner (every choice gives the same results: the forward's winner is
// order-free and both backward paths keep the reference's per-triangle order).
struct TuneParams {
    float wave_work = 16384.0f;  // forward: candidate estimate walked by one wave (else queued as slabs)
    uint32_t fwd_waves = 4096;   // forward: target wave count (sets triangles per wave, 1..64)
    uint32_t tiled = 0;          // forward with outputs: 1 = binned into raster tiles (LDS keys), 0 = global keys
    uint32_t bin_tiles = 4;      // tiled forward: most raster tiles a binned triangle may touch (else slabs)
    float flat_work = 512.0f;    // backward: region estimate walked in the concatenated row list
    uint32_t bwd_waves = 4096;   // backward: target wave count
    uint32_t order = 0;          // forward candidates: 0 = index order, 1 = per view, the nearer end first
};

hipError_t raster_candidates(const ViewTable& views, const uint32_t* idx, uint32_t T, uint32_t V, uint32_t cull,
                             const RasterScratch& s, const TuneParams& tp, hipStream_t st);
// Tiled forward with outputs (raster, derivatives): bins the triangles into
// raster tiles, walks each tile's triangles with LDS keys and resolves the
// tile; triangles touching more than tp.bin_tiles tiles take the slab path
// through `keys` first.  Capacity of bin_list: tp.bin_tiles * views * T.
size_t bin_scan_temp_bytes(uint32_t n);
size_t tri_record_bytes();
hipError_t raster_tiled(const ViewTable& views, const uint32_t* idx, uint32_t T, uint32_t V, uint32_t cull,
                        int want_deriv, const RasterScratch& s, const TuneParams& tp, hipStream_t st);
// k_raster_resolve: raster / derivatives from the keys; resets the keys and the queue counters
hipError_t raster_resolve(const ViewTable& views, const uint32_t* idx, int want_deriv, const RasterScratch& s,
                          hipStream_t st);
size_t csr_total_words(uint32_t V);
// *flag |= 1 when any of the n words of a and b differ.
hipError_t mark_if_differ(const uint32_t* a, const uint32_t* b, size_t n, uint32_t* flag, hipStream_t st);
hipError_t build_vertex_csr(const uint32_t* idx, uint32_t T, uint32_t V, const VertexCsr& csr, hipStream_t st);
// k_bwd_prepass: per-pixel compared ids and per-view whole-box flags (raster, grad_raster of each view)
hipError_t backward_prepass(const ViewTable& views, const uint32_t* idx, uint32_t T, uint32_t V,
                            const RasterScratch& s, hipStream_t st);
// corner: [views * 3T] float4 per-triangle corner gradients (after backward_prepass);
// ordered: boxes over 65,536 pixels summed in the reference's row-major order
// (else a fixed workgroup tree)
hipError_t corner_gradients(const ViewTable& views, const uint32_t* idx, uint32_t T, uint32_t V,
                            const RasterScratch& s, float4* corner, const TuneParams& tp, bool ordered,
                            hipStream_t st);
// (also clears won, view_flags and the large-box queue counters for the next call)
hipError_t vertex_reduce(const ViewTable& views, uint32_t T, uint32_t V, const VertexCsr& csr, const float4* corner,
                         uint8_t* won, uint32_t* view_flags, uint32_t* counters, hipStream_t st);

// ---- interp_texture.hip
hipError_t interp_forward(const PixelTable& views, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                          uint32_t T, hipStream_t st);
hipError_t interp_backward(const PixelTable& views, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                           uint32_t T, float* grad_attr, hipStream_t st);
hipError_t texture_forward(const PixelTable& views, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                           uint32_t mode, hipStream_t st);
hipError_t texture_backward(const PixelTable& views, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                            uint32_t mode, int compute_uv, float* grad_tex, hipStream_t st);
// PDR_FLAG_DETERMINISTIC: grad_attr / grad_tex rewritten by ordered gathers
// (bit-reproducible, the reference's summation order).  Scratch: keys and
// values in / out [entries = 3 or 4 per pixel], starts [V + 1 or texels + 1],
// radix-sort temp storage of deterministic_temp_bytes(entries).
struct DeterministicScratch {
    uint32_t *keys_in, *keys_out, *vals_in, *vals_out, *starts;
    void* temp;
    size_t temp_bytes;
};
size_t deterministic_temp_bytes(uint32_t entries);
hipError_t interp_backward_deterministic(const PixelTable& views, uint32_t V, uint32_t A, const uint32_t* idx,
                                         uint32_t T, float* grad_attr, const DeterministicScratch& s, hipStream_t st);
hipError_t texture_backward_deterministic(const PixelTable& views, uint32_t tw, uint32_t th, uint32_t C,
                                          uint32_t mode, float* grad_tex, const DeterministicScratch& s,
                                          hipStream_t st);
// uv validity check of texture_backward's covered pixels; flags[0] |= 1 non-finite, 2 too large
hipError_t texture_check_uv(const PixelTable& views, uint32_t tw, uint32_t th, uint32_t* flags, hipStream_t st);
// max index check; flags[0] = 1 when any index >= V
hipError_t check_indices(const uint32_t* idx, size_t n, uint32_t V, uint32_t* flags, hipStream_t st);

// ---- pdr_noalloc.hip
// Takes the device address of the no-allocation scratch array on `device`
// (once per device; the device must be current).
void noalloc_scratch_init(int device);

// ---- common kernels (raster.hip)
hipError_t fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st);
hipError_t fill_u64(unsigned long long* p, unsigned long long value, size_t n, hipStream_t st);
hipError_t fill_f32(float* p, float value, size_t n, hipStream_t st);

// ---- refine.hip
struct RefineStepTables {
    const float* lr;        // [steps]
    const float* bc1;       // [steps]
    const float* bc2;       // [steps]
    const float* norm;      // [steps]
    const uint32_t* views;  // [steps * batch]
};

// One view of the refinement, reduced to its pixels with mask > 0 ("active"),
// in raster order.
struct RefineView {
    const float4* uvm;  // [active] (u, v, mask, raster index as bits)
    const float4* rgb;  // [active] target (r, g, b, 0)
    float* loss;        // [pixels] per-pixel loss term of the view's last step (0 without mask)
    uint32_t active;    // pixels with mask > 0 (known on the host once precompute has read the counts back)
    uint32_t pixels;
};

// Per-group photometric losses of the last `period` steps: step s writes row
// s % period, slot b at b * max_groups (row = batch * max_groups floats, groups
// past a view's last hold 0; max_groups is a multiple of 4).
struct LossParts {
    float* part;
    uint32_t period;
    uint32_t max_groups;
    size_t row;
};

// The step a refinement kernel works on: the device step counter plus the
// step's position in a captured block of steps (the counter moves on once per
// block, or once per step outside graphs).
struct StepIndex {
    const uint32_t* counter;
    uint32_t add;
};
__device__ __forceinline__ uint32_t step_of(const StepIndex& s) { return *s.counter + s.add; }

// The bilinear taps of one recurring step batch (the views of its batch slots)
// merged and grouped by texel, ascending.  Entry = (slot * stride + active
// pixel) * 4 + tap; within a texel the entries keep slot-then-pixel order.
struct RefineBatchTaps {
    const uint32_t* texels;       // [runs] texels the batch samples
    const uint32_t* tap_offsets;  // [runs + 1] entry range of each texel
    const uint32_t* taps;         // [entries], ascending within a texel
    const float* tap_weights;     // [entries] bilinear weight of the tap
    const uint32_t* runs;         // device count of runs
};

hipError_t transform_positions(const float* pos3, uint32_t V, const float* m16_host, float* out4, hipStream_t st);
// clips + k * V * 4 = [pos, 1] transformed by the k-th row-major matrix of m16_dev (device), k < nviews.
hipError_t transform_views(const float* pos3, uint32_t V, const float* m16_dev, uint32_t nviews, float* clips,
                           hipStream_t st);
uint32_t refine_partial_count(uint32_t P);
// One view of a refinement precompute group: its inputs, per-view scratch
// and the packed outputs in the view arena.
struct PrecomputeView {
    const float* image;    // [P, channels]
    const float* vis;      // [P] visibility mask or nullptr
    const float* raster;   // [P, 4]
    const float* uv;       // [P, 2] interpolated atlas uv
    float* target;         // [P, 3]
    float* mask;           // [P]
    double* partials;      // [refine_partial_count(P)] block sums of the mask
    double* weight;        // the view's mask sum (fixed-order double)
    uint32_t* tflags;      // [tile slots] active flags, 8x8-tile order
    uint32_t* tpos;        // [tile slots] their exclusive scan: packed index
    uint32_t* rflags;      // [P] active flags, raster order
    uint32_t* rpos;        // [P] their exclusive scan: raster rank
    uint32_t* count;       // active pixels
    float4* uvm;           // [P] packed (u, v, mask, pixel bits)
    float4* rgb;           // [P] packed target
    float* lterm;          // [P] per-pixel loss terms (cleared)
    uint32_t* order;       // [P] raster rank -> packed index
    uint32_t channels, width, height;
};
// Mask sums and active-pixel packing of every view of a group (one launch per
// pass, views on the grid's y axis); totals: [2 * nviews * active_scan_tiles(
// max(max_pixels, max_slots))] scan tile totals.
uint32_t active_scan_tiles(uint32_t n);
hipError_t precompute_views(const PrecomputeView* views_dev, uint32_t nviews, uint32_t max_pixels,
                            uint32_t max_slots, uint32_t* totals, hipStream_t st);
// Residual pass over the step's views, then the texel gather of the step's
// batch (batches[step % nbatches]); resid holds batch slots of `stride` entries.
// The joint seam term's residual (seam_step's first kernel), run as one more
// slice of the photometric residual launch when npairs > 0; no loss is kept.
struct SeamResidualArgs {
    const float* pairs = nullptr;
    uint32_t npairs = 0;
    float eps = 0.0f, norm = 0.0f, weight = 0.0f;
    float4* resid = nullptr;
};
// views_dev: [steps * batch] view records, one per step slot (slot b of step s at s * batch + b).
hipError_t photometric_step(const RefineView* views_dev, uint32_t batch, uint32_t max_active,
                            const RefineBatchTaps* batches_dev, uint32_t nbatches, uint32_t max_runs,
                            const float* tex, uint32_t tw, uint32_t th, float eps, const RefineStepTables& tables,
                            StepIndex step_at, float4* resid, uint32_t stride, float* grad,
                            const SeamResidualArgs& seam, hipStream_t st);
// The step's photometric group losses, computed by extra blocks of the Adam
// kernel that follows the step's gather (blocks = 0: none).
struct PhotoLossArgs {
    const RefineView* views = nullptr;
    uint32_t batch = 0;
    LossParts parts{};
    uint32_t blocks = 0;
};
PhotoLossArgs photo_loss_args(const RefineView* views_dev, uint32_t batch, const LossParts& parts);
// loss[s] for steps first .. first + count - 1 (count <= parts.period), from their loss partials.
hipError_t loss_fold(const LossParts& parts, uint32_t first, uint32_t count, float* loss, hipStream_t st);
// Slots of a view in 8x8-tile order (whole tiles, >= its pixels).
uint32_t tile_slots(uint32_t W, uint32_t H);
uint32_t pixel_groups(uint32_t n);
// norm[s] = 1 / (3 * sum of the step's views' mask sums) (0 when that sum is 0)
hipError_t norm_table(const double* view_weight, const uint32_t* view_index, uint32_t steps, uint32_t batch,
                      float* norm, hipStream_t st);
// The batch tap lists key texels in 8x8-tile order when the key space fits 32
// bits: tiles per row (0: linear texel keys) and the number of keys.
uint32_t texel_key_tiles(uint32_t tw, uint32_t th);
uint64_t texel_key_range(uint32_t tw, uint32_t th);
// One slot of a recurring step batch: its view's packed pixels, raster-order
// index and active count, and where its tap entries start in the batch's arrays.
struct BatchSlot {
    const float4* uvm;
    const uint32_t* order;
    uint32_t count;
    uint32_t offset;
};
// Tap entries of every slot of one batch (slots_dev: [nslots] device array),
// each slot's generated in raster order (order[raster rank] = packed index).
hipError_t batch_taps(const BatchSlot* slots_dev, uint32_t nslots, uint32_t max_count, uint32_t stride, uint32_t tw,
                      uint32_t th, uint32_t* keys, uint32_t* vals, hipStream_t st);
// Weights of sorted batch tap entries; slot_uvm[slot] = that slot's packed pixels (device array).
hipError_t batch_tap_weights(const uint32_t* vals, uint32_t E, const float4* const* slot_uvm, uint32_t stride,
                             uint32_t tw, uint32_t th, float* w, hipStream_t st);
// Sorts n (texel, entry) pairs by texel (stably) and run-length encodes the
// first E of them: texels, offsets ([runs + 1]) and the device run count.
size_t photo_group_temp_bytes(uint32_t n);
hipError_t photo_group_taps(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                            uint32_t n, int end_bit, uint32_t E, uint32_t* texels, uint32_t* runlen,
                            uint32_t* offsets, uint32_t* runs, void* temp, size_t temp_bytes, hipStream_t st);
// Taps of fixed sample points grouped by texel (as in RefineBatchTaps).
struct TapGroups {
    const uint32_t* texels;
    const uint32_t* offsets;
    const uint32_t* taps;
    const float* weights;
    uint32_t ntexels;
};
// loss_part (or null): [step][group] losses of the shader's 64-pair groups, step = device counter.
hipError_t seam_step(const float* tex, uint32_t tw, uint32_t th, const float* pairs, uint32_t npairs, float eps,
                     float norm, float weight, StepIndex step_at, const TapGroups& taps, float4* resid,
                     float* grad, float* loss_part, bool with_residual, hipStream_t st);
// loss[k] for seam steps k < steps from their group partials.
hipError_t seam_loss_fold(const float* part, uint32_t npairs, uint32_t steps, float* loss, hipStream_t st);
hipError_t seam_taps(const float* pairs, uint32_t npairs, uint32_t tw, uint32_t th, uint32_t* keys, uint32_t* vals,
                     hipStream_t st);
hipError_t seam_tap_weights(const uint32_t* vals, uint32_t E, const float* pairs, uint32_t tw, uint32_t th, float* w,
                            hipStream_t st);
hipError_t adam_tv_step(const float* tex, float* tex_out, float* grad, float* m, float* v, uint32_t tw, uint32_t th,
                        const uint8_t* tv_mask, float tv_eps, float tv_scale, const float* lr_table,
                        const float* bc1_table, const float* bc2_table, float b1, float b2, float eps, float cmin,
                        float cmax, StepIndex step_at, const PhotoLossArgs& ploss, hipStream_t st);
hipError_t advance_step(uint32_t* step_counter, uint32_t count, hipStream_t st);
// Per-texel neighbour mask of the chart-aware TV (chart may be null: one chart).
hipError_t tv_mask(const int32_t* chart, uint32_t tw, uint32_t th, uint8_t* mask, hipStream_t st);
hipError_t tv_gradient(const float* tex, uint32_t tw, uint32_t th, const int32_t* chart, float eps, float scale,
                       float* grad, hipStream_t st);
hipError_t adam_step(float* tex, float* grad, float* m, float* v, uint32_t n, const float* lr_table,
                     const float* bc1_table, const float* bc2_table, float b1, float b2, float eps, float cmin,
                     float cmax, StepIndex step_at, const PhotoLossArgs& ploss, hipStream_t st);
hipError_t tv_pair_count(const int32_t* chart, uint32_t tw, uint32_t th, unsigned long long* out, hipStream_t st);

// ---- bake.hip
// One projection view of the bake, in device memory (read by k_bake_project).
struct BakeView {
    const unsigned long long* shadow;  // the view's raster keys (order-mapped depth << 32 | id, ~0 empty), at key_index
    const float* photo;   // [h,w,channels]
    const float* mask;    // [h,w] or null
    float m[16];          // world_to_clip, row-major
    float cam[3];
    float weight;         // effective weight (times the softmax scale in softmax mode)
    float depth_bias, min_cos, mask_floor;
    float vp[4];          // resolved viewport
    uint32_t width, height, channels, index;
    uint32_t key_tx;      // key tiles per row of the shadow keys
};

hipError_t atlas_clip_positions(const float* uv, uint32_t V, float* out4, hipStream_t st);
hipError_t bake_init(uint32_t P, float* acc3, float* accw, float* best, uint32_t* src, hipStream_t st);
hipError_t bake_project(const float* araster, const float* apos, const float* anrm, uint32_t P, const BakeView* views,
                        uint32_t nviews, uint32_t blend, uint32_t pcf, float* acc3, float* accw, float* best,
                        uint32_t* src, hipStream_t st);
hipError_t bake_finish(uint32_t P, uint32_t blend, const float* araster, const float* acc3, const float* accw,
                       float* color4, float* confidence, float* valid, float* coverage, uint32_t* face_ids,
                       hipStream_t st);

// Warm-start features [P, 40]: zero before the first launch, then one launch per view chunk.
hipError_t bake_features(const float* araster, const float* apos, const float* anrm, uint32_t P,
                         const BakeView* views, uint32_t nviews, uint32_t pcf, float* feats, hipStream_t st);

// ---- mlp.hip: the warm-start MLP 40 -> 64 -> 64 -> 3
constexpr uint32_t kMlpFeatures = 40;
constexpr uint32_t kMlpParams = 6979;
constexpr uint32_t kMlpPartStride = kMlpParams + 2;  // per-wave partials: gradient, loss sum, valid count
constexpr uint32_t kMlpMaxWaves = 2048;
uint32_t mlp_grad_waves(uint64_t n);
hipError_t mlp_infer(const float* x, uint32_t n, const float* w, float* out, hipStream_t st);
hipError_t mlp_grad(const float* x, const float* target, const float* valid, uint32_t n, const float* w, float* part,
                    uint32_t waves, hipStream_t st);
hipError_t mlp_reduce(const float* part, uint32_t waves, float* grad, float* loss_count, float* params, float* m1,
                      float* m2, float lr, float beta1, float beta2, float eps, float b1c, float b2c, hipStream_t st);

// ---- padding.hip: pad_texture_atlas
// counters: valid, gutter, unobserved, remaining, filled (zeroed by the caller).
hipError_t pad_prepare(uint32_t P, const float* valid, const float* coverage, float* filled, float threshold,
                       uint8_t* source, unsigned long long* counters, hipStream_t st);
// The 8SSEDT field (field: W * H + 1 entries, the last written as a discard
// slot), rows in bands of 64 on separate workgroups handing their last rows on
// through edge [pad_band_count(H) * W] (tagged 64-bit words), status[1] (set to
// 1 when a band gave up waiting for the band above, after max_polls polls (0:
// the default bound); the field is then invalid).
uint32_t pad_band_count(uint32_t H);
hipError_t pad_field_bands(uint32_t W, uint32_t H, const uint8_t* source, uint32_t* field, unsigned long long* edge,
                           uint32_t* status, uint32_t max_polls, hipStream_t st);
// The same field with one launch per band, in order (no band waits on a band
// still running): the fallback when pad_field_bands gave up.
hipError_t pad_field_bands_ordered(uint32_t W, uint32_t H, const uint8_t* source, uint32_t* field,
                                   unsigned long long* edge, uint32_t* status, hipStream_t st);
hipError_t pad_fill(uint32_t W, uint32_t H, uint32_t C, const uint32_t* field, const uint8_t* source,
                    const float* coverage, uint32_t margin, int fill_unobserved, uint32_t max_fill,
                    const float* src_color, float* color, float* filled, unsigned long long* counters, hipStream_t st);

}  // namespace pdr
