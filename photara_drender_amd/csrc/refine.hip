// Atlas refinement kernels (src/texture_refiner.cpp:174-454).
//
//  k_transform        clip = M [p,1] per vertex (texture_refiner.cpp:36-50)
//  k_view_inputs      target RGB, mask and per-block mask sums of one view
//                     (texture_refiner.cpp:228-242)
//  k_photo_residual   atlas_photometric_gradient.hlsl:45-108 for all views of a
//                     step in one launch, up to the per-pixel output gradient
//                     d / r * m * norm and each pixel's loss term, over each
//                     view's pixels with mask > 0 packed in 8x8-tile order at
//                     precompute (k_active_flags, scans, k_active_pack); the
//                     per-step scalars come from device tables indexed by a
//                     device step counter so the step loop can be captured
//                     once and replayed
//  photo_loss_groups  (in the Adam kernels' extra blocks) the shader's 64-lane
//                     tree sum of each 64-pixel raster group's loss terms
//  k_loss_fold        each step's group losses added one after another in
//                     (batch slot, group) order, as the shader's one atomic
//                     add per group would if the groups ran in order
//  k_photo_gather     the texel gradient as a gather: the step batches recur
//                     (step s uses batch s % L), so each batch's bilinear taps
//                     are inverted once at precompute (k_batch_taps in raster
//                     order of the pixels, a stable radix sort by texel,
//                     run-length offsets, k_batch_tap_weights),
//                     and a step sums og * w over each texel's taps from all
//                     its views with one plain read-modify-write - the same
//                     terms as the reference's per-pixel atomic scatter, in a
//                     fixed order and without atomics
//  k_seam_residual /  atlas_seam_gradient.hlsl:50-100, in the same residual +
//  k_seam_gather      gather form over the (fixed) seam pairs' taps
//  k_adam             atlas_adam_update.hlsl:20-43, float4-vectorised, the
//                     shader's expression with correctly rounded sqrt and
//                     division
//  k_adam_tv          the same update with the gradient of the chart-aware
//                     total variation of python/photara_drender/atlas.py:
//                     84-115 (extension) added per texel first (hardware
//                     square root and reciprocal)
//  k_advance_step     moves the device step counter on after a step, or after
//                     a captured block of steps (each kernel of the block adds
//                     its step's position in the block: StepIndex)
//
// No gradient or loss uses float atomics (the reference emulates them with
// compare-exchange loops): every texel's terms are added in (batch slot,
// raster pixel, tap) order and every loss in (slot, group) order, so the
// parity-mode refinement is deterministic and equals the CPU restatement bit
// for bit.
#include <hipcub/hipcub.hpp>

#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr uint32_t kBlock = 256;

struct M16 {
    float m[16];
};

__global__ void k_transform(const float* __restrict__ pos3, uint32_t V, M16 M, float4* __restrict__ out) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const float x = pos3[v * 3 + 0], y = pos3[v * 3 + 1], z = pos3[v * 3 + 2];
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = M.m[k * 4] * x + M.m[k * 4 + 1] * y + M.m[k * 4 + 2] * z + M.m[k * 4 + 3];
    out[v] = make_float4(r[0], r[1], r[2], r[3]);
}

// k_transform over a group of views: view blockIdx.y with matrix mats[y], into clips + y * V.
__global__ void k_transform_views(const float* __restrict__ pos3, uint32_t V, const M16* __restrict__ mats,
                                  float4* __restrict__ clips) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const M16& M = mats[blockIdx.y];
    const float x = pos3[v * 3 + 0], y = pos3[v * 3 + 1], z = pos3[v * 3 + 2];
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = M.m[k * 4] * x + M.m[k * 4 + 1] * y + M.m[k * 4 + 2] * z + M.m[k * 4 + 3];
    clips[static_cast<size_t>(blockIdx.y) * V + v] = make_float4(r[0], r[1], r[2], r[3]);
}

// Per view (blockIdx.y) of a precompute group: target RGB, mask and block
// partials of the mask sum in double (fixed order: lanes, then waves).
__global__ __launch_bounds__(kBlock) void k_view_inputs(const PrecomputeView* __restrict__ views) {
    __shared__ double wsum[kBlock];
    const PrecomputeView& pv = views[blockIdx.y];
    const uint32_t P = pv.width * pv.height;
    if (blockIdx.x * kBlock >= P) return;  // whole block past this view
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    double w = 0.0;
    if (p < P) {
        const float* image = pv.image;
        const uint32_t channels = pv.channels;
        pv.target[static_cast<size_t>(p) * 3 + 0] = image[static_cast<size_t>(p) * channels + 0];
        pv.target[static_cast<size_t>(p) * 3 + 1] = image[static_cast<size_t>(p) * channels + 1];
        pv.target[static_cast<size_t>(p) * 3 + 2] = image[static_cast<size_t>(p) * channels + 2];
        float m = 0.0f;
        if (pv.raster[static_cast<size_t>(p) * 4 + 3] != 0.0f) m = pv.vis ? clampf(pv.vis[p], 0.0f, 1.0f) : 1.0f;
        pv.mask[p] = m;
        w = m;
    }
    wsum[threadIdx.x] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (uint32_t i = 0; i < kBlock; ++i) s += wsum[i];
        pv.partials[blockIdx.x] = s;
    }
}

// Fixed-order sum of a view's block partials (one workgroup per view): 256
// lanes add contiguous chunks, lane 0 adds the chunk sums in order.
__global__ __launch_bounds__(kBlock) void k_sum_partials(const PrecomputeView* __restrict__ views) {
    __shared__ double chunk[kBlock];
    const PrecomputeView& pv = views[blockIdx.x];
    const uint32_t n = (pv.width * pv.height + kBlock - 1) / kBlock;
    const uint32_t per = (n + kBlock - 1) / kBlock;
    const uint32_t b = threadIdx.x * per, e = b + per < n ? b + per : n;
    double s = 0.0;
    for (uint32_t i = b; i < e; ++i) s += pv.partials[i];
    chunk[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t i = 0; i < kBlock; ++i) t += chunk[i];
        *pv.weight = t;
    }
}

struct AtlasTap {
    uint32_t t[4];
    float w[4];
};

// atlas_photometric_gradient.hlsl:98-117 / atlas_seam_gradient.hlsl:182-199 (clamp addressing)
__device__ __forceinline__ AtlasTap atlas_tap(float u, float v, uint32_t tw, uint32_t th) {
    const float tx = u * static_cast<float>(tw) - 0.5f, ty = v * static_cast<float>(th) - 0.5f;
    const float flx = floorf(tx), fly = floorf(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    const float fx = tx - flx, fy = ty - fly;
    const int32_t wm = static_cast<int32_t>(tw) - 1, hm = static_cast<int32_t>(th) - 1;
    auto cl = [](int32_t c, int32_t hi) { return static_cast<uint32_t>(c < 0 ? 0 : (c > hi ? hi : c)); };
    const uint32_t x0 = cl(bx, wm), x1 = cl(add1(bx), wm), y0 = cl(by, hm), y1 = cl(add1(by), hm);
    AtlasTap a;
    a.t[0] = y0 * tw + x0;
    a.t[1] = y0 * tw + x1;
    a.t[2] = y1 * tw + x0;
    a.t[3] = y1 * tw + x1;
    a.w[0] = (1.0f - fx) * (1.0f - fy);
    a.w[1] = fx * (1.0f - fy);
    a.w[2] = (1.0f - fx) * fy;
    a.w[3] = fx * fy;
    return a;
}

// One seam pair's residual: the output gradient d / r * norm (* weight) per
// channel into resid[q]; returns the pair's loss sum_ch r * norm.
__device__ __forceinline__ float seam_pair_residual(uint32_t q, const float* __restrict__ tex, uint32_t tw, uint32_t th,
                                                    const float* __restrict__ pairs, float eps, float norm,
                                                    float weight, float4* __restrict__ resid) {
    const float4 pr = reinterpret_cast<const float4*>(pairs)[q];
    const AtlasTap a0 = atlas_tap(pr.x, pr.y, tw, th);
    const AtlasTap a1 = atlas_tap(pr.z, pr.w, tw, th);
    float g3[3];
    float local = 0.0f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float c0 = 0.0f, c1 = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            c0 += tex[static_cast<size_t>(a0.t[k]) * 3 + ch] * a0.w[k];
            c1 += tex[static_cast<size_t>(a1.t[k]) * 3 + ch] * a1.w[k];
        }
        const float d = c0 - c1;
        const float r = sqrtf(d * d + eps * eps);
        float g = d / r * norm;
        local += r * norm;
        if (weight != 1.0f) g = g * weight;
        g3[ch] = g;
    }
    resid[q] = make_float4(g3[0], g3[1], g3[2], 0.0f);
    return local;
}

// Hand-off of LDS data between lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One lane per active pixel of each batch view (packed in raster order): the
// output gradient into resid, and the pixel's loss term sum_ch r * m * norm into
// the view's dense per-pixel loss array (zero at pixels without mask, set once
// at precompute) for k_photo_loss.
__global__ __launch_bounds__(kBlock) void k_photo_residual(const RefineView* __restrict__ views, uint32_t batch,
                                                           const float* __restrict__ tex, uint32_t tw, uint32_t th,
                                                           float eps, RefineStepTables tables,
                                                           StepIndex step_at,
                                                           float4* __restrict__ resid, uint32_t max_active,
                                                           SeamResidualArgs seam) {
    if (blockIdx.y == batch) {  // the joint seam term's residual, beside the views' (no loss)
        for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < seam.npairs; q += gridDim.x * kBlock)
            (void)seam_pair_residual(q, tex, tw, th, seam.pairs, seam.eps, seam.norm, seam.weight, seam.resid);
        return;
    }
    const uint32_t step = step_of(step_at);
    const uint32_t b = blockIdx.y;
    const RefineView& rv = views[static_cast<size_t>(step) * batch + b];  // the step slot's view record
    const uint32_t a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= rv.active) return;
    const float norm = tables.norm[step];
    const float4 q = rv.uvm[a];
    const float4 target = rv.rgb[a];
    const float m = q.z;
    const AtlasTap at = atlas_tap(q.x, q.y, tw, th);
    const float tgt[3] = {target.x, target.y, target.z};
    float og[3];
    float local = 0.0f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float rendered = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) rendered += tex[static_cast<size_t>(at.t[k]) * 3 + ch] * at.w[k];
        const float d = rendered - tgt[ch];
        const float r = sqrtf(d * d + eps * eps);
        og[ch] = d / r * m * norm;
        local += r * m * norm;
    }
    resid[static_cast<size_t>(b) * max_active + a] = make_float4(og[0], og[1], og[2], 0.0f);
    rv.loss[__float_as_uint(q.w)] = local;
}

// The loss of 64-pixel raster group g of each batch view (the shader's thread
// group of pixels 64g .. 64g + 63): the shader's 64-lane tree sum of the
// pixels' loss terms, into the step's row of loss partials.  Groups past the
// view's last (up to the widest view of the batch) write 0.  Run by the first
// blocks of the Adam kernels (after the residual pass has written the terms):
// wave w of those blocks takes groups 4w .. 4w + 3 of the row (slot
// q / max_groups, group q % max_groups; max_groups is a multiple of 4).
constexpr uint32_t kLossGroupsPerWave = 4;
__device__ __forceinline__ void photo_loss_groups(const RefineView* __restrict__ views, uint32_t batch, uint32_t step,
                                                  const LossParts& parts, uint32_t w) {
    const uint32_t q0 = w * kLossGroupsPerWave;
    const uint32_t b = q0 / parts.max_groups, g0 = q0 % parts.max_groups;
    if (b >= batch) return;  // wave-uniform
    const RefineView& rv = views[static_cast<size_t>(step) * batch + b];
    const uint32_t lane = threadIdx.x & 63u;
    float v[kLossGroupsPerWave];
#pragma unroll
    for (uint32_t j = 0; j < kLossGroupsPerWave; ++j) {
        const uint32_t p = (g0 + j) * 64u + lane;
        v[j] = p < rv.pixels ? rv.loss[p] : 0.0f;
    }
    float t[kLossGroupsPerWave];
#pragma unroll
    for (uint32_t j = 0; j < kLossGroupsPerWave; ++j) t[j] = wave_tree_sum(v[j]);
    if (lane == 0) {
        float* out = parts.part + static_cast<size_t>(step % parts.period) * parts.row +
                     static_cast<size_t>(b) * parts.max_groups + g0;
        *reinterpret_cast<float4*>(out) = make_float4(t[0], t[1], t[2], t[3]);
    }
}

// loss[step] of steps first .. first + count - 1: each step's group losses
// (rows of batch * max_groups partials) added one after another in (batch slot,
// group) order, as the shader's one add per group lands when the groups run in
// that order.  The shader skips zero group sums; every partial is >= +0 (a sum
// of r * m * norm terms) and so is the running sum, so adding a zero changes
// nothing.  One workgroup per step: waves 1-3 stream the row through LDS in
// chunks of kFoldChunk float4 (double-buffered) while lane 0 of wave 0 adds
// the previous chunk in order, so the serial chain never waits on memory.
constexpr uint32_t kFoldChunk = 1024;

__global__ __launch_bounds__(kBlock) void k_loss_fold(LossParts parts, uint32_t first, float* __restrict__ loss) {
    __shared__ float4 chunk[2][kFoldChunk];
    const uint32_t step = first + blockIdx.x, tid = threadIdx.x;
    const float4* row = reinterpret_cast<const float4*>(parts.part + static_cast<size_t>(step % parts.period) * parts.row);
    const uint32_t n4 = static_cast<uint32_t>(parts.row / 4u);
    constexpr uint32_t kLoaders = kBlock - 64u;
    auto stage = [&](uint32_t c, uint32_t buf) {  // chunk starting at float4 c into chunk[buf] (waves 1-3)
        if (tid < 64u) return;
        for (uint32_t i = tid - 64u; i < kFoldChunk && c + i < n4; i += kLoaders) chunk[buf][i] = row[c + i];
    };
    stage(0, 0);
    __syncthreads();
    float acc = 0.0f;
    uint32_t buf = 0;
    for (uint32_t c = 0; c < n4; c += kFoldChunk, buf ^= 1u) {
        stage(c + kFoldChunk, buf ^ 1u);
        if (tid == 0) {
            const uint32_t m = n4 - c < kFoldChunk ? n4 - c : kFoldChunk;
            for (uint32_t i = 0; i < m; ++i) {
                const float4 v = chunk[buf][i];
                acc += v.x;
                acc += v.y;
                acc += v.z;
                acc += v.w;
            }
        }
        __syncthreads();
    }
    if (tid == 0) loss[step] = acc;
}

// Texel t of a tw-wide atlas as a key in 8x8-tile order (tiles_x = tiles per
// row; 0: the key is t itself).  Sorting taps by it makes the gather's 64 lanes
// take an 8x8 texel block, whose taps come from a compact image region.
__device__ __forceinline__ uint32_t texel_key(uint32_t t, uint32_t tw, uint32_t tiles_x) {
    if (!tiles_x) return t;
    const uint32_t x = t % tw, y = t / tw;
    return ((y >> 3) * tiles_x + (x >> 3)) * 64u + ((y & 7u) << 3) + (x & 7u);
}
__device__ __forceinline__ uint32_t key_texel(uint32_t k, uint32_t tw, uint32_t tiles_x) {
    if (!tiles_x) return k;
    const uint32_t tile = k >> 6;
    return ((tile / tiles_x) * 8u + ((k >> 3) & 7u)) * tw + (tile % tiles_x) * 8u + (k & 7u);
}

// Texel gradient of the step's batch: one lane per texel the batch samples,
// its taps from all slots added in list order.  A wave owns 64 consecutive
// texels of the list, whose tap entries are contiguous: it walks them 256 at a
// time, four entries per lane (coalesced tap and weight loads, four residual
// gathers in flight per lane, the next window's entries loaded before this
// window's products are summed), stages the products og * w in LDS, and each
// lane adds those of its own texel in entry order - the same sums, in the same
// order, as one lane walking its texel's entries.
constexpr uint32_t kGatherPer = 4;  // entries per lane per window
constexpr uint32_t kGatherWindow = 64u * kGatherPer;
// The gather of one wave over texels [first, first + 64) of a texel list
// (texel keys, tap offsets with one entry past the last texel, taps, tap
// weights; tap >> kShift indexes the residuals; tiles_x = 0: keys are texels).
template <uint32_t kShift>
__device__ __forceinline__ void texel_gather(const uint32_t* __restrict__ texels, const uint32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ taps, const float* __restrict__ weights,
                                             uint32_t runs, const float4* __restrict__ resid, uint32_t tw,
                                             uint32_t tiles_x, float* __restrict__ grad,
                                             float4 (*__restrict__ prod)[kGatherWindow]) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t first = blockIdx.x * kBlock + wave * 64u;
    if (first >= runs) return;  // wave-uniform
    const uint32_t i = first + lane, end = first + 64u < runs ? first + 64u : runs;
    const bool live = i < runs;
    const uint32_t e_lo = offsets[live ? i : end], e_hi = offsets[live ? i + 1u : end];
    const uint32_t E0 = offsets[first], E1 = offsets[end];
    float a[3] = {0.0f, 0.0f, 0.0f};
    if (E1 > E0) {
        const uint32_t top = E1 - 1u;  // clamp: every load stays inside the wave's entries
        auto entry = [&](uint32_t e) { return e < top ? e : top; };
        uint32_t tp[kGatherPer], en[kGatherPer];
        float w[kGatherPer];
#pragma unroll
        for (uint32_t h = 0; h < kGatherPer; ++h) {
            en[h] = entry(E0 + 64u * h + lane);
            tp[h] = taps[en[h]];
            w[h] = weights[en[h]];
        }
        for (uint32_t c = E0; c < E1; c += kGatherWindow) {
            float4 r[kGatherPer];
            float wc[kGatherPer];
#pragma unroll
            for (uint32_t h = 0; h < kGatherPer; ++h) {
                r[h] = resid[tp[h] >> kShift];
                wc[h] = c + 64u * h + lane < E1 ? w[h] : 0.0f;
            }
#pragma unroll
            for (uint32_t h = 0; h < kGatherPer; ++h) {
                en[h] = entry(c + kGatherWindow + 64u * h + lane);
                tp[h] = taps[en[h]];
                w[h] = weights[en[h]];
            }
#pragma unroll
            for (uint32_t h = 0; h < kGatherPer; ++h)
                prod[wave][64u * h + lane] = make_float4(r[h].x * wc[h], r[h].y * wc[h], r[h].z * wc[h], 0.0f);
            wave_lds_sync();
            const uint32_t s = e_lo > c ? e_lo : c, t = e_hi < c + kGatherWindow ? e_hi : c + kGatherWindow;
            for (uint32_t k = s; k < t; ++k) {
                const float4 q = prod[wave][k - c];
                a[0] += q.x;
                a[1] += q.y;
                a[2] += q.z;
            }
            wave_lds_sync();
        }
    }
    if (!live) return;
    const size_t t = static_cast<size_t>(key_texel(texels[i], tw, tiles_x)) * 3;
    grad[t] += a[0];
    grad[t + 1] += a[1];
    grad[t + 2] += a[2];
}

__global__ __launch_bounds__(kBlock) void k_photo_gather(const RefineBatchTaps* __restrict__ batches, uint32_t nbatches,
                                                         uint32_t tw, uint32_t tiles_x,
                                                         StepIndex step_at,
                                                         const float4* __restrict__ resid, float* __restrict__ grad) {
    __shared__ float4 prod[kBlock / 64][kGatherWindow];
    const RefineBatchTaps& bt = batches[step_of(step_at) % nbatches];
    texel_gather<2>(bt.texels, bt.tap_offsets, bt.taps, bt.tap_weights, *bt.runs, resid, tw, tiles_x, grad, prod);
}

// Tile-order slots of a W x H view (whole 8x8 tiles).
__device__ __forceinline__ uint32_t tile_slots_of(uint32_t W, uint32_t H) {
    return ((W + 7u) / 8u) * ((H + 7u) / 8u) * 64u;
}

// Tile-order slot q of a W x H view: tile q / 64 (row-major over 8x8 tiles),
// pixel (q % 64) row-major inside the tile.
__device__ __forceinline__ bool tile_pixel(uint32_t q, uint32_t W, uint32_t H, uint32_t& p) {
    const uint32_t tiles_x = (W + 7u) / 8u, tile = q >> 6, lane = q & 63u;
    const uint32_t x = (tile % tiles_x) * 8u + (lane & 7u), y = (tile / tiles_x) * 8u + (lane >> 3);
    p = y * W + x;
    return x < W && y < H;
}

// Active flags of a view (blockIdx.y) over its 8x8-tile-order slots and over raster order.
__global__ void k_active_flags(const PrecomputeView* __restrict__ views) {
    const PrecomputeView& pv = views[blockIdx.y];
    const uint32_t W = pv.width, H = pv.height;
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q < tile_slots_of(W, H)) {
        uint32_t p = 0;
        const bool in = tile_pixel(q, W, H, p);
        pv.tflags[q] = in && pv.mask[p] > 0.0f ? 1u : 0u;
    }
    if (q < W * H) pv.rflags[q] = pv.mask[q] > 0.0f ? 1u : 0u;
}

// Exclusive scans of both flag arrays of every view of the group (segment
// 2 y + z: z = 0 tile order, with the view's active count; z = 1 raster
// order), in tiles of kScanTile: each tile's total, then each tile's scan
// offset by the totals of the tiles before it.
constexpr uint32_t kScanTile = 2048;  // 256 lanes x 8

__device__ __forceinline__ void scan_segment(const PrecomputeView& pv, uint32_t z, const uint32_t*& in, uint32_t*& out,
                                             uint32_t& n) {
    in = z ? pv.rflags : pv.tflags;
    out = z ? pv.rpos : pv.tpos;
    n = z ? pv.width * pv.height : tile_slots_of(pv.width, pv.height);
}

__global__ __launch_bounds__(kBlock) void k_active_scan_totals(const PrecomputeView* __restrict__ views,
                                                               uint32_t tiles, uint32_t* __restrict__ totals) {
    using Reduce = hipcub::BlockReduce<uint32_t, kBlock>;
    __shared__ typename Reduce::TempStorage tmp;
    const uint32_t* in;
    uint32_t* out;
    uint32_t n;
    scan_segment(views[blockIdx.y], blockIdx.z, in, out, n);
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 8;
    if (blockIdx.x * kScanTile >= n) return;
    uint32_t s = 0;
    for (uint32_t k = 0; k < 8; ++k)
        if (base + k < n) s += in[base + k];
    s = Reduce(tmp).Sum(s);
    if (threadIdx.x == 0) totals[(static_cast<size_t>(blockIdx.y) * 2 + blockIdx.z) * tiles + blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_active_scan_out(const PrecomputeView* __restrict__ views, uint32_t tiles,
                                                            const uint32_t* __restrict__ totals) {
    using Reduce = hipcub::BlockReduce<uint32_t, kBlock>;
    using Scan = hipcub::BlockScan<uint32_t, kBlock>;
    __shared__ typename Reduce::TempStorage rtmp;
    __shared__ typename Scan::TempStorage stmp;
    __shared__ uint32_t prefix;
    const PrecomputeView& pv = views[blockIdx.y];
    const uint32_t* in;
    uint32_t* out;
    uint32_t n;
    scan_segment(pv, blockIdx.z, in, out, n);
    if (blockIdx.x * kScanTile >= n) return;
    const uint32_t* tot = totals + (static_cast<size_t>(blockIdx.y) * 2 + blockIdx.z) * tiles;
    uint32_t s = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kBlock) s += tot[b];
    s = Reduce(rtmp).Sum(s);
    if (threadIdx.x == 0) prefix = s;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 8;
    uint32_t c[8], t = 0;
    for (uint32_t k = 0; k < 8; ++k) {
        c[k] = base + k < n ? in[base + k] : 0u;
        t += c[k];
    }
    uint32_t run;
    Scan(stmp).ExclusiveSum(t, run);
    __syncthreads();
    run += prefix;
    for (uint32_t k = 0; k < 8; ++k) {
        if (base + k < n) out[base + k] = run;
        run += c[k];
    }
    // the last tile's last lane holds the segment total: the view's active count
    if (blockIdx.z == 0 && (blockIdx.x + 1) * kScanTile >= n && threadIdx.x == kBlock - 1) *pv.count = run;
}

// Packed record a = tpos[q] of the active pixel p at tile-order slot q:
// (u, v, mask, p as bits) and the target; order[raster rank of p] = a; every
// pixel's loss term cleared.
__global__ void k_active_pack(const PrecomputeView* __restrict__ views) {
    const PrecomputeView& pv = views[blockIdx.y];
    const uint32_t W = pv.width, H = pv.height;
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= tile_slots_of(W, H)) return;
    uint32_t p = 0;
    if (!tile_pixel(q, W, H, p)) return;
    pv.lterm[p] = 0.0f;
    if (!pv.tflags[q]) return;
    const uint32_t a = pv.tpos[q];
    const float* uv = pv.uv;
    const float* target = pv.target;
    pv.uvm[a] = make_float4(uv[static_cast<size_t>(p) * 2], uv[static_cast<size_t>(p) * 2 + 1], pv.mask[p],
                            __uint_as_float(p));
    pv.rgb[a] = make_float4(target[static_cast<size_t>(p) * 3], target[static_cast<size_t>(p) * 3 + 1],
                            target[static_cast<size_t>(p) * 3 + 2], 0.0f);
    pv.order[pv.rpos[p]] = a;
}

// Tap entries of one recurring batch, slot blockIdx.y over its view's active
// pixels, in raster order (order[r] = packed index of the pixel of raster
// rank r), so that a stable sort by texel leaves every texel's entries in
// (slot, raster pixel, tap) order: key = texel, value = (slot * stride + a) *
// 4 + tap, written from the slot's offset into the batch's arrays.
__global__ void k_batch_taps(const BatchSlot* __restrict__ slots, uint32_t stride, uint32_t tw, uint32_t th,
                             uint32_t tiles_x, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const BatchSlot& bs = slots[blockIdx.y];
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= bs.count) return;
    const uint32_t a = bs.order[r];
    const float4 q = bs.uvm[a];
    const AtlasTap at = atlas_tap(q.x, q.y, tw, th);
    const uint32_t base = (blockIdx.y * stride + a) * 4u;
    const size_t o = static_cast<size_t>(bs.offset) + static_cast<size_t>(r) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        keys[o + k] = texel_key(at.t[k], tw, tiles_x);
        vals[o + k] = base + static_cast<uint32_t>(k);
    }
}

// Bilinear weight of every sorted batch tap entry (the same expression as the scatter's).
__global__ void k_batch_tap_weights(const uint32_t* __restrict__ vals, uint32_t E,
                                    const float4* const* __restrict__ slot_uvm, uint32_t stride, uint32_t tw,
                                    uint32_t th, float* __restrict__ w) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    const uint32_t v = vals[e], r = v >> 2;
    const float4 q = slot_uvm[r / stride][r % stride];
    const AtlasTap at = atlas_tap(q.x, q.y, tw, th);
    w[e] = at.w[v & 3u];
}

// src/texture_refiner.cpp:315-330: the per-step normalisation from the views' mask sums.
__global__ void k_norm_table(const double* __restrict__ view_weight, const uint32_t* __restrict__ view_index,
                             uint32_t steps, uint32_t batch, float* __restrict__ norm) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= steps) return;
    double bw = 0.0;
    for (uint32_t b = 0; b < batch; ++b) bw += view_weight[view_index[static_cast<size_t>(s) * batch + b]];
    norm[s] = bw > 0.0 ? static_cast<float>(1.0 / (bw * 3.0)) : 0.0f;
}

// offsets[U] = E (the run starts come from an exclusive scan of the run lengths).
__global__ void k_set_last(uint32_t* __restrict__ offsets, const uint32_t* __restrict__ runs, uint32_t E) {
    if (threadIdx.x == 0 && blockIdx.x == 0) offsets[*runs] = E;
}

// Seam pairs (atlas_seam_gradient.hlsl:50-100) up to the per-pair output
// gradient g = d / r * norm (* weight); the texel gradient is gathered by
// k_seam_gather over the pairs' taps grouped by texel (k_seam_taps at setup).
// With loss partials, each wave (the shader's group of 64 pairs) writes its
// tree-summed loss to part[step * groups + group] for k_seam_loss_fold.
__global__ __launch_bounds__(kBlock) void k_seam_residual(const float* __restrict__ tex, uint32_t tw, uint32_t th,
                                                          const float* __restrict__ pairs, uint32_t npairs, float eps,
                                                          float norm, float weight,
                                                          StepIndex step_at,
                                                          float4* __restrict__ resid, float* __restrict__ part) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    const float local = q < npairs ? seam_pair_residual(q, tex, tw, th, pairs, eps, norm, weight, resid) : 0.0f;
    const float s = wave_tree_sum(local);
    const uint32_t groups = (npairs + 63u) / 64u;
    if (part && (threadIdx.x & 63u) == 0) part[static_cast<size_t>(step_of(step_at)) * groups + (q >> 6)] = s;
}

// loss[k] of seam step k: its group losses added in group order, zeros
// skipped.  One workgroup per step: the row is staged through LDS in coalesced
// chunks and lane 0 adds each chunk from LDS, so the serial chain of adds never
// waits on a global load.
constexpr uint32_t kSeamFoldChunk = 4096;

__global__ __launch_bounds__(kBlock) void k_seam_loss_fold(const float* __restrict__ part, uint32_t groups,
                                                           float* __restrict__ loss) {
    __shared__ float chunk[kSeamFoldChunk];
    const uint32_t k = blockIdx.x, tid = threadIdx.x;
    const float* p = part + static_cast<size_t>(k) * groups;
    float acc = 0.0f;
    for (uint32_t c = 0; c < groups; c += kSeamFoldChunk) {
        const uint32_t m = groups - c < kSeamFoldChunk ? groups - c : kSeamFoldChunk;
        for (uint32_t i = tid; i < m; i += kBlock) chunk[i] = p[c + i];
        __syncthreads();
        if (tid == 0)
            for (uint32_t i = 0; i < m; ++i)
                if (chunk[i] != 0.0f) acc += chunk[i];
        __syncthreads();
    }
    if (tid == 0) loss[k] = acc;
}

// Texel gradient of the seam term: entries (pair * 8 + tap * 2 + side) with the
// tap weight, negated on the second side, grouped by texel.  Within a pair the
// entries keep the shader's order: tap k of side 0, then tap k of side 1.
__global__ __launch_bounds__(kBlock) void k_seam_gather(const uint32_t* __restrict__ texels,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ taps,
                                                        const float* __restrict__ weights, uint32_t ntexels,
                                                        const float4* __restrict__ resid, float* __restrict__ grad) {
    __shared__ float4 prod[kBlock / 64][kGatherWindow];
    texel_gather<3>(texels, offsets, taps, weights, ntexels, resid, 0u, 0u, grad, prod);
}

__global__ void k_seam_taps(const float* __restrict__ pairs, uint32_t npairs, uint32_t tw, uint32_t th,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= npairs) return;
    const float4 pr = reinterpret_cast<const float4*>(pairs)[q];
    const AtlasTap a0 = atlas_tap(pr.x, pr.y, tw, th);
    const AtlasTap a1 = atlas_tap(pr.z, pr.w, tw, th);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        keys[static_cast<size_t>(q) * 8 + 2 * k] = a0.t[k];
        vals[static_cast<size_t>(q) * 8 + 2 * k] = q * 8u + 2u * static_cast<uint32_t>(k);
        keys[static_cast<size_t>(q) * 8 + 2 * k + 1] = a1.t[k];
        vals[static_cast<size_t>(q) * 8 + 2 * k + 1] = q * 8u + 2u * static_cast<uint32_t>(k) + 1u;
    }
}

__global__ void k_seam_tap_weights(const uint32_t* __restrict__ vals, uint32_t E, const float* __restrict__ pairs,
                                   uint32_t tw, uint32_t th, float* __restrict__ w) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    const uint32_t v = vals[e];
    const float4 pr = reinterpret_cast<const float4*>(pairs)[v >> 3];
    const bool second = (v & 1u) != 0u;
    const AtlasTap a = second ? atlas_tap(pr.z, pr.w, tw, th) : atlas_tap(pr.x, pr.y, tw, th);
    const float wt = a.w[(v >> 1) & 3u];
    w[e] = second ? -wt : wt;
}

// Native-precision square root, reciprocal and reciprocal square root
// (v_sqrt_f32, v_rcp_f32, v_rsq_f32: about 1 ulp, as the shaders' sqrt and
// division compile on a GPU).  The correctly rounded library forms take about
// 15 instructions each and left k_adam_tv bound by the VALU.  The rsq operand
// is kept at or above FLT_MIN, where the instruction is defined for every input.
__device__ __forceinline__ float native_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float native_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float charbonnier_slope(float d, float eps2) {  // d / sqrt(d^2 + eps^2)
    return d * __builtin_amdgcn_rsqf(fmaxf(d * d + eps2, 1.17549435e-38f));
}

// Neighbour mask of texel i for the chart-aware TV: bit 0 the texel lies in a
// chart (id >= 0; every texel when chart is null), bits 1-4 its left, right,
// upper and lower neighbours exist and lie in the same chart.
__device__ __forceinline__ uint32_t tv_mask_of(const int32_t* __restrict__ chart, uint32_t tw, uint32_t th,
                                               uint32_t i) {
    const uint32_t x = i % tw, y = i / tw;
    const int32_t c = chart ? chart[i] : 0;
    if (c < 0) return 0u;
    auto same = [&](uint32_t j) { return (chart ? chart[j] : 0) == c; };
    uint32_t m = 1u;
    if (x > 0 && same(i - 1)) m |= 2u;
    if (x + 1 < tw && same(i + 1)) m |= 4u;
    if (y > 0 && same(i - tw)) m |= 8u;
    if (y + 1 < th && same(i + tw)) m |= 16u;
    return m;
}

// d/dt of weight * sum_pairs sqrt((t_b - t_a)^2 + eps^2) / (pairs * 3) over
// horizontal and vertical neighbour pairs inside one chart, at texel i with
// neighbour mask `mask` (pairs left, right, up, down, added in that order).
// All five texels are loaded before any arithmetic: a neighbour outside the
// mask reads the texel itself and adds exactly zero.
__device__ __forceinline__ void tv_texel(const float* __restrict__ tex, uint32_t tw, uint32_t mask, float eps2,
                                         uint32_t i, float g[3]) {
    const uint32_t nb[4] = {mask & 2u ? i - 1u : i, mask & 4u ? i + 1u : i, mask & 8u ? i - tw : i,
                            mask & 16u ? i + tw : i};
    float t[3], u[4][3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) t[ch] = tex[static_cast<size_t>(i) * 3 + ch];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) u[q][ch] = tex[static_cast<size_t>(nb[q]) * 3 + ch];
    g[0] = g[1] = g[2] = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool on = (mask & (2u << q)) != 0u;
        const float sign = (q & 1) ? -1.0f : 1.0f;  // left / up: d = t - u, +d/r; right / down: d = u - t, -d/r
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float d = (q & 1) ? u[q][ch] - t[ch] : t[ch] - u[q][ch];
            const float s = sign * charbonnier_slope(d, eps2);
            g[ch] += on ? s : 0.0f;
        }
    }
}

__global__ void k_tv_mask(const int32_t* __restrict__ chart, uint32_t tw, uint32_t th, uint8_t* __restrict__ mask) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < tw * th) mask[i] = static_cast<uint8_t>(tv_mask_of(chart, tw, th, i));
}

__global__ __launch_bounds__(kBlock) void k_tv_pairs(const int32_t* __restrict__ chart, uint32_t tw, uint32_t th,
                                                     unsigned long long* __restrict__ out) {
    __shared__ unsigned long long wsum[kBlock / 64];
    unsigned long long n = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < tw * th; i += gridDim.x * kBlock) {
        const uint32_t x = i % tw, y = i / tw;
        const int32_t c = chart ? chart[i] : 0;
        if (c >= 0) {
            if (x + 1 < tw && (chart ? chart[i + 1] : 0) == c) ++n;
            if (y + 1 < th && (chart ? chart[i + tw] : 0) == c) ++n;
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) n += __shfl_xor(n, m, 64);
    if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < kBlock / 64; ++w) t += wsum[w];
        if (t) atomicAdd(out, t);  // one atomic per workgroup (the grid is capped)
    }
}

// atlas_adam_update.hlsl:27-42 as written: p - lr * c1 / (sqrt(c2) + eps),
// with correctly rounded division and square root.
__device__ __forceinline__ float adam_one(float p, float& g, float& m, float& v, float lr, float b1, float b2,
                                          float eps, float bc1, float bc2, float cmin, float cmax) {
    const float m1 = b1 * m + (1.0f - b1) * g;
    const float m2 = b2 * v + (1.0f - b2) * g * g;
    m = m1;
    v = m2;
    g = 0.0f;
    const float c1 = m1 / bc1, c2 = m2 / bc2;
    return clampf(p - lr * c1 / (sqrtf(c2) + eps), cmin, cmax);
}

// The same update with the hardware square root and reciprocal (about 1 ulp
// each) in place of the correctly rounded forms, for the extended objective's
// fused TV + Adam pass, which they keep off the VALU bound.
__device__ __forceinline__ float adam_one_native(float p, float& g, float& m, float& v, float lr, float b1, float b2,
                                                 float eps, float bc1, float bc2, float cmin, float cmax) {
    const float m1 = b1 * m + (1.0f - b1) * g;
    const float m2 = b2 * v + (1.0f - b2) * g * g;
    m = m1;
    v = m2;
    g = 0.0f;
    const float c1 = m1 / bc1, c2 = m2 / bc2;
    return clampf(p - lr * c1 * native_rcp(native_sqrt(c2) + eps), cmin, cmax);
}

// Adam with the chart-TV gradient (scale = weight / (pairs * 3)) added to each
// texel's gradient first, as the separate TV pass would: one lane per texel.
// Every texel reads its neighbours' pre-step values: the updated atlas goes to
// tex_out (a second buffer), and the caller swaps the two.
__global__ __launch_bounds__(kBlock) void k_adam_tv(const float* __restrict__ tex, float* __restrict__ tex_out,
                                                    float* __restrict__ grad, float* __restrict__ mom1,
                                                    float* __restrict__ mom2, uint32_t tw, uint32_t th,
                                                    const uint8_t* __restrict__ tv_mask, float tv_eps2, float tv_scale,
                                                    const float* __restrict__ lr_t, const float* __restrict__ bc1_t,
                                                    const float* __restrict__ bc2_t, float b1, float b2, float eps,
                                                    float cmin, float cmax, StepIndex step_at,
                                                    PhotoLossArgs ploss) {
    const uint32_t step = step_of(step_at);
    if (blockIdx.x < ploss.blocks) {  // the step's photometric group losses first
        photo_loss_groups(ploss.views, ploss.batch, step, ploss.parts, blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6));
        return;
    }
    const uint32_t blk = blockIdx.x - ploss.blocks, nadam = gridDim.x - ploss.blocks;
    const float lr = lr_t[step], bc1 = bc1_t[step], bc2 = bc2_t[step];
    for (uint32_t i = blk * kBlock + threadIdx.x; i < tw * th; i += nadam * kBlock) {
        const size_t b = static_cast<size_t>(i) * 3;
        float gr[3], m[3], v[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            gr[ch] = grad[b + ch];
            m[ch] = mom1[b + ch];
            v[ch] = mom2[b + ch];
        }
        float g[3];
        tv_texel(tex, tw, tv_mask[i], tv_eps2, i, g);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float gj = gr[ch] + g[ch] * tv_scale;
            tex_out[b + ch] = adam_one_native(tex[b + ch], gj, m[ch], v[ch], lr, b1, b2, eps, bc1, bc2, cmin, cmax);
            grad[b + ch] = gj;
            mom1[b + ch] = m[ch];
            mom2[b + ch] = v[ch];
        }
    }
}

// The TV gradient alone (scale * d/dt per texel), for tests of tv_texel.
__global__ void k_tv_gradient(const float* __restrict__ tex, uint32_t tw, uint32_t th,
                              const int32_t* __restrict__ chart, float eps, float scale, float* __restrict__ grad) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= tw * th) return;
    float g[3];
    tv_texel(tex, tw, tv_mask_of(chart, tw, th, i), eps * eps, i, g);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) grad[static_cast<size_t>(i) * 3 + ch] = g[ch] * scale;
}

__global__ void k_advance_step(uint32_t* step_counter, uint32_t count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_counter += count;
}

__global__ __launch_bounds__(kBlock) void k_adam(float* __restrict__ tex, float* __restrict__ grad,
                                                 float* __restrict__ mom1, float* __restrict__ mom2, uint32_t n,
                                                 const float* __restrict__ lr_t, const float* __restrict__ bc1_t,
                                                 const float* __restrict__ bc2_t, float b1, float b2, float eps,
                                                 float cmin, float cmax, StepIndex step_at,
                                                 PhotoLossArgs ploss) {
    const uint32_t step = step_of(step_at);
    if (blockIdx.x < ploss.blocks) {  // the step's photometric group losses first
        photo_loss_groups(ploss.views, ploss.batch, step, ploss.parts, blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6));
        return;
    }
    const uint32_t blk = blockIdx.x - ploss.blocks, nadam = gridDim.x - ploss.blocks;
    const float lr = lr_t[step], bc1 = bc1_t[step], bc2 = bc2_t[step];
    const uint32_t n4 = n / 4;
    for (uint32_t i = blk * kBlock + threadIdx.x; i < n4; i += nadam * kBlock) {
        float4 p = reinterpret_cast<float4*>(tex)[i];
        float4 g = reinterpret_cast<float4*>(grad)[i];
        float4 m = reinterpret_cast<float4*>(mom1)[i];
        float4 v = reinterpret_cast<float4*>(mom2)[i];
        p.x = adam_one(p.x, g.x, m.x, v.x, lr, b1, b2, eps, bc1, bc2, cmin, cmax);
        p.y = adam_one(p.y, g.y, m.y, v.y, lr, b1, b2, eps, bc1, bc2, cmin, cmax);
        p.z = adam_one(p.z, g.z, m.z, v.z, lr, b1, b2, eps, bc1, bc2, cmin, cmax);
        p.w = adam_one(p.w, g.w, m.w, v.w, lr, b1, b2, eps, bc1, bc2, cmin, cmax);
        reinterpret_cast<float4*>(tex)[i] = p;
        reinterpret_cast<float4*>(grad)[i] = g;
        reinterpret_cast<float4*>(mom1)[i] = m;
        reinterpret_cast<float4*>(mom2)[i] = v;
    }
    if (blk == 0) {
        for (uint32_t i = n4 * 4 + threadIdx.x; i < n; i += kBlock)
            tex[i] = adam_one(tex[i], grad[i], mom1[i], mom2[i], lr, b1, b2, eps, bc1, bc2, cmin, cmax);
    }
}

}  // namespace

hipError_t transform_positions(const float* pos3, uint32_t V, const float* m16, float* out4, hipStream_t st) {
    M16 M;
    for (int k = 0; k < 16; ++k) M.m[k] = m16[k];
    k_transform<<<(V + kBlock - 1) / kBlock, kBlock, 0, st>>>(pos3, V, M, reinterpret_cast<float4*>(out4));
    return hipGetLastError();
}

uint32_t refine_partial_count(uint32_t P) { return (P + kBlock - 1) / kBlock; }

hipError_t transform_views(const float* pos3, uint32_t V, const float* m16_dev, uint32_t nviews, float* clips,
                           hipStream_t st) {
    if (!V || !nviews) return hipSuccess;
    k_transform_views<<<dim3((V + kBlock - 1) / kBlock, nviews), kBlock, 0, st>>>(
        pos3, V, reinterpret_cast<const M16*>(m16_dev), reinterpret_cast<float4*>(clips));
    return hipGetLastError();
}

uint32_t active_scan_tiles(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

hipError_t precompute_views(const PrecomputeView* views_dev, uint32_t nviews, uint32_t max_pixels,
                            uint32_t max_slots, uint32_t* totals, hipStream_t st) {
    if (!nviews) return hipSuccess;
    const uint32_t pix_blocks = (max_pixels + kBlock - 1) / kBlock;
    k_view_inputs<<<dim3(pix_blocks ? pix_blocks : 1, nviews), kBlock, 0, st>>>(views_dev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_sum_partials<<<nviews, kBlock, 0, st>>>(views_dev);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t span = max_slots > max_pixels ? max_slots : max_pixels;
    k_active_flags<<<dim3((span + kBlock - 1) / kBlock, nviews), kBlock, 0, st>>>(views_dev);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t tiles = active_scan_tiles(span);
    k_active_scan_totals<<<dim3(tiles, nviews, 2), kBlock, 0, st>>>(views_dev, tiles, totals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_active_scan_out<<<dim3(tiles, nviews, 2), kBlock, 0, st>>>(views_dev, tiles, totals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_active_pack<<<dim3((max_slots + kBlock - 1) / kBlock, nviews), kBlock, 0, st>>>(views_dev);
    return hipGetLastError();
}

uint32_t texel_key_tiles(uint32_t tw, uint32_t th) {
    const uint64_t tx = (tw + 7u) / 8u, ty = (th + 7u) / 8u;
    return tx * ty * 64u <= 0xFFFFFFFFull ? static_cast<uint32_t>(tx) : 0u;
}

uint64_t texel_key_range(uint32_t tw, uint32_t th) {
    const uint32_t tx = texel_key_tiles(tw, th);
    return tx ? static_cast<uint64_t>(tx) * ((th + 7u) / 8u) * 64u : static_cast<uint64_t>(tw) * th;
}

hipError_t photometric_step(const RefineView* views_dev, uint32_t batch, uint32_t max_active,
                            const RefineBatchTaps* batches_dev, uint32_t nbatches, uint32_t max_runs,
                            const float* tex, uint32_t tw, uint32_t th, float eps, const RefineStepTables& tables,
                            StepIndex step_at, float4* resid, uint32_t stride, float* grad,
                            const SeamResidualArgs& seam, hipStream_t st) {
    const uint32_t blocks = (max_active + kBlock - 1) / kBlock;
    const dim3 grid(blocks ? blocks : 1, batch + (seam.npairs ? 1u : 0u));
    k_photo_residual<<<grid, kBlock, 0, st>>>(views_dev, batch, tex, tw, th, eps, tables, step_at, resid, stride,
                                              seam);
    const uint32_t gblocks = (max_runs + kBlock - 1) / kBlock;
    k_photo_gather<<<gblocks ? gblocks : 1, kBlock, 0, st>>>(batches_dev, nbatches, tw, texel_key_tiles(tw, th),
                                                             step_at, resid, grad);
    return hipGetLastError();
}

uint32_t tile_slots(uint32_t W, uint32_t H) { return ((W + 7u) / 8u) * ((H + 7u) / 8u) * 64u; }

uint32_t pixel_groups(uint32_t n) { return (n + 63u) / 64u; }

PhotoLossArgs photo_loss_args(const RefineView* views_dev, uint32_t batch, const LossParts& parts) {
    const uint64_t waves = static_cast<uint64_t>(batch) * parts.max_groups / kLossGroupsPerWave;
    return PhotoLossArgs{views_dev, batch, parts, static_cast<uint32_t>((waves + kBlock / 64u - 1u) / (kBlock / 64u))};
}

hipError_t loss_fold(const LossParts& parts, uint32_t first, uint32_t count, float* loss, hipStream_t st) {
    if (!count) return hipSuccess;
    k_loss_fold<<<count, kBlock, 0, st>>>(parts, first, loss);
    return hipGetLastError();
}

hipError_t seam_loss_fold(const float* part, uint32_t npairs, uint32_t steps, float* loss, hipStream_t st) {
    if (!steps || !npairs) return hipSuccess;
    k_seam_loss_fold<<<steps, kBlock, 0, st>>>(part, (npairs + 63u) / 64u, loss);
    return hipGetLastError();
}

hipError_t batch_taps(const BatchSlot* slots_dev, uint32_t nslots, uint32_t max_count, uint32_t stride, uint32_t tw,
                      uint32_t th, uint32_t* keys, uint32_t* vals, hipStream_t st) {
    if (!nslots || !max_count) return hipSuccess;
    k_batch_taps<<<dim3((max_count + kBlock - 1) / kBlock, nslots), kBlock, 0, st>>>(
        slots_dev, stride, tw, th, texel_key_tiles(tw, th), keys, vals);
    return hipGetLastError();
}

hipError_t batch_tap_weights(const uint32_t* vals, uint32_t E, const float4* const* slot_uvm, uint32_t stride,
                             uint32_t tw, uint32_t th, float* w, hipStream_t st) {
    if (!E) return hipSuccess;
    k_batch_tap_weights<<<(E + kBlock - 1) / kBlock, kBlock, 0, st>>>(vals, E, slot_uvm, stride, tw, th, w);
    return hipGetLastError();
}

hipError_t norm_table(const double* view_weight, const uint32_t* view_index, uint32_t steps, uint32_t batch,
                      float* norm, hipStream_t st) {
    if (!steps) return hipSuccess;
    k_norm_table<<<(steps + kBlock - 1) / kBlock, kBlock, 0, st>>>(view_weight, view_index, steps, batch, norm);
    return hipGetLastError();
}

size_t photo_group_temp_bytes(uint32_t n) {
    size_t a = 0, b = 0, c = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<int>(n));
    (void)hipcub::DeviceRunLengthEncode::Encode(nullptr, b, static_cast<const uint32_t*>(nullptr),
                                                static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                                static_cast<uint32_t*>(nullptr), static_cast<int>(n));
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), static_cast<int>(n));
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// Sorts n tap entries by texel (stable), then groups the first E (the taps of
// pixels with mask > 0): texels [runs], run lengths, offsets [runs] (the last
// offset is set by set_last_offset once runs is known on the device).
hipError_t photo_group_taps(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                            uint32_t n, int end_bit, uint32_t E, uint32_t* texels, uint32_t* runlen,
                            uint32_t* offsets, uint32_t* runs, void* temp, size_t temp_bytes, hipStream_t st) {
    size_t bytes = temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, vals_in, vals_out,
                                                      static_cast<int>(n), 0, end_bit, st);
    if (e != hipSuccess || E == 0) return e;
    bytes = temp_bytes;
    e = hipcub::DeviceRunLengthEncode::Encode(temp, bytes, keys_out, texels, runlen, runs, static_cast<int>(E), st);
    if (e != hipSuccess) return e;
    bytes = temp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(temp, bytes, runlen, offsets, static_cast<int>(E), st);
    if (e != hipSuccess) return e;
    k_set_last<<<1, 64, 0, st>>>(offsets, runs, E);
    return hipGetLastError();
}

hipError_t seam_step(const float* tex, uint32_t tw, uint32_t th, const float* pairs, uint32_t npairs, float eps,
                     float norm, float weight, StepIndex step_at, const TapGroups& taps, float4* resid,
                     float* grad, float* loss_part, bool with_residual, hipStream_t st) {
    if (!npairs) return hipSuccess;
    if (with_residual)
        k_seam_residual<<<(npairs + kBlock - 1) / kBlock, kBlock, 0, st>>>(tex, tw, th, pairs, npairs, eps, norm,
                                                                          weight, step_at, resid, loss_part);
    if (taps.ntexels)
        k_seam_gather<<<(taps.ntexels + kBlock - 1) / kBlock, kBlock, 0, st>>>(
            taps.texels, taps.offsets, taps.taps, taps.weights, taps.ntexels, resid, grad);
    return hipGetLastError();
}

hipError_t seam_taps(const float* pairs, uint32_t npairs, uint32_t tw, uint32_t th, uint32_t* keys, uint32_t* vals,
                     hipStream_t st) {
    if (!npairs) return hipSuccess;
    k_seam_taps<<<(npairs + kBlock - 1) / kBlock, kBlock, 0, st>>>(pairs, npairs, tw, th, keys, vals);
    return hipGetLastError();
}

hipError_t seam_tap_weights(const uint32_t* vals, uint32_t E, const float* pairs, uint32_t tw, uint32_t th, float* w,
                            hipStream_t st) {
    if (!E) return hipSuccess;
    k_seam_tap_weights<<<(E + kBlock - 1) / kBlock, kBlock, 0, st>>>(vals, E, pairs, tw, th, w);
    return hipGetLastError();
}

hipError_t adam_tv_step(const float* tex, float* tex_out, float* grad, float* m, float* v, uint32_t tw, uint32_t th,
                        const uint8_t* tv_mask, float tv_eps, float tv_scale, const float* lr_table,
                        const float* bc1_table, const float* bc2_table, float b1, float b2, float eps, float cmin,
                        float cmax, StepIndex step_at, const PhotoLossArgs& ploss, hipStream_t st) {
    uint32_t grid = (tw * th + kBlock - 1) / kBlock;
    if (grid > 8192) grid = 8192;
    k_adam_tv<<<grid + ploss.blocks, kBlock, 0, st>>>(tex, tex_out, grad, m, v, tw, th, tv_mask, tv_eps * tv_eps,
                                                      tv_scale, lr_table, bc1_table, bc2_table, b1, b2, eps, cmin,
                                                      cmax, step_at, ploss);
    return hipGetLastError();
}

hipError_t tv_mask(const int32_t* chart, uint32_t tw, uint32_t th, uint8_t* mask, hipStream_t st) {
    k_tv_mask<<<(tw * th + kBlock - 1) / kBlock, kBlock, 0, st>>>(chart, tw, th, mask);
    return hipGetLastError();
}

hipError_t tv_gradient(const float* tex, uint32_t tw, uint32_t th, const int32_t* chart, float eps, float scale,
                       float* grad, hipStream_t st) {
    k_tv_gradient<<<(tw * th + kBlock - 1) / kBlock, kBlock, 0, st>>>(tex, tw, th, chart, eps, scale, grad);
    return hipGetLastError();
}

hipError_t advance_step(uint32_t* step_counter, uint32_t count, hipStream_t st) {
    k_advance_step<<<1, 64, 0, st>>>(step_counter, count);
    return hipGetLastError();
}

hipError_t tv_pair_count(const int32_t* chart, uint32_t tw, uint32_t th, unsigned long long* out, hipStream_t st) {
    uint32_t grid = (tw * th + kBlock - 1) / kBlock;
    if (grid > 512) grid = 512;
    k_tv_pairs<<<grid ? grid : 1, kBlock, 0, st>>>(chart, tw, th, out);
    return hipGetLastError();
}

hipError_t adam_step(float* tex, float* grad, float* m, float* v, uint32_t n, const float* lr_table,
                     const float* bc1_table, const float* bc2_table, float b1, float b2, float eps, float cmin,
                     float cmax, StepIndex step_at, const PhotoLossArgs& ploss, hipStream_t st) {
    const uint32_t n4 = n / 4;
    uint32_t grid = (n4 + kBlock - 1) / kBlock;
    if (grid > 8192) grid = 8192;
    if (grid == 0) grid = 1;
    k_adam<<<grid + ploss.blocks, kBlock, 0, st>>>(tex, grad, m, v, n, lr_table, bc1_table, bc2_table, b1, b2, eps,
                                                   cmin, cmax, step_at, ploss);
    return hipGetLastError();
}

}  // namespace pdr
