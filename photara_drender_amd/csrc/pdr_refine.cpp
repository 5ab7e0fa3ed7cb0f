// pdr_refine: TextureRefiner::refine (src/texture_refiner.cpp:174-454) on the GPU.
//
// Precompute per view (texture_refiner.cpp:206-254): transform -> bin + raster
// (no derivatives) -> interpolate uv -> target / mask / mask weight, all on the
// device.  The optimisation loop keeps the reference's schedule exactly
// (round-robin views, cosine learning rate, bias corrections computed on the
// host with the same float library calls, seam polish with a reset Adam) but
// reads every per-step scalar from device tables indexed by a device step
// counter, so a block of steps is captured once as a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <numeric>

#include "pdr_host.hpp"

using namespace pdrh;

namespace {

struct GraphExec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphExec() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

}  // namespace

extern "C" {

void pdr_refine_options_default(pdr_refine_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // include/photara_drender/types.hpp:267-283
    o->width = 1024;
    o->height = 1024;
    o->steps = 1000;
    o->batch_size = 4;
    o->learning_rate = 5e-3f;
    o->minimum_learning_rate = 2.5e-4f;
    o->photometric_epsilon = 1e-3f;
    o->adam_beta1 = 0.9f;
    o->adam_beta2 = 0.999f;
    o->adam_epsilon = 1e-8f;
    o->clamp_min = 0.0f;
    o->clamp_max = 1.0f;
    o->seam_polish_steps = 30;
    o->seam_learning_rate = 1e-3f;
    o->seam_epsilon = 1e-4f;
    o->tv_epsilon = 1e-4f;
    o->use_graph = 1;
}

pdr_status pdr_refine(pdr_context ctx, const float* initial_texture, const float* positions, uint32_t V,
                      const float* uv, const uint32_t* idx, uint32_t T, const pdr_refine_view* views, uint32_t nviews,
                      const float* seam_pairs, uint32_t npairs, const pdr_refine_options* options, float* color_out,
                      float* loss_out, float* seam_loss_out, double* times_out2) {
    return guarded([&] {
        require_ctx(ctx);
        pdr_refine_options o;
        if (options)
            o = *options;
        else
            pdr_refine_options_default(&o);
        // src/texture_refiner.cpp:182-204
        const bool idx_device = is_device_pointer(idx);
        if (!positions || V == 0 || !uv || !idx || T == 0)
            raise(PDR_INVALID_ARGUMENT, "texture refinement mesh arrays have inconsistent shapes");
        if (!idx_device)
            check_indices_host(idx, static_cast<size_t>(T) * 3, V, PDR_INVALID_ARGUMENT,
                               "texture refinement mesh arrays have inconsistent shapes");
        const uint64_t texels = static_cast<uint64_t>(o.width) * o.height;
        const uint64_t nvalues = texels * 3;
        if (o.width == 0 || o.height == 0 || !initial_texture || !views || nviews == 0 || o.steps == 0 ||
            o.batch_size == 0 || !std::isfinite(o.learning_rate) || o.learning_rate <= 0.0f ||
            !std::isfinite(o.minimum_learning_rate) || o.minimum_learning_rate < 0.0f ||
            o.minimum_learning_rate > o.learning_rate || !std::isfinite(o.photometric_epsilon) ||
            o.photometric_epsilon <= 0.0f || !(o.adam_beta1 > 0.0f && o.adam_beta1 < 1.0f) ||
            !(o.adam_beta2 > 0.0f && o.adam_beta2 < 1.0f) || !(o.clamp_min < o.clamp_max) ||
            (npairs && !seam_pairs) || !std::isfinite(o.seam_learning_rate) || o.seam_learning_rate <= 0.0f ||
            !std::isfinite(o.seam_epsilon) || o.seam_epsilon <= 0.0f)
            raise(PDR_INVALID_ARGUMENT, "texture refinement options are invalid");
        if (nvalues > 0xFFFFFFFFull) raise(PDR_OVERFLOW, "texture refinement atlas exceeds the shader element limit");
        for (uint32_t k = 0; k < nviews; ++k) {
            const pdr_refine_view& v = views[k];
            if (v.width == 0 || v.height == 0 || (v.channels != 3 && v.channels != 4))
                raise(PDR_INVALID_ARGUMENT, "texture refinement images must contain RGB or RGBA data");
            if (!v.image || static_cast<uint64_t>(v.width) * v.height > 0xFFFFFFFFull)
                raise(PDR_INVALID_ARGUMENT, "texture refinement image or mask has an invalid shape");
        }
        const float tv_eps = o.tv_epsilon > 0.0f ? o.tv_epsilon : 1e-4f;

        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(false, nullptr);
        DeviceBuffers bufs(ctx, "refine");
        if (idx_device)
            check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_INVALID_ARGUMENT,
                                 "texture refinement mesh arrays have inconsistent shapes");

        // ---------------------------------------------------------- precompute
        // device inputs may still be in flight on other streams of the caller
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const auto t0 = std::chrono::steady_clock::now();
        const float* dpos = device_view(bufs, positions, static_cast<size_t>(V) * 3, st);
        const float* duv = device_view(bufs, uv, static_cast<size_t>(V) * 2, st);
        const uint32_t* didx = device_view(bufs, idx, static_cast<size_t>(T) * 3, st);
        uint32_t max_pixels = 0;
        for (uint32_t k = 0; k < nviews; ++k) max_pixels = std::max(max_pixels, views[k].width * views[k].height);
        // views are processed in groups of up to kMaxViews, one launch per pass
        const uint32_t group = std::min<uint32_t>(nviews, pdr::kMaxViews);
        uint32_t max_slots = 0;  // tile-order slots (>= pixels)
        for (uint32_t k = 0; k < nviews; ++k)
            max_slots = std::max(max_slots, pdr::tile_slots(views[k].width, views[k].height));
        const size_t gp = static_cast<size_t>(group) * max_pixels, gs = static_cast<size_t>(group) * max_slots;
        float* clips = bufs.alloc<float>(static_cast<size_t>(group) * V * 4);
        float* mats = bufs.alloc<float>(static_cast<size_t>(group) * 16);
        float* rasters = bufs.alloc<float>(gp * 4);
        float* vuv = bufs.alloc<float>(gp * 2);
        float* target = bufs.alloc<float>(gp * 3);
        float* mask = bufs.alloc<float>(gp);
        const uint32_t nparts = pdr::refine_partial_count(max_pixels);
        double* partials = bufs.alloc<double>(static_cast<size_t>(group) * nparts);
        uint32_t* act_flags = bufs.alloc<uint32_t>(gs);  // tile order
        uint32_t* act_pos = bufs.alloc<uint32_t>(gs);
        uint32_t* rank_flags = bufs.alloc<uint32_t>(gp);  // raster order
        uint32_t* rank_pos = bufs.alloc<uint32_t>(gp);
        const uint32_t scan_tiles = pdr::active_scan_tiles(std::max(max_pixels, max_slots));
        uint32_t* scan_totals = bufs.alloc<uint32_t>(static_cast<size_t>(group) * 2 * scan_tiles);
        pdr::PrecomputeView* dpre = bufs.alloc<pdr::PrecomputeView>(group);
        double* view_weight = bufs.alloc<double>(nviews);
        uint32_t* view_counts = bufs.alloc<uint32_t>(nviews);  // active pixels
        std::vector<pdr::RefineView> rviews(nviews);
        std::vector<const uint32_t*> view_order(nviews);  // raster rank -> packed index, per view
        // every view's packed pixels in one allocation, sized for every pixel so
        // the per-view passes need no host readback
        auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
        size_t total_bytes = 0;
        uint32_t max_groups = 0;
        for (uint32_t k = 0; k < nviews; ++k) {
            const uint32_t P = views[k].width * views[k].height;
            total_bytes += 2 * up(P * sizeof(float4)) + 2 * up(P * sizeof(float));
            max_groups = std::max(max_groups, (pdr::pixel_groups(P) + 3u) & ~3u);
        }
        unsigned char* view_arena = bufs.alloc<unsigned char>(std::max<size_t>(total_bytes, 256));
        size_t arena_off = 0;
        std::vector<pdr::PrecomputeView> hpre(group);
        std::vector<float> hmats(static_cast<size_t>(group) * 16);
        // views the last bake on this context rasterised for its shadow maps (same
        // mesh, same position in the view list, size, viewport and matrix): their
        // keys are resolved instead of rasterising the views again
        std::vector<bool> cached(nviews, false);
        {
            const auto& c = ctx->view_keys;
            bool any = false;
            if (c.T == T && c.V == V)
                for (uint32_t k = 0; k < nviews && k < c.width.size(); ++k) {
                    const pdr_refine_view& v = views[k];
                    float vp[4];
                    resolve_viewport(v.width, v.height, v.has_viewport ? v.viewport : nullptr, vp);
                    cached[k] = c.width[k] == v.width && c.height[k] == v.height &&
                                std::memcmp(c.viewport[k].data(), vp, sizeof(vp)) == 0 &&
                                std::memcmp(c.matrix[k].data(), v.world_to_clip, 16 * sizeof(float)) == 0;
                    any |= cached[k];
                }
            if (any) {  // the mesh itself, compared on the device
                uint32_t* flag = ctx->get_t<uint32_t>("view_keys_flag", 1);
                hip_check(pdr::fill_u32(flag, 0u, 1, st), "fill");
                hip_check(pdr::mark_if_differ(reinterpret_cast<const uint32_t*>(dpos),
                                              ctx->get_t<uint32_t>("view_keys_pos", static_cast<size_t>(V) * 3),
                                              static_cast<size_t>(V) * 3, flag, st),
                          "mark_if_differ");
                hip_check(pdr::mark_if_differ(didx, ctx->get_t<uint32_t>("view_keys_idx", static_cast<size_t>(T) * 3),
                                              static_cast<size_t>(T) * 3, flag, st),
                          "mark_if_differ");
                hip_check(hipMemcpyAsync(ctx->pinned, flag, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
                sync(st);
                if (ctx->pinned[0]) std::fill(cached.begin(), cached.end(), false);
            }
        }
        for (uint32_t g0 = 0; g0 < nviews; g0 += group) {
            const uint32_t gn = std::min(group, nviews - g0);
            for (uint32_t b = 0; b < gn; ++b) std::memcpy(&hmats[b * 16], views[g0 + b].world_to_clip, 16 * 4);
            hip_check(hipMemcpyAsync(mats, hmats.data(), gn * 16 * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
            hip_check(pdr::transform_views(dpos, V, mats, gn, clips, st), "transform_views");
            pdr::ViewTable vt{};
            vt.n = gn;
            pdr::PixelTable pt{};
            pt.n = gn;
            for (uint32_t b = 0; b < gn; ++b) {
                const pdr_refine_view& v = views[g0 + b];
                const uint32_t P = v.width * v.height;
                float vp[4];
                resolve_viewport(v.width, v.height, v.has_viewport ? v.viewport : nullptr, vp);
                vt.v[b] = make_view(clips + static_cast<size_t>(b) * V * 4, v.width, v.height, vp);
                vt.v[b].raster = rasters + static_cast<size_t>(b) * max_pixels * 4;
                pdr::PixelView& pv = pt.v[b];
                pv.raster = vt.v[b].raster;
                pv.out0 = vuv + static_cast<size_t>(b) * max_pixels * 2;
                pv.width = v.width;
                pv.height = v.height;
                pv.tiles_x = (v.width + pdr::kGradTile - 1) / pdr::kGradTile;
                pv.tile_base = pt.total_tiles;
                pv.pixel_base = pt.total_pixels;
                pt.total_tiles += pv.tiles_x * ((v.height + pdr::kGradTile - 1) / pdr::kGradTile);
                pt.total_pixels += P;
                // the pixels with mask > 0 packed in 8x8-tile order (atlas locality
                // of a wave's taps), their raster ranks (the order in which the
                // reference's pixels add their taps), and the view's per-pixel loss
                // terms (zero where the mask is)
                pdr::PrecomputeView& pre = hpre[b];
                const size_t po = static_cast<size_t>(b) * max_pixels, so = static_cast<size_t>(b) * max_slots;
                pre.image = device_view(bufs, v.image, static_cast<size_t>(P) * v.channels, st);
                pre.vis = device_view(bufs, v.visibility_mask, P, st);
                pre.raster = vt.v[b].raster;
                pre.uv = pv.out0;
                pre.target = target + po * 3;
                pre.mask = mask + po;
                pre.partials = partials + static_cast<size_t>(b) * nparts;
                pre.weight = view_weight + g0 + b;
                pre.tflags = act_flags + so;
                pre.tpos = act_pos + so;
                pre.rflags = rank_flags + po;
                pre.rpos = rank_pos + po;
                pre.count = view_counts + g0 + b;
                pre.uvm = reinterpret_cast<float4*>(view_arena + arena_off);
                arena_off += up(static_cast<size_t>(P) * sizeof(float4));
                pre.rgb = reinterpret_cast<float4*>(view_arena + arena_off);
                arena_off += up(static_cast<size_t>(P) * sizeof(float4));
                pre.lterm = reinterpret_cast<float*>(view_arena + arena_off);
                arena_off += up(static_cast<size_t>(P) * sizeof(float));
                pre.order = reinterpret_cast<uint32_t*>(view_arena + arena_off);
                arena_off += up(static_cast<size_t>(P) * sizeof(uint32_t));
                pre.channels = v.channels;
                pre.width = v.width;
                pre.height = v.height;
                view_order[g0 + b] = pre.order;
                rviews[g0 + b] = pdr::RefineView{pre.uvm, pre.rgb, pre.lterm, 0u, P};  // active: set once read back
            }
            finalize_table(vt);
            bool group_cached = true;
            std::vector<size_t> key_offsets(gn);
            for (uint32_t b = 0; b < gn; ++b) {
                group_cached = group_cached && cached[g0 + b];
                if (cached[g0 + b]) key_offsets[b] = ctx->view_keys.offset[g0 + b];
            }
            if (group_cached)
                run_raster_resolve_keys(ctx, vt, didx, T, 0, ctx->get_t<unsigned long long>("view_keys", 1),
                                        key_offsets, st);
            else
                run_raster_forward(ctx, vt, didx, T, V, 0, 0, st);
            hip_check(pdr::interp_forward(pt, duv, V, 2, didx, T, st), "interp_forward");
            hip_check(hipMemcpyAsync(dpre, hpre.data(), gn * sizeof(pdr::PrecomputeView), hipMemcpyHostToDevice, st),
                      "hipMemcpyAsync");
            hip_check(pdr::precompute_views(dpre, gn, max_pixels, max_slots, scan_totals, st), "precompute_views");
            if (g0 + group < nviews) sync(st);  // the next group rewrites the group's scratch and view records
        }
        std::vector<uint32_t> host_counts(nviews);
        hip_check(hipMemcpyAsync(host_counts.data(), view_counts, host_counts.size() * 4, hipMemcpyDeviceToHost, st),
                  "hipMemcpyAsync");
        sync(st);
        uint32_t max_active = 0;
        for (uint32_t k = 0; k < nviews; ++k) {
            max_active = std::max(max_active, host_counts[k]);
            rviews[k].active = host_counts[k];
        }
        if (static_cast<uint64_t>(o.batch_size) * max_active * 4u > 0x7fffffffull)
            raise(PDR_OVERFLOW, "texture refinement batch exceeds 2^29 active pixels");

        // Gather form of the photometric gradient.  Step s samples views
        // (s * batch + b) % nviews, so the step batches recur with period
        // nviews / gcd(nviews, batch): each recurring batch's bilinear taps are
        // merged and grouped by texel once here (a stable radix sort by texel).
        uint32_t nbatches = nviews / std::gcd(nviews, o.batch_size);
        nbatches = std::max<uint32_t>(1, std::min(nbatches, o.steps));
        auto slot_view = [&](uint32_t c, uint32_t b) {
            return static_cast<uint32_t>((static_cast<uint64_t>(c) * o.batch_size + b) % nviews);
        };
        std::vector<uint32_t> batch_entries(nbatches, 0);
        uint32_t max_entries = 0;
        size_t tap_bytes = 0;
        for (uint32_t c = 0; c < nbatches; ++c) {
            for (uint32_t b = 0; b < o.batch_size; ++b) batch_entries[c] += host_counts[slot_view(c, b)] * 4u;
            max_entries = std::max(max_entries, batch_entries[c]);
            const size_t E = batch_entries[c];
            tap_bytes += 3 * up(E * 4) + up((E + 1) * 4);
        }
        int key_bits = 1;
        while (key_bits < 32 && (1ull << key_bits) < pdr::texel_key_range(o.width, o.height)) ++key_bits;
        int seam_key_bits = 1;
        while (seam_key_bits < 32 && (1ull << seam_key_bits) < static_cast<unsigned long long>(o.width) * o.height)
            ++seam_key_bits;
        uint32_t* tap_keys = bufs.alloc<uint32_t>(std::max<uint32_t>(max_entries, 1));
        uint32_t* tap_vals = bufs.alloc<uint32_t>(std::max<uint32_t>(max_entries, 1));
        uint32_t* tap_keys_sorted = bufs.alloc<uint32_t>(std::max<uint32_t>(max_entries, 1));
        uint32_t* group_runlen = bufs.alloc<uint32_t>(std::max<uint32_t>(max_entries, 1));
        const size_t group_temp_bytes = pdr::photo_group_temp_bytes(max_entries);
        void* group_temp = bufs.alloc<unsigned char>(std::max<size_t>(group_temp_bytes, 1));
        unsigned char* tap_arena = bufs.alloc<unsigned char>(std::max<size_t>(tap_bytes, 256));
        uint32_t* batch_runs = bufs.alloc<uint32_t>(nbatches);
        const float4** slot_uvm = bufs.alloc<const float4*>(static_cast<size_t>(nbatches) * o.batch_size);
        std::vector<const float4*> host_slot_uvm(static_cast<size_t>(nbatches) * o.batch_size);
        for (uint32_t c = 0; c < nbatches; ++c)
            for (uint32_t b = 0; b < o.batch_size; ++b)
                host_slot_uvm[static_cast<size_t>(c) * o.batch_size + b] = rviews[slot_view(c, b)].uvm;
        hip_check(hipMemcpyAsync(slot_uvm, host_slot_uvm.data(), host_slot_uvm.size() * sizeof(const float4*),
                                 hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
        hip_check(pdr::fill_u32(batch_runs, 0u, nbatches, st), "fill");
        std::vector<pdr::RefineBatchTaps> rbatches(nbatches);
        size_t tap_off = 0;
        auto carve = [&](size_t bytes) {
            void* p = tap_arena + tap_off;
            tap_off += up(bytes);
            return p;
        };
        // each batch's slots: view records, active counts and entry offsets
        std::vector<pdr::BatchSlot> hslots(static_cast<size_t>(nbatches) * o.batch_size);
        std::vector<uint32_t> slot_max(nbatches, 0);
        for (uint32_t c = 0; c < nbatches; ++c) {
            uint32_t off = 0;
            for (uint32_t b = 0; b < o.batch_size; ++b) {
                const uint32_t k = slot_view(c, b);
                hslots[static_cast<size_t>(c) * o.batch_size + b] =
                    pdr::BatchSlot{rviews[k].uvm, view_order[k], host_counts[k], off};
                off += host_counts[k] * 4u;
                slot_max[c] = std::max(slot_max[c], host_counts[k]);
            }
        }
        pdr::BatchSlot* dslots = bufs.alloc<pdr::BatchSlot>(hslots.size());
        hip_check(hipMemcpyAsync(dslots, hslots.data(), hslots.size() * sizeof(pdr::BatchSlot), hipMemcpyHostToDevice,
                                 st),
                  "hipMemcpyAsync");
        for (uint32_t c = 0; c < nbatches; ++c) {
            const uint32_t E = batch_entries[c];
            hip_check(pdr::batch_taps(dslots + static_cast<size_t>(c) * o.batch_size, o.batch_size, slot_max[c],
                                      max_active, o.width, o.height, tap_keys, tap_vals, st),
                      "batch_taps");
            uint32_t* texels_c = static_cast<uint32_t*>(carve(static_cast<size_t>(E) * 4));
            uint32_t* offsets_c = static_cast<uint32_t*>(carve((static_cast<size_t>(E) + 1) * 4));
            uint32_t* taps_c = static_cast<uint32_t*>(carve(static_cast<size_t>(E) * 4));
            float* weights_c = static_cast<float*>(carve(static_cast<size_t>(E) * 4));
            hip_check(pdr::photo_group_taps(tap_keys, tap_vals, tap_keys_sorted, taps_c, E, key_bits, E, texels_c,
                                            group_runlen, offsets_c, batch_runs + c, group_temp, group_temp_bytes, st),
                      "photo_group_taps");
            hip_check(pdr::batch_tap_weights(taps_c, E, slot_uvm + static_cast<size_t>(c) * o.batch_size, max_active,
                                             o.width, o.height, weights_c, st),
                      "batch_tap_weights");
            rbatches[c] = pdr::RefineBatchTaps{texels_c, offsets_c, taps_c, weights_c, batch_runs + c};
        }
        pdr::RefineBatchTaps* dbatches = bufs.alloc<pdr::RefineBatchTaps>(nbatches);
        hip_check(hipMemcpyAsync(dbatches, rbatches.data(), nbatches * sizeof(pdr::RefineBatchTaps),
                                 hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
        std::vector<uint32_t> host_runs(nbatches);
        hip_check(hipMemcpyAsync(host_runs.data(), batch_runs, nbatches * 4, hipMemcpyDeviceToHost, st),
                  "hipMemcpyAsync");
        sync(st);
        const uint32_t max_runs = *std::max_element(host_runs.begin(), host_runs.end());
        const auto t1 = std::chrono::steady_clock::now();

        // ---------------------------------------------------------- state + schedule
        const uint32_t n = static_cast<uint32_t>(nvalues);
        const uint32_t seam_steps = npairs ? o.seam_polish_steps : 0;
        float* tex = bufs.alloc<float>(n);
        float* grad = bufs.alloc<float>(n);
        float* m1 = bufs.alloc<float>(n);
        float* m2 = bufs.alloc<float>(n);
        float* loss = bufs.alloc<float>(o.steps + seam_steps);
        float4* resid = bufs.alloc<float4>(static_cast<size_t>(o.batch_size) * max_active);
        // The loss partials of `fold` steps stay in memory and are folded into
        // their step losses between blocks (outside the captured graphs):
        // `fold` is a multiple of the block length, up to 1,024 steps or
        // 256 MiB of partials.
        const uint32_t block = o.checkpoint_every ? o.checkpoint_every : std::min<uint32_t>(o.steps, 50);
        pdr::LossParts parts{nullptr, block, max_groups, static_cast<size_t>(o.batch_size) * max_groups};
        {
            const size_t block_bytes = parts.row * block * sizeof(float);
            const size_t by_mem = block_bytes ? (size_t{256} << 20) / block_bytes : 1;
            const size_t blocks = std::max<size_t>(1, std::min<size_t>({by_mem, 1024u / block, (o.steps + block - 1) / block}));
            parts.period = static_cast<uint32_t>(blocks * block);
        }
        parts.part = bufs.alloc<float>(std::max<size_t>(parts.row * parts.period, 4));
        float4* seam_resid = bufs.alloc<float4>(npairs);
        float* seam_part = bufs.alloc<float>(std::max<size_t>(static_cast<size_t>((npairs + 63u) / 64u) *
                                                                  (npairs ? o.seam_polish_steps : 0u),
                                                              1));
        uint32_t* counters = bufs.alloc<uint32_t>(4);  // step, done
        copy_any(tex, initial_texture, static_cast<size_t>(n) * 4, st);
        hip_check(pdr::fill_f32(grad, 0.0f, n, st), "fill");
        hip_check(pdr::fill_f32(m1, 0.0f, n, st), "fill");
        hip_check(pdr::fill_f32(m2, 0.0f, n, st), "fill");
        hip_check(pdr::fill_f32(loss, 0.0f, o.steps + seam_steps, st), "fill");
        hip_check(pdr::fill_u32(counters, 0u, 4, st), "fill");
        const float* dpairs = device_view(bufs, seam_pairs, static_cast<size_t>(npairs) * 4, st);
        // the seam pairs' taps grouped by texel (gather form of the seam gradient)
        pdr::TapGroups seam_groups{nullptr, nullptr, nullptr, nullptr, 0};
        if (npairs) {
            const uint32_t E = npairs * 8u;
            uint32_t* keys = bufs.alloc<uint32_t>(E);
            uint32_t* vals = bufs.alloc<uint32_t>(E);
            uint32_t* keys_sorted = bufs.alloc<uint32_t>(E);
            uint32_t* vals_sorted = bufs.alloc<uint32_t>(E);
            uint32_t* texels_tmp = bufs.alloc<uint32_t>(E);
            uint32_t* runlen = bufs.alloc<uint32_t>(E);
            uint32_t* offsets = bufs.alloc<uint32_t>(static_cast<size_t>(E) + 1);
            uint32_t* runs_dev = bufs.alloc<uint32_t>(1);
            const size_t temp_bytes = pdr::photo_group_temp_bytes(E);
            void* temp = bufs.alloc<unsigned char>(temp_bytes);
            hip_check(pdr::seam_taps(dpairs, npairs, o.width, o.height, keys, vals, st), "seam_taps");
            hip_check(pdr::photo_group_taps(keys, vals, keys_sorted, vals_sorted, E, seam_key_bits, E, texels_tmp,
                                            runlen, offsets, runs_dev, temp, temp_bytes, st),
                      "photo_group_taps");
            uint32_t runs = 0;
            hip_check(hipMemcpyAsync(&runs, runs_dev, sizeof(runs), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            sync(st);
            float* weights = bufs.alloc<float>(E);
            hip_check(pdr::seam_tap_weights(vals_sorted, E, dpairs, o.width, o.height, weights, st),
                      "seam_tap_weights");
            seam_groups = pdr::TapGroups{texels_tmp, offsets, vals_sorted, weights, runs};
        }

        // src/texture_refiner.cpp:315-356 (the normalisation from the mask sums on the device)
        std::vector<float> lr(o.steps), bc1(o.steps), bc2(o.steps);
        std::vector<uint32_t> vidx(static_cast<size_t>(o.steps) * o.batch_size);
        const float pi = std::acos(-1.0F);
        for (uint32_t s = 0; s < o.steps; ++s) {
            for (uint32_t b = 0; b < o.batch_size; ++b) {
                const size_t vi = (static_cast<size_t>(s) * o.batch_size + b) % nviews;
                vidx[static_cast<size_t>(s) * o.batch_size + b] = static_cast<uint32_t>(vi);
            }
            const float phase = static_cast<float>(s) / static_cast<float>(o.steps);
            lr[s] = o.minimum_learning_rate +
                    0.5F * (o.learning_rate - o.minimum_learning_rate) * (1.0F + std::cos(pi * phase));
            bc1[s] = 1.0F - std::pow(o.adam_beta1, static_cast<float>(s + 1));
            bc2[s] = 1.0F - std::pow(o.adam_beta2, static_cast<float>(s + 1));
        }
        float* d_lr = bufs.alloc<float>(o.steps);
        float* d_bc1 = bufs.alloc<float>(o.steps);
        float* d_bc2 = bufs.alloc<float>(o.steps);
        float* d_norm = bufs.alloc<float>(o.steps);
        uint32_t* d_vidx = bufs.alloc<uint32_t>(vidx.size());
        copy_any(d_lr, lr.data(), lr.size() * 4, st);
        copy_any(d_bc1, bc1.data(), bc1.size() * 4, st);
        copy_any(d_bc2, bc2.data(), bc2.size() * 4, st);
        copy_any(d_vidx, vidx.data(), vidx.size() * 4, st);
        hip_check(pdr::norm_table(view_weight, d_vidx, o.steps, o.batch_size, d_norm, st), "norm_table");
        // each step slot's view record itself (the residual pass then needs no view-index load)
        std::vector<pdr::RefineView> step_views(vidx.size());
        for (size_t q = 0; q < vidx.size(); ++q) step_views[q] = rviews[vidx[q]];
        pdr::RefineView* dviews = bufs.alloc<pdr::RefineView>(step_views.size());
        copy_any(dviews, step_views.data(), step_views.size() * sizeof(pdr::RefineView), st);
        pdr::RefineStepTables tables{d_lr, d_bc1, d_bc2, d_norm, d_vidx};

        // extended objective (python/photara_drender/atlas.py:84-144)
        const int32_t* dchart = nullptr;
        uint8_t* dtv_mask = nullptr;
        float tv_scale = 0.0f;
        if (o.tv_weight != 0.0f) {
            dchart = device_view(bufs, o.chart_ids, static_cast<size_t>(texels), st);
            dtv_mask = bufs.alloc<uint8_t>(static_cast<size_t>(texels));
            hip_check(pdr::tv_mask(dchart, o.width, o.height, dtv_mask, st), "tv_mask");
            unsigned long long* cnt = bufs.alloc<unsigned long long>(1);
            hip_check(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st), "hipMemsetAsync");
            hip_check(pdr::tv_pair_count(dchart, o.width, o.height, cnt, st), "tv_pair_count");
            unsigned long long pairs = 0;
            hip_check(hipMemcpyAsync(&pairs, cnt, sizeof(pairs), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            sync(st);
            tv_scale = o.tv_weight / (static_cast<float>(std::max<unsigned long long>(pairs, 1)) * 3.0f);
        }
        const float seam_norm = npairs ? 1.0F / static_cast<float>(npairs * 3) : 0.0f;
        sync(st);

        // ---------------------------------------------------------- optimisation
        // With the TV term, Adam reads each texel's neighbours from the pre-step
        // atlas and writes the new one to the other buffer: `cur` is the buffer
        // holding the current atlas (host-tracked; a captured block of steps
        // starts from a fixed parity).
        const bool fused_tv = o.tv_weight != 0.0f;
        float* texbuf[2] = {tex, fused_tv ? bufs.alloc<float>(n) : tex};
        const auto t2 = std::chrono::steady_clock::now();
        const bool joint_seam = o.joint_seam_weight != 0.0f && npairs;
        const pdr::PhotoLossArgs ploss = pdr::photo_loss_args(dviews, o.batch_size, parts);
        // k: the step's position in a captured block (its kernels add it to the
        // device step counter, which moves on once per block); outside graphs
        // k = 0 and the counter moves on after every step
        auto one_step = [&](hipStream_t s, uint32_t c, uint32_t k, bool advance) {
            const pdr::StepIndex at{counters, k};
            pdr::SeamResidualArgs seam_res;  // the joint seam residual rides in the photometric residual launch
            if (joint_seam)
                seam_res = pdr::SeamResidualArgs{dpairs, npairs, o.seam_epsilon, seam_norm, o.joint_seam_weight,
                                                 seam_resid};
            hip_check(pdr::photometric_step(dviews, o.batch_size, max_active, dbatches, nbatches, max_runs, texbuf[c],
                                            o.width, o.height, o.photometric_epsilon, tables, at, resid,
                                            max_active, grad, seam_res, s),
                      "photometric_step");
            if (joint_seam)
                hip_check(pdr::seam_step(texbuf[c], o.width, o.height, dpairs, npairs, o.seam_epsilon, seam_norm,
                                         o.joint_seam_weight, at, seam_groups, seam_resid, grad, nullptr,
                                         false, s),
                          "seam_step");
            if (fused_tv)
                hip_check(pdr::adam_tv_step(texbuf[c], texbuf[c ^ 1u], grad, m1, m2, o.width, o.height, dtv_mask, tv_eps,
                                            tv_scale, d_lr, d_bc1, d_bc2, o.adam_beta1, o.adam_beta2, o.adam_epsilon,
                                            o.clamp_min, o.clamp_max, at, ploss, s),
                          "adam_tv_step");
            else
                hip_check(pdr::adam_step(texbuf[c], grad, m1, m2, n, d_lr, d_bc1, d_bc2, o.adam_beta1, o.adam_beta2,
                                         o.adam_epsilon, o.clamp_min, o.clamp_max, at, ploss, s),
                          "adam_step");
            if (advance) hip_check(pdr::advance_step(counters, 1u, s), "advance_step");
            return fused_tv ? c ^ 1u : c;
        };
        auto capture = [&](uint32_t count, uint32_t c, GraphExec& g) {
            hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            for (uint32_t k = 0; k < count; ++k) c = one_step(st, c, k, false);
            hip_check(pdr::advance_step(counters, count, st), "advance_step");
            hip_check(hipStreamEndCapture(st, &g.graph), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        };
        const uint32_t run = o.stop_after && o.stop_after < o.steps ? o.stop_after : o.steps;  // steps executed
        const uint32_t nblocks = run / block, rest = run % block;
        const uint32_t flips = fused_tv ? 1u : 0u;
        GraphExec full[2], tail;
        if (o.use_graph) {
            if (nblocks) capture(block, 0u, full[0]);
            if (nblocks > 1 && (block & flips)) capture(block, 1u, full[1]);
            if (rest) capture(rest, (nblocks * block) & flips, tail);
        }
        uint32_t done = 0, ckpt = 0, cur = 0, folded = 0;
        const size_t atlas_bytes = static_cast<size_t>(n) * 4;
        for (uint32_t b = 0; b <= nblocks; ++b) {
            const uint32_t count = b < nblocks ? block : rest;
            if (!count) continue;
            if (o.use_graph) {
                hip_check(hipGraphLaunch(b < nblocks ? full[cur].exec : tail.exec, st), "hipGraphLaunch");
                cur ^= count & flips;
            } else {
                for (uint32_t k = 0; k < count; ++k) cur = one_step(st, cur, 0u, true);
            }
            done += count;
            if (done - folded == parts.period || done == run) {
                hip_check(pdr::loss_fold(parts, folded, done - folded, loss, st), "loss_fold");
                folded = done;
            }
            if (o.checkpoint_every && o.checkpoints && done % o.checkpoint_every == 0)
                copy_any(reinterpret_cast<char*>(o.checkpoints) + static_cast<size_t>(ckpt++) * atlas_bytes,
                         texbuf[cur], atlas_bytes, st);
        }
        tex = texbuf[cur];
        if (seam_steps) {
            // texture_refiner.cpp:371-415: reset Adam, constant seam learning rate
            std::vector<float> slr(seam_steps, o.seam_learning_rate), sb1(seam_steps), sb2(seam_steps);
            for (uint32_t k = 0; k < seam_steps; ++k) {
                sb1[k] = 1.0F - std::pow(o.adam_beta1, static_cast<float>(k + 1));
                sb2[k] = 1.0F - std::pow(o.adam_beta2, static_cast<float>(k + 1));
            }
            float* d_slr = bufs.alloc<float>(seam_steps);
            float* d_sb1 = bufs.alloc<float>(seam_steps);
            float* d_sb2 = bufs.alloc<float>(seam_steps);
            copy_any(d_slr, slr.data(), seam_steps * 4, st);
            copy_any(d_sb1, sb1.data(), seam_steps * 4, st);
            copy_any(d_sb2, sb2.data(), seam_steps * 4, st);
            hip_check(pdr::fill_f32(grad, 0.0f, n, st), "fill");
            hip_check(pdr::fill_f32(m1, 0.0f, n, st), "fill");
            hip_check(pdr::fill_f32(m2, 0.0f, n, st), "fill");
            hip_check(pdr::fill_u32(counters, 0u, 2, st), "fill");
            for (uint32_t k = 0; k < seam_steps; ++k) {
                hip_check(pdr::seam_step(tex, o.width, o.height, dpairs, npairs, o.seam_epsilon, seam_norm, 1.0f,
                                         pdr::StepIndex{counters, 0u}, seam_groups, seam_resid, grad, seam_part, true,
                                         st),
                          "seam_step");
                hip_check(pdr::adam_step(tex, grad, m1, m2, n, d_slr, d_sb1, d_sb2, o.adam_beta1, o.adam_beta2,
                                         o.adam_epsilon, o.clamp_min, o.clamp_max, pdr::StepIndex{counters, 0u},
                                         pdr::PhotoLossArgs{}, st),
                          "adam_step");
                hip_check(pdr::advance_step(counters, 1u, st), "advance_step");
            }
            hip_check(pdr::seam_loss_fold(seam_part, npairs, seam_steps, loss + o.steps, st), "seam_loss_fold");
        }
        sync(st);
        const auto t3 = std::chrono::steady_clock::now();
        copy_any(color_out, tex, atlas_bytes, st);
        copy_any(loss_out, loss, static_cast<size_t>(o.steps) * 4, st);
        if (seam_steps) copy_any(seam_loss_out, loss + o.steps, static_cast<size_t>(seam_steps) * 4, st);
        sync(st);
        if (times_out2) {
            times_out2[0] = std::chrono::duration<double>(t1 - t0).count();
            times_out2[1] = std::chrono::duration<double>(t3 - t2).count();
        }
    });
}

// Self-test hook outside the public header: the chart-TV gradient that the
// refiner's fused Adam step adds (scale * d/dt per texel), on device arrays,
// synchronously on the default stream.  Returns a hipError_t.
int pdr_testing_tv_gradient(const float* tex, uint32_t tw, uint32_t th, const int32_t* chart, float eps, float scale,
                            float* grad) {
    hipError_t e = pdr::tv_gradient(tex, tw, th, chart, eps, scale, grad, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return static_cast<int>(e);
}

}  // extern "C"
