// pdr_pad_texture_atlas: pad_texture_atlas (src/texture_padding.cpp:226-351)
// with its TexturePaddingOptions / TexturePaddingStats
// (include/photara_drender/texture_padding.hpp:33-57); kernels in padding.hip.
#include <atomic>
#include <chrono>

#include "pdr_host.hpp"

using namespace pdrh;

namespace {
// Test hooks: a poll bound for the parallel band sweep (0: the kernel's own
// default) and the number of times the ordered fallback ran.
std::atomic<uint32_t> g_pad_poll_limit{0};
std::atomic<uint64_t> g_pad_fallbacks{0};
}  // namespace

extern "C" {

void pdr_padding_options_default(pdr_padding_options* o) {
    if (!o) return;
    // include/photara_drender/texture_padding.hpp:33-44
    o->margin = 4;
    o->fill_unobserved = 1;
    o->maximum_fill_distance = 0;
    o->valid_threshold = 0.5f;
}

pdr_status pdr_pad_texture_atlas(pdr_context ctx, float* color, uint32_t width, uint32_t height, uint32_t channels,
                                 const float* valid_mask, const float* coverage_mask, float* filled_mask,
                                 const pdr_padding_options* options, pdr_padding_stats* stats_out, void* stream) {
    return guarded([&] {
        const auto start = std::chrono::steady_clock::now();
        require_ctx(ctx);
        pdr_padding_options o;
        if (options)
            o = *options;
        else
            pdr_padding_options_default(&o);
        // src/texture_padding.cpp:228-252
        if (width == 0 || height == 0) raise(PDR_INVALID_ARGUMENT, "cannot pad an empty atlas");
        if (width > 16384u || height > 16384u)
            raise(PDR_INVALID_ARGUMENT, "atlas dimensions exceed the supported maximum");
        if (channels != 3 && channels != 4) raise(PDR_INVALID_ARGUMENT, "atlas color must have 3 or 4 channels");
        require_ptr(color, "color");
        if (!valid_mask) raise(PDR_INVALID_ARGUMENT, "atlas valid_mask must match the atlas dimensions");
        if (!std::isfinite(o.valid_threshold) || o.valid_threshold <= 0.0f || o.valid_threshold > 1.0f)
            raise(PDR_INVALID_ARGUMENT, "valid_threshold must be within (0, 1]");
        const uint32_t W = width, H = height, C = channels;
        const size_t P = static_cast<size_t>(W) * H;
        const bool device = classify({color, valid_mask, coverage_mask, filled_mask});
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        float* dcolor = color;
        const float *dvalid = valid_mask, *dcov = coverage_mask;
        float* dfilled = filled_mask;
        if (!device) {
            dcolor = const_cast<float*>(stage_in(ctx, "pad_color", static_cast<const float*>(color), P * C, st));
            dvalid = stage_in(ctx, "pad_valid", valid_mask, P, st);
            dcov = stage_in(ctx, "pad_coverage", coverage_mask, P, st);
            dfilled = const_cast<float*>(stage_in(ctx, "pad_filled", static_cast<const float*>(filled_mask), P, st));
        }
        uint8_t* source = ctx->get_t<uint8_t>("pad_source", P);
        uint32_t* field = ctx->get_t<uint32_t>("pad_field", P + 1);  // + the sweeps' discard slot
        unsigned long long* counters = ctx->get_t<unsigned long long>("pad_counters", 5);
        hip_check(pdr::fill_u64(counters, 0ull, 5, st), "fill_u64");
        ctx->timed("pad_prepare", st, [&] {
            hip_check(pdr::pad_prepare(static_cast<uint32_t>(P), dvalid, dcov, dfilled, o.valid_threshold, source,
                                       counters, st),
                      "pad_prepare");
        });
        unsigned long long c[5];
        hip_check(hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        sync(st);
        if (c[0] != 0 || c[4] != 0) {  // :277-281: nothing to extrapolate from otherwise
            // 8SSEDT field: rows in bands of 64 on separate workgroups, each waiting for the band above
            const uint32_t bands = pdr::pad_band_count(H);
            unsigned long long* edge = ctx->get_t<unsigned long long>("pad_band_edge", static_cast<size_t>(bands) * W);
            uint32_t* status = ctx->get_t<uint32_t>("pad_band_status", 1);
            ctx->timed("pad_field", st, [&] {
                hip_check(pdr::pad_field_bands(W, H, source, field, edge, status, g_pad_poll_limit.load(), st),
                          "pad_field_bands");
            });
            uint32_t gave_up = 0;
            hip_check(hipMemcpyAsync(&gave_up, status, sizeof(gave_up), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            sync(st);
            if (gave_up) {
                // a band waited too long (the GPU time-sliced with other work, or bands
                // dispatched out of order): the same sweeps one band per launch, in order
                g_pad_fallbacks.fetch_add(1);
                ctx->timed("pad_field_ordered", st, [&] {
                    hip_check(pdr::pad_field_bands_ordered(W, H, source, field, edge, status, st),
                              "pad_field_bands_ordered");
                });
                hip_check(hipMemcpyAsync(&gave_up, status, sizeof(gave_up), hipMemcpyDeviceToHost, st),
                          "hipMemcpyAsync");
                sync(st);
                if (gave_up) raise(PDR_RUNTIME, "texture padding: a row band found the band above incomplete");
            }
            // A far sentinel offset can land inside atlases wider or taller than it: read colours from a copy then.
            const float* src_color = dcolor;
            if (W > 12000u || H > 12000u) {
                float* copy = ctx->get_t<float>("pad_color_copy", P * C);
                hip_check(hipMemcpyAsync(copy, dcolor, P * C * sizeof(float), hipMemcpyDeviceToDevice, st),
                          "hipMemcpyAsync");
                src_color = copy;
            }
            ctx->timed("pad_fill", st, [&] {
                hip_check(pdr::pad_fill(W, H, C, field, source, dcov, o.margin, o.fill_unobserved ? 1 : 0,
                                        o.maximum_fill_distance, src_color, dcolor, dfilled, counters, st),
                          "pad_fill");
            });
            hip_check(hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        }
        if (!device) {
            copy_out(color, dcolor, P * C * sizeof(float), st);
            if (filled_mask) copy_out(filled_mask, dfilled, P * sizeof(float), st);
        }
        sync(st);
        if (stats_out) {
            stats_out->valid_texels = c[0];
            stats_out->gutter_texels = c[1];
            stats_out->unobserved_texels = c[2];
            stats_out->remaining_unobserved = c[3];
            stats_out->filled_texels = c[4];
            stats_out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        }
    });
}

// Test hooks (outside the public header).
void pdr_testing_set_pad_poll_limit(uint32_t polls) { g_pad_poll_limit.store(polls); }
uint64_t pdr_testing_pad_fallbacks(void) { return g_pad_fallbacks.load(); }

}  // extern "C"
