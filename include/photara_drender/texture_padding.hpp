// pad_texture_atlas — include/photara_drender/texture_padding.hpp:33-64 of the
// reference (TexturePaddingOptions, TexturePaddingStats, pad_texture_atlas),
// run on the GPU through pdr_pad_texture_atlas.  The reference's function
// takes no context; the overload without one uses a process-wide context on
// device 0.
#pragma once

#include <cstdint>
#include <stdexcept>

#include "pdr.h"
#include "photara_drender/context.hpp"
#include "photara_drender/types.hpp"

namespace photara_drender {

struct TexturePaddingOptions {
    std::uint32_t margin = 4;
    bool fill_unobserved = true;
    std::uint32_t maximum_fill_distance = 0;
    float valid_threshold = 0.5F;
};

struct TexturePaddingStats {
    std::uint64_t valid_texels = 0;
    std::uint64_t gutter_texels = 0;
    std::uint64_t unobserved_texels = 0;
    std::uint64_t remaining_unobserved = 0;
    std::uint64_t filled_texels = 0;
    double seconds = 0.0;
};

// Extends output.color in place and records guard / unobserved fills in
// output.filled_mask; output.valid_mask is never widened.
[[nodiscard]] inline TexturePaddingStats pad_texture_atlas(Context& context, TextureBakeOutput& output,
                                                           const TexturePaddingOptions& options = {}) {
    const std::size_t pixels = static_cast<std::size_t>(output.width) * output.height;
    if (output.width == 0 || output.height == 0) throw std::invalid_argument("cannot pad an empty atlas");
    if (output.color.size() % pixels != 0) throw std::invalid_argument("atlas color size does not match its dimensions");
    if (output.valid_mask.size() != pixels)
        throw std::invalid_argument("atlas valid_mask must match the atlas dimensions");
    const auto channels = static_cast<std::uint32_t>(output.color.size() / pixels);
    if (output.filled_mask.size() != pixels) output.filled_mask.assign(pixels, 0.0F);
    const float* coverage = output.coverage_mask.size() == pixels ? output.coverage_mask.data() : nullptr;
    pdr_padding_options o{options.margin, options.fill_unobserved ? 1 : 0, options.maximum_fill_distance,
                          options.valid_threshold};
    pdr_padding_stats s{};
    detail::check(pdr_pad_texture_atlas(context.handle(), output.color.data(), output.width, output.height, channels,
                                        output.valid_mask.data(), coverage, output.filled_mask.data(), &o, &s,
                                        nullptr));
    return TexturePaddingStats{s.valid_texels, s.gutter_texels, s.unobserved_texels, s.remaining_unobserved,
                               s.filled_texels, s.seconds};
}

[[nodiscard]] inline TexturePaddingStats pad_texture_atlas(TextureBakeOutput& output,
                                                           const TexturePaddingOptions& options = {}) {
    static Context context;
    return pad_texture_atlas(context, output, options);
}

}  // namespace photara_drender
