// TextureRefiner — include/photara_drender/texture_refiner.hpp:11-33 of the
// reference; refine() runs pdr_refine (GPU precompute, graph-captured Adam loop).
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "pdr.h"
#include "photara_drender/context.hpp"
#include "photara_drender/types.hpp"

namespace photara_drender {

class TextureRefiner {
public:
    explicit TextureRefiner(Context& context) : context_(context.handle()) {}

    [[nodiscard]] TextureRefineOutput refine(std::span<const float> initial_texture, std::span<const float> positions,
                                             std::span<const float> uv,
                                             std::span<const std::uint32_t> triangle_indices,
                                             std::span<const ProjectionView> views,
                                             std::span<const float> seam_uv_pairs = {},
                                             const TextureRefineOptions& options = {}) {
        const std::size_t vertex_count = positions.size() / 3;
        if (positions.empty() || positions.size() % 3 != 0 || uv.size() != vertex_count * 2 ||
            triangle_indices.empty() || triangle_indices.size() % 3 != 0)
            throw std::invalid_argument("texture refinement mesh arrays have inconsistent shapes");
        const std::size_t values = static_cast<std::size_t>(options.width) * options.height * 3;
        if (initial_texture.size() != values || seam_uv_pairs.size() % 4 != 0)
            throw std::invalid_argument("texture refinement options are invalid");
        std::vector<pdr_refine_view> raw(views.size());
        for (std::size_t k = 0; k < views.size(); ++k) {
            const ProjectionView& v = views[k];
            const std::size_t pixels = static_cast<std::size_t>(v.width) * v.height;
            if (v.width == 0 || v.height == 0 || (v.channel_count != 3 && v.channel_count != 4))
                throw std::invalid_argument("texture refinement images must contain RGB or RGBA data");
            if (v.image.size() != pixels * v.channel_count ||
                (!v.visibility_mask.empty() && v.visibility_mask.size() != pixels))
                throw std::invalid_argument("texture refinement image or mask has an invalid shape");
            pdr_refine_view& r = raw[k];
            r.width = v.width;
            r.height = v.height;
            r.channels = v.channel_count;
            r.image = v.image.data();
            r.visibility_mask = v.visibility_mask.empty() ? nullptr : v.visibility_mask.data();
            for (int i = 0; i < 16; ++i) r.world_to_clip[i] = v.world_to_clip[i];
            r.has_viewport = v.viewport.has_value() ? 1 : 0;
            if (v.viewport) {
                r.viewport[0] = v.viewport->x;
                r.viewport[1] = v.viewport->y;
                r.viewport[2] = v.viewport->width;
                r.viewport[3] = v.viewport->height;
            }
        }
        pdr_refine_options o;
        pdr_refine_options_default(&o);
        o.width = options.width;
        o.height = options.height;
        o.steps = options.steps;
        o.batch_size = options.batch_size;
        o.learning_rate = options.learning_rate;
        o.minimum_learning_rate = options.minimum_learning_rate;
        o.photometric_epsilon = options.photometric_epsilon;
        o.adam_beta1 = options.adam_beta1;
        o.adam_beta2 = options.adam_beta2;
        o.adam_epsilon = options.adam_epsilon;
        o.clamp_min = options.clamp_min;
        o.clamp_max = options.clamp_max;
        o.seam_polish_steps = options.seam_polish_steps;
        o.seam_learning_rate = options.seam_learning_rate;
        o.seam_epsilon = options.seam_epsilon;
        const std::uint32_t pairs = static_cast<std::uint32_t>(seam_uv_pairs.size() / 4);
        TextureRefineOutput out;
        out.width = options.width;
        out.height = options.height;
        out.color.resize(values);
        out.loss_history.resize(options.steps);
        out.seam_loss_history.resize(pairs ? options.seam_polish_steps : 0);
        double times[2] = {0.0, 0.0};
        detail::check(pdr_refine(context_, initial_texture.data(), positions.data(),
                                 static_cast<std::uint32_t>(vertex_count), uv.data(), triangle_indices.data(),
                                 static_cast<std::uint32_t>(triangle_indices.size() / 3), raw.data(),
                                 static_cast<std::uint32_t>(raw.size()),
                                 seam_uv_pairs.empty() ? nullptr : seam_uv_pairs.data(), pairs, &o, out.color.data(),
                                 out.loss_history.data(), out.seam_loss_history.data(), times));
        out.precompute_seconds = times[0];
        out.optimization_seconds = times[1];
        return out;
    }

private:
    pdr_context context_;
};

} // namespace photara_drender
