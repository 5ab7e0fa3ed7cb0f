// TextureBaker::bake — include/photara_drender/texture_baker.hpp of the
// reference; bake() runs pdr_bake (GPU atlas raster, shadow maps, projection).
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "pdr.h"
#include "photara_drender/context.hpp"
#include "photara_drender/types.hpp"

namespace photara_drender {

class TextureBaker {
public:
    explicit TextureBaker(Context& context) : context_(context.handle()) {}

    // src/texture_baker.cpp:156-361
    [[nodiscard]] TextureBakeOutput bake(std::span<const float> positions, std::span<const float> normals,
                                         std::span<const float> uv, std::span<const std::uint32_t> triangle_indices,
                                         std::span<const ProjectionView> views,
                                         const TextureBakeOptions& options = {}) {
        if (positions.empty() || positions.size() % 3 != 0)
            throw std::invalid_argument("positions must have shape [vertex_count, 3]");
        const std::size_t vertex_count = positions.size() / 3;
        if (normals.size() != positions.size()) throw std::invalid_argument("normals must have shape [vertex_count, 3]");
        if (uv.size() != vertex_count * 2) throw std::invalid_argument("uv must have shape [vertex_count, 2]");
        if (triangle_indices.empty() || triangle_indices.size() % 3 != 0)
            throw std::invalid_argument("triangle_indices must have shape [triangle_count, 3]");
        std::vector<pdr_bake_view> raw(views.size());
        for (std::size_t k = 0; k < views.size(); ++k) {
            const ProjectionView& v = views[k];
            const std::size_t pixels = static_cast<std::size_t>(v.width) * v.height;
            if (v.width == 0 || v.height == 0 || (v.channel_count != 3 && v.channel_count != 4))
                throw std::invalid_argument("projection image must have nonzero dimensions and 3 or 4 channels");
            if (v.image.size() != pixels * v.channel_count)
                throw std::invalid_argument("projection image data size does not match its descriptor");
            if (!v.visibility_mask.empty() && v.visibility_mask.size() != pixels)
                throw std::invalid_argument("projection visibility_mask must have shape [height, width]");
            pdr_bake_view& r = raw[k];
            pdr_bake_view_default(&r);
            r.width = v.width;
            r.height = v.height;
            r.channels = v.channel_count;
            r.image = v.image.data();
            r.visibility_mask = v.visibility_mask.empty() ? nullptr : v.visibility_mask.data();
            for (int i = 0; i < 16; ++i) r.world_to_clip[i] = v.world_to_clip[i];
            for (int i = 0; i < 3; ++i) r.camera_position[i] = v.camera_position[i];
            r.has_viewport = v.viewport.has_value() ? 1 : 0;
            if (v.viewport) {
                r.viewport[0] = v.viewport->x;
                r.viewport[1] = v.viewport->y;
                r.viewport[2] = v.viewport->width;
                r.viewport[3] = v.viewport->height;
            }
            r.weight = v.weight;
            r.depth_bias = v.depth_bias;
            r.min_view_cosine = v.min_view_cosine;
        }
        pdr_bake_options o;
        pdr_bake_options_default(&o);
        o.width = options.width;
        o.height = options.height;
        o.blend_mode = static_cast<std::uint32_t>(options.blend_mode);
        o.visibility_mode = static_cast<std::uint32_t>(options.visibility_mode);
        o.pcf_radius = options.pcf_radius;
        o.ray_origin_bias = options.ray_origin_bias;
        o.allow_visibility_fallback = options.allow_visibility_fallback ? 1 : 0;
        o.mask_floor = options.mask_floor;
        o.softmax_scale = options.softmax_scale;
        const std::size_t P = static_cast<std::size_t>(options.width) * options.height;
        TextureBakeOutput out;
        out.width = options.width;
        out.height = options.height;
        out.color.resize(P * 4);
        out.confidence.resize(P);
        out.source_view.resize(P);
        out.valid_mask.resize(P);
        out.coverage_mask.resize(P);
        out.face_ids.resize(P);
        int used = 0;
        detail::check(pdr_bake(context_, positions.data(), normals.data(), uv.data(),
                               static_cast<std::uint32_t>(vertex_count), triangle_indices.data(),
                               static_cast<std::uint32_t>(triangle_indices.size() / 3), raw.data(),
                               static_cast<std::uint32_t>(raw.size()), &o, out.color.data(), out.confidence.data(),
                               out.source_view.data(), out.valid_mask.data(), out.coverage_mask.data(),
                               out.face_ids.data(), &used));
        out.used_ray_query = used != 0;
        return out;
    }

private:
    pdr_context context_;
};

// src/texture_padding.cpp:187-222
[[nodiscard]] inline float suggest_softmax_scale(std::span<const float> positions,
                                                 std::span<const float> camera_positions) {
    if (positions.size() < 3 || positions.size() % 3 != 0)
        throw std::invalid_argument("positions must have shape [vertex_count, 3]");
    if (camera_positions.size() < 3 || camera_positions.size() % 3 != 0)
        throw std::invalid_argument("camera_positions must have shape [camera_count, 3]");
    return pdr_suggest_softmax_scale(positions.data(), static_cast<std::uint32_t>(positions.size() / 3),
                                     camera_positions.data(), static_cast<std::uint32_t>(camera_positions.size() / 3));
}

}  // namespace photara_drender
