// Rasterizer — include/photara_drender/rasterizer.hpp:12-64 of the reference,
// same signatures, outputs and exceptions; each call runs the libpdr kernels
// through the C ABI of pdr.h (host spans: staged, synchronous).
#pragma once

#include <cstdint>
#include <span>
#include <stdexcept>
#include <vector>

#include "pdr.h"
#include "photara_drender/context.hpp"
#include "photara_drender/types.hpp"

namespace photara_drender {

class Rasterizer {
public:
    explicit Rasterizer(Context& context) : context_(context.handle()) {}

    [[nodiscard]] RasterizeOutput forward(std::span<const float> clip_positions,
                                          std::span<const std::uint32_t> triangle_indices,
                                          const RasterizeOptions& options) {
        validate_mesh_shapes(clip_positions, triangle_indices);  // src/rasterizer.cpp:29-40
        RasterizeOutput result;
        result.width = options.width;
        result.height = options.height;
        const std::size_t pixels = static_cast<std::size_t>(options.width) * options.height;
        result.raster.resize(pixels * 4);
        if (options.output_barycentric_derivatives) result.barycentric_derivatives.resize(pixels * 4);
        float viewport[4];
        const float* requested = nullptr;
        if (options.viewport) {
            viewport[0] = options.viewport->x;
            viewport[1] = options.viewport->y;
            viewport[2] = options.viewport->width;
            viewport[3] = options.viewport->height;
            requested = viewport;
        }
        float resolved[4] = {0, 0, 0, 0};
        detail::check(pdr_rasterize_forward(
            context_, clip_positions.data(), static_cast<std::uint32_t>(clip_positions.size() / 4),
            triangle_indices.data(), static_cast<std::uint32_t>(triangle_indices.size() / 3), options.width,
            options.height, static_cast<std::uint32_t>(options.cull_mode),
            options.output_barycentric_derivatives ? 1 : 0, requested, result.raster.data(),
            options.output_barycentric_derivatives ? result.barycentric_derivatives.data() : nullptr, resolved,
            nullptr));
        result.viewport = Viewport{resolved[0], resolved[1], resolved[2], resolved[3]};
        return result;
    }

    [[nodiscard]] std::vector<float> backward(std::span<const float> clip_positions,
                                              std::span<const std::uint32_t> triangle_indices,
                                              const RasterizeOutput& forward_output,
                                              std::span<const float> grad_raster,
                                              std::span<const float> grad_barycentric_derivatives = {}) {
        validate_mesh_shapes(clip_positions, triangle_indices);
        const std::size_t values = static_cast<std::size_t>(forward_output.width) * forward_output.height * 4;
        if (forward_output.width == 0 || forward_output.height == 0 || forward_output.raster.size() != values)
            throw std::invalid_argument("forward_output has inconsistent dimensions");
        if (grad_raster.size() != values) throw std::invalid_argument("grad_raster must have shape [height, width, 4]");
        if (!grad_barycentric_derivatives.empty())
            throw std::invalid_argument(
                "Gradient propagation through barycentric derivatives is not implemented in the portable baseline "
                "backend");
        std::vector<float> result(clip_positions.size());
        const float vp[4] = {forward_output.viewport.x, forward_output.viewport.y, forward_output.viewport.width,
                             forward_output.viewport.height};
        detail::check(pdr_rasterize_backward(
            context_, clip_positions.data(), static_cast<std::uint32_t>(clip_positions.size() / 4),
            triangle_indices.data(), static_cast<std::uint32_t>(triangle_indices.size() / 3),
            forward_output.raster.data(), forward_output.width, forward_output.height, grad_raster.data(),
            forward_output.viewport.width > 0.0F ? vp : nullptr, result.data(), nullptr));
        return result;
    }

    [[nodiscard]] InterpolateOutput interpolate_forward(std::span<const float> vertex_attributes,
                                                        std::uint32_t attribute_count,
                                                        std::span<const std::uint32_t> triangle_indices,
                                                        const RasterizeOutput& raster_output) {
        if (attribute_count == 0 || vertex_attributes.empty() || vertex_attributes.size() % attribute_count != 0)
            throw std::invalid_argument("vertex_attributes must have shape [vertex_count, attribute_count]");
        if (triangle_indices.empty() || triangle_indices.size() % 3 != 0)
            throw std::invalid_argument("triangle_indices must have shape [triangle_count, 3]");
        check_raster(raster_output);
        InterpolateOutput result;
        result.width = raster_output.width;
        result.height = raster_output.height;
        result.attribute_count = attribute_count;
        result.values.resize(static_cast<std::size_t>(raster_output.width) * raster_output.height * attribute_count);
        detail::check(pdr_interpolate_forward(
            context_, vertex_attributes.data(), static_cast<std::uint32_t>(vertex_attributes.size() / attribute_count),
            attribute_count, triangle_indices.data(), static_cast<std::uint32_t>(triangle_indices.size() / 3),
            raster_output.raster.data(), raster_output.width, raster_output.height, result.values.data(), nullptr));
        return result;
    }

    [[nodiscard]] InterpolateGradients interpolate_backward(std::span<const float> vertex_attributes,
                                                            std::uint32_t attribute_count,
                                                            std::span<const std::uint32_t> triangle_indices,
                                                            const RasterizeOutput& raster_output,
                                                            std::span<const float> grad_interpolated) {
        if (attribute_count == 0 || vertex_attributes.empty() || vertex_attributes.size() % attribute_count != 0)
            throw std::invalid_argument("vertex_attributes must have shape [vertex_count, attribute_count]");
        if (triangle_indices.empty() || triangle_indices.size() % 3 != 0)
            throw std::invalid_argument("triangle_indices is incompatible with vertex_attributes");
        check_raster(raster_output);
        const std::size_t pixels = static_cast<std::size_t>(raster_output.width) * raster_output.height;
        if (grad_interpolated.size() != pixels * attribute_count)
            throw std::invalid_argument("grad_interpolated must have shape [height, width, attribute_count]");
        InterpolateGradients result;
        result.attributes.resize(vertex_attributes.size());
        result.raster.resize(pixels * 4);
        detail::check(pdr_interpolate_backward(
            context_, vertex_attributes.data(), static_cast<std::uint32_t>(vertex_attributes.size() / attribute_count),
            attribute_count, triangle_indices.data(), static_cast<std::uint32_t>(triangle_indices.size() / 3),
            raster_output.raster.data(), raster_output.width, raster_output.height, grad_interpolated.data(),
            result.attributes.data(), result.raster.data(), nullptr));
        return result;
    }

    [[nodiscard]] TextureOutput texture_forward(std::span<const float> texture, const TextureDesc& texture_desc,
                                                std::span<const float> uv, const RasterizeOutput& raster_output) {
        validate_texture(texture, texture_desc, uv, raster_output);
        TextureOutput result;
        result.width = raster_output.width;
        result.height = raster_output.height;
        result.channel_count = texture_desc.channel_count;
        result.values.resize(static_cast<std::size_t>(raster_output.width) * raster_output.height *
                             texture_desc.channel_count);
        detail::check(pdr_texture_forward(context_, texture.data(), texture_desc.width, texture_desc.height,
                                          texture_desc.channel_count,
                                          static_cast<std::uint32_t>(texture_desc.address_mode), uv.data(),
                                          raster_output.raster.data(), raster_output.width, raster_output.height,
                                          result.values.data(), nullptr));
        return result;
    }

    [[nodiscard]] TextureGradients texture_backward(std::span<const float> texture, const TextureDesc& texture_desc,
                                                    std::span<const float> uv, const RasterizeOutput& raster_output,
                                                    std::span<const float> grad_sampled,
                                                    bool compute_uv_gradient = true) {
        validate_texture(texture, texture_desc, uv, raster_output);
        const std::size_t pixels = static_cast<std::size_t>(raster_output.width) * raster_output.height;
        if (grad_sampled.size() != pixels * texture_desc.channel_count)
            throw std::invalid_argument("grad_sampled must have shape [image_height, image_width, channel_count]");
        TextureGradients result;
        result.texture.resize(texture.size());
        if (compute_uv_gradient) result.uv.resize(uv.size());
        detail::check(pdr_texture_backward(
            context_, texture.data(), texture_desc.width, texture_desc.height, texture_desc.channel_count,
            static_cast<std::uint32_t>(texture_desc.address_mode), uv.data(), raster_output.raster.data(),
            raster_output.width, raster_output.height, grad_sampled.data(), compute_uv_gradient ? 1 : 0,
            result.texture.data(), compute_uv_gradient ? result.uv.data() : nullptr, nullptr));
        return result;
    }

private:
    static void validate_mesh_shapes(std::span<const float> positions, std::span<const std::uint32_t> indices) {
        if (positions.empty() || positions.size() % 4 != 0)
            throw std::invalid_argument("clip_positions must have shape [vertex_count, 4]");
        if (indices.empty() || indices.size() % 3 != 0)
            throw std::invalid_argument("triangle_indices must have shape [triangle_count, 3]");
    }

    static void check_raster(const RasterizeOutput& r) {
        if (r.width == 0 || r.height == 0 ||
            r.raster.size() != static_cast<std::size_t>(r.width) * r.height * 4)
            throw std::invalid_argument("raster_output has inconsistent dimensions");
    }

    // src/rasterizer.cpp:260-289 (shape checks; range / overflow checks in libpdr)
    static void validate_texture(std::span<const float> texture, const TextureDesc& desc, std::span<const float> uv,
                                 const RasterizeOutput& r) {
        if (desc.width == 0 || desc.height == 0 || desc.channel_count == 0)
            throw std::invalid_argument("Texture width, height, and channel_count must be positive");
        if (desc.width > 0x3FFFFFFFu || desc.height > 0x3FFFFFFFu)
            throw std::out_of_range("Texture dimensions exceed portable addressing limits");
        if (texture.size() != static_cast<std::size_t>(desc.width) * desc.height * desc.channel_count)
            throw std::invalid_argument("texture must have shape [texture_height, texture_width, channel_count]");
        check_raster(r);
        if (uv.size() != static_cast<std::size_t>(r.width) * r.height * 2)
            throw std::invalid_argument("uv must have shape [image_height, image_width, 2]");
    }

    pdr_context context_;
};

} // namespace photara_drender
