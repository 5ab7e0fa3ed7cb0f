// _photara_drender: pybind11 module with the Python signatures of the
// reference's nanobind module (python/bindings.cpp:414-646 PythonRasterizer,
// :289-404 TextureBaker.refine_texture, :650-837 module definition) for the
// raster / interpolate / texture / refine path.  Arrays are C-contiguous numpy
// float32 / uint32; resolution is (height, width); C++ exceptions map to
// ValueError / IndexError / OverflowError / RuntimeError as in the reference.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "photara_drender/photara_drender.hpp"

namespace py = pybind11;
using namespace photara_drender;

namespace {

using FloatInput = py::array_t<float, py::array::c_style | py::array::forcecast>;
using U32Input = py::array_t<std::uint32_t, py::array::c_style | py::array::forcecast>;

template <class T>
std::span<const T> as_span(const py::array_t<T, py::array::c_style | py::array::forcecast>& a) {
    return {a.data(), static_cast<std::size_t>(a.size())};
}

void validate_positions(const FloatInput& a) {
    if (a.ndim() != 2 || a.shape(1) != 4 || a.shape(0) <= 0)
        throw std::invalid_argument("positions must have shape [vertex_count, 4]");
}

void validate_indices(const U32Input& a) {
    if (a.ndim() != 2 || a.shape(1) != 3 || a.shape(0) <= 0)
        throw std::invalid_argument("indices must have shape [triangle_count, 3]");
}

void validate_image(const FloatInput& a, std::uint32_t width, std::uint32_t height, const char* name) {
    if (a.ndim() != 3 || a.shape(0) != height || a.shape(1) != width || a.shape(2) != 4)
        throw std::invalid_argument(std::string(name) + " must have shape [height, width, 4]");
}

CullMode parse_cull_mode(const std::string& value) {
    if (value == "none") return CullMode::none;
    if (value == "back") return CullMode::back;
    if (value == "front") return CullMode::front;
    throw std::invalid_argument("cull_mode must be 'none', 'back', or 'front'");
}

AddressMode parse_address_mode(const std::string& value) {
    if (value == "clamp") return AddressMode::clamp;
    if (value == "wrap") return AddressMode::wrap;
    if (value == "mirror") return AddressMode::mirror;
    throw std::invalid_argument("address_mode must be 'clamp', 'wrap', or 'mirror'");
}

py::array_t<float> to_array(std::vector<float>&& values, std::vector<py::ssize_t> shape) {
    auto* data = new std::vector<float>(std::move(values));
    py::capsule owner(data, [](void* p) { delete static_cast<std::vector<float>*>(p); });
    return py::array_t<float>(shape, data->data(), owner);
}

RasterizeOutput raster_from_array(const FloatInput& raster) {
    if (raster.ndim() != 3 || raster.shape(2) != 4)
        throw std::invalid_argument("raster must have shape [height, width, 4]");
    RasterizeOutput out;
    out.height = static_cast<std::uint32_t>(raster.shape(0));
    out.width = static_cast<std::uint32_t>(raster.shape(1));
    out.raster.assign(raster.data(), raster.data() + raster.size());
    return out;
}

class PythonRasterizer {
public:
    PythonRasterizer(std::uint32_t device_index, bool enable_validation)
        : context_(ContextOptions{device_index, enable_validation}), rasterizer_(context_) {}

    py::tuple forward(const FloatInput& positions, const U32Input& indices,
                      const std::pair<std::uint32_t, std::uint32_t>& resolution, const std::string& cull_mode,
                      bool output_barycentric_derivatives, const std::optional<std::array<float, 4>>& viewport) {
        validate_positions(positions);
        validate_indices(indices);
        RasterizeOptions options;
        options.height = resolution.first;
        options.width = resolution.second;
        options.cull_mode = parse_cull_mode(cull_mode);
        options.output_barycentric_derivatives = output_barycentric_derivatives;
        if (viewport) options.viewport = Viewport{(*viewport)[0], (*viewport)[1], (*viewport)[2], (*viewport)[3]};
        RasterizeOutput output;
        {
            py::gil_scoped_release release;
            output = rasterizer_.forward(as_span(positions), as_span(indices), options);
        }
        const py::ssize_t h = options.height, w = options.width;
        auto raster = to_array(std::move(output.raster), {h, w, 4});
        py::array_t<float> derivatives = output_barycentric_derivatives
                                             ? to_array(std::move(output.barycentric_derivatives), {h, w, 4})
                                             : to_array(std::vector<float>{}, {0, 0, 4});
        return py::make_tuple(raster, derivatives);
    }

    py::array_t<float> backward(const FloatInput& positions, const U32Input& indices, const FloatInput& raster,
                                const FloatInput& grad_raster, const std::optional<std::array<float, 4>>& viewport) {
        validate_positions(positions);
        validate_indices(indices);
        RasterizeOutput forward_output = raster_from_array(raster);
        validate_image(grad_raster, forward_output.width, forward_output.height, "grad_raster");
        if (viewport)
            forward_output.viewport = Viewport{(*viewport)[0], (*viewport)[1], (*viewport)[2], (*viewport)[3]};
        std::vector<float> gradient;
        {
            py::gil_scoped_release release;
            gradient = rasterizer_.backward(as_span(positions), as_span(indices), forward_output, as_span(grad_raster));
        }
        return to_array(std::move(gradient), {positions.shape(0), 4});
    }

    py::array_t<float> interpolate_forward(const FloatInput& attributes, const U32Input& indices,
                                           const FloatInput& raster) {
        if (attributes.ndim() != 2 || attributes.shape(0) <= 0 || attributes.shape(1) <= 0)
            throw std::invalid_argument("vertex_attributes must have shape [vertex_count, attribute_count]");
        validate_indices(indices);
        const RasterizeOutput r = raster_from_array(raster);
        InterpolateOutput output;
        {
            py::gil_scoped_release release;
            output = rasterizer_.interpolate_forward(as_span(attributes), static_cast<std::uint32_t>(attributes.shape(1)),
                                                     as_span(indices), r);
        }
        return to_array(std::move(output.values), {static_cast<py::ssize_t>(output.height),
                                                   static_cast<py::ssize_t>(output.width),
                                                   static_cast<py::ssize_t>(output.attribute_count)});
    }

    py::tuple interpolate_backward(const FloatInput& attributes, const U32Input& indices, const FloatInput& raster,
                                   const FloatInput& grad_interpolated) {
        if (attributes.ndim() != 2 || attributes.shape(0) <= 0 || attributes.shape(1) <= 0)
            throw std::invalid_argument("vertex_attributes must have shape [vertex_count, attribute_count]");
        validate_indices(indices);
        const RasterizeOutput r = raster_from_array(raster);
        if (grad_interpolated.ndim() != 3 || grad_interpolated.shape(0) != raster.shape(0) ||
            grad_interpolated.shape(1) != raster.shape(1) || grad_interpolated.shape(2) != attributes.shape(1))
            throw std::invalid_argument("grad_interpolated must have shape [height, width, attribute_count]");
        InterpolateGradients g;
        {
            py::gil_scoped_release release;
            g = rasterizer_.interpolate_backward(as_span(attributes), static_cast<std::uint32_t>(attributes.shape(1)),
                                                 as_span(indices), r, as_span(grad_interpolated));
        }
        auto ga = to_array(std::move(g.attributes), {attributes.shape(0), attributes.shape(1)});
        auto gr = to_array(std::move(g.raster), {raster.shape(0), raster.shape(1), 4});
        return py::make_tuple(ga, gr);
    }

    py::array_t<float> texture_forward(const FloatInput& texture, const FloatInput& uv, const FloatInput& raster,
                                       const std::string& address_mode) {
        if (texture.ndim() != 3 || texture.shape(0) <= 0 || texture.shape(1) <= 0 || texture.shape(2) <= 0)
            throw std::invalid_argument("texture must have shape [texture_height, texture_width, channel_count]");
        if (raster.ndim() != 3 || raster.shape(2) != 4)
            throw std::invalid_argument("raster must have shape [image_height, image_width, 4]");
        if (uv.ndim() != 3 || uv.shape(0) != raster.shape(0) || uv.shape(1) != raster.shape(1) || uv.shape(2) != 2)
            throw std::invalid_argument("uv must have shape [image_height, image_width, 2]");
        const RasterizeOutput r = raster_from_array(raster);
        const TextureDesc desc{static_cast<std::uint32_t>(texture.shape(1)), static_cast<std::uint32_t>(texture.shape(0)),
                               static_cast<std::uint32_t>(texture.shape(2)), parse_address_mode(address_mode)};
        TextureOutput output;
        {
            py::gil_scoped_release release;
            output = rasterizer_.texture_forward(as_span(texture), desc, as_span(uv), r);
        }
        return to_array(std::move(output.values), {static_cast<py::ssize_t>(output.height),
                                                   static_cast<py::ssize_t>(output.width),
                                                   static_cast<py::ssize_t>(output.channel_count)});
    }

    py::tuple texture_backward(const FloatInput& texture, const FloatInput& uv, const FloatInput& raster,
                               const FloatInput& grad_sampled, const std::string& address_mode,
                               bool compute_uv_gradient) {
        if (texture.ndim() != 3 || texture.shape(0) <= 0 || texture.shape(1) <= 0 || texture.shape(2) <= 0)
            throw std::invalid_argument("texture must have shape [texture_height, texture_width, channel_count]");
        if (raster.ndim() != 3 || raster.shape(2) != 4 || uv.ndim() != 3 || uv.shape(0) != raster.shape(0) ||
            uv.shape(1) != raster.shape(1) || uv.shape(2) != 2)
            throw std::invalid_argument("uv and raster dimensions are incompatible");
        if (grad_sampled.ndim() != 3 || grad_sampled.shape(0) != raster.shape(0) ||
            grad_sampled.shape(1) != raster.shape(1) || grad_sampled.shape(2) != texture.shape(2))
            throw std::invalid_argument("grad_sampled must have shape [image_height, image_width, channel_count]");
        const RasterizeOutput r = raster_from_array(raster);
        const TextureDesc desc{static_cast<std::uint32_t>(texture.shape(1)), static_cast<std::uint32_t>(texture.shape(0)),
                               static_cast<std::uint32_t>(texture.shape(2)), parse_address_mode(address_mode)};
        TextureGradients g;
        {
            py::gil_scoped_release release;
            g = rasterizer_.texture_backward(as_span(texture), desc, as_span(uv), r, as_span(grad_sampled),
                                             compute_uv_gradient);
        }
        auto gt = to_array(std::move(g.texture), {texture.shape(0), texture.shape(1), texture.shape(2)});
        if (!compute_uv_gradient) return py::make_tuple(gt, py::none());
        auto gu = to_array(std::move(g.uv), {uv.shape(0), uv.shape(1), 2});
        return py::make_tuple(gt, gu);
    }

    [[nodiscard]] const DeviceInfo& device_info() const noexcept { return context_.device_info(); }
    [[nodiscard]] std::uintptr_t context_handle() const noexcept {
        return reinterpret_cast<std::uintptr_t>(context_.handle());
    }

private:
    Context context_;
    Rasterizer rasterizer_;
};

class PythonTextureBaker {
public:
    PythonTextureBaker(std::uint32_t device_index, bool enable_validation)
        : context_(ContextOptions{device_index, enable_validation}), baker_(context_), refiner_(context_) {}

    // python/bindings.cpp:161-287
    py::tuple bake(const FloatInput& positions, const FloatInput& normals, const FloatInput& uv,
                   const U32Input& indices, const py::list& images, const FloatInput& world_to_clip,
                   const FloatInput& camera_positions, const py::object& visibility_masks, const py::object& viewports,
                   const std::pair<std::uint32_t, std::uint32_t>& resolution, const std::string& blend_mode,
                   const std::string& visibility_mode, std::uint32_t pcf_radius, bool allow_visibility_fallback,
                   float softmax_scale, float mask_floor) {
        if (positions.ndim() != 2 || positions.shape(1) != 3 || positions.shape(0) <= 0)
            throw std::invalid_argument("positions must have shape [vertex_count, 3]");
        if (normals.ndim() != 2 || normals.shape(0) != positions.shape(0) || normals.shape(1) != 3)
            throw std::invalid_argument("normals must have shape [vertex_count, 3]");
        if (uv.ndim() != 2 || uv.shape(0) != positions.shape(0) || uv.shape(1) != 2)
            throw std::invalid_argument("uv must have shape [vertex_count, 2]");
        validate_indices(indices);
        const auto view_count = static_cast<py::ssize_t>(images.size());
        if (view_count <= 0 || world_to_clip.ndim() != 3 || world_to_clip.shape(0) != view_count ||
            world_to_clip.shape(1) != 4 || world_to_clip.shape(2) != 4)
            throw std::invalid_argument("world_to_clip must have shape [view_count, 4, 4]");
        if (camera_positions.ndim() != 2 || camera_positions.shape(0) != view_count || camera_positions.shape(1) != 3)
            throw std::invalid_argument("camera_positions must have shape [view_count, 3]");
        py::list mask_list, viewport_list;
        if (!visibility_masks.is_none()) {
            mask_list = visibility_masks.cast<py::list>();
            if (static_cast<py::ssize_t>(mask_list.size()) != view_count)
                throw std::invalid_argument("visibility_masks must contain one entry per view");
        }
        if (!viewports.is_none()) {
            viewport_list = viewports.cast<py::list>();
            if (static_cast<py::ssize_t>(viewport_list.size()) != view_count)
                throw std::invalid_argument("viewports must contain one entry per view");
        }
        std::vector<ProjectionView> views(static_cast<std::size_t>(view_count));
        for (py::ssize_t i = 0; i < view_count; ++i) {
            const auto image = images[static_cast<std::size_t>(i)].cast<FloatInput>();
            if (image.ndim() != 3 || image.shape(0) <= 0 || image.shape(1) <= 0 ||
                (image.shape(2) != 3 && image.shape(2) != 4))
                throw std::invalid_argument("each image must have shape [height, width, 3 or 4]");
            auto& view = views[static_cast<std::size_t>(i)];
            view.height = static_cast<std::uint32_t>(image.shape(0));
            view.width = static_cast<std::uint32_t>(image.shape(1));
            view.channel_count = static_cast<std::uint32_t>(image.shape(2));
            view.image.assign(image.data(), image.data() + image.size());
            std::copy_n(world_to_clip.data() + i * 16, 16, view.world_to_clip.begin());
            std::copy_n(camera_positions.data() + i * 3, 3, view.camera_position.begin());
            if (!visibility_masks.is_none() && !mask_list[static_cast<std::size_t>(i)].is_none()) {
                const auto mask = mask_list[static_cast<std::size_t>(i)].cast<FloatInput>();
                if (mask.ndim() != 2 || mask.shape(0) != image.shape(0) || mask.shape(1) != image.shape(1))
                    throw std::invalid_argument("each visibility mask must have shape [image_height, image_width]");
                view.visibility_mask.assign(mask.data(), mask.data() + mask.size());
            }
            if (!viewports.is_none() && !viewport_list[static_cast<std::size_t>(i)].is_none()) {
                const auto v = viewport_list[static_cast<std::size_t>(i)].cast<std::array<float, 4>>();
                view.viewport = Viewport{v[0], v[1], v[2], v[3]};
            }
        }
        TextureBakeOptions options;
        options.height = resolution.first;
        options.width = resolution.second;
        options.pcf_radius = pcf_radius;
        options.allow_visibility_fallback = allow_visibility_fallback;
        if (blend_mode == "best_view")
            options.blend_mode = ProjectionBlendMode::best_view;
        else if (blend_mode == "weighted_average")
            options.blend_mode = ProjectionBlendMode::weighted_average;
        else if (blend_mode == "softmax")
            options.blend_mode = ProjectionBlendMode::softmax;
        else
            throw std::invalid_argument("blend_mode must be 'best_view', 'weighted_average', or 'softmax'");
        options.softmax_scale = softmax_scale;
        options.mask_floor = mask_floor;
        if (visibility_mode == "shadow_map")
            options.visibility_mode = VisibilityMode::shadow_map;
        else if (visibility_mode == "hybrid_ray_query")
            options.visibility_mode = VisibilityMode::hybrid_ray_query;
        else if (visibility_mode == "ray_query")
            options.visibility_mode = VisibilityMode::ray_query;
        else
            throw std::invalid_argument("visibility_mode must be 'shadow_map', 'hybrid_ray_query', or 'ray_query'");
        TextureBakeOutput out;
        {
            py::gil_scoped_release release;
            out = baker_.bake(as_span(positions), as_span(normals), as_span(uv), as_span(indices), views, options);
        }
        const auto H = static_cast<py::ssize_t>(out.height), W = static_cast<py::ssize_t>(out.width);
        auto color = to_array(std::move(out.color), {H, W, 4});
        auto confidence = to_array(std::move(out.confidence), {H, W});
        py::array_t<std::uint32_t> source({H, W});
        std::copy(out.source_view.begin(), out.source_view.end(), source.mutable_data());
        auto valid = to_array(std::move(out.valid_mask), {H, W});
        auto coverage = to_array(std::move(out.coverage_mask), {H, W});
        return py::make_tuple(color, confidence, source, valid, coverage, out.used_ray_query);
    }

    // python/bindings.cpp:289-404
    py::tuple refine_texture(const FloatInput& initial_texture, const FloatInput& positions, const FloatInput& uv,
                             const U32Input& indices, const py::list& images, const FloatInput& world_to_clip,
                             const py::object& visibility_masks, const py::object& viewports,
                             const py::object& seam_uv_pairs, std::uint32_t steps, std::uint32_t batch_size,
                             float learning_rate, float minimum_learning_rate, float photometric_epsilon,
                             std::uint32_t seam_polish_steps, float seam_learning_rate) {
        if (initial_texture.ndim() != 3 || initial_texture.shape(0) <= 0 || initial_texture.shape(1) <= 0 ||
            initial_texture.shape(2) != 3)
            throw std::invalid_argument("initial_texture must have shape [height, width, 3]");
        if (positions.ndim() != 2 || positions.shape(0) <= 0 || positions.shape(1) != 3 || uv.ndim() != 2 ||
            uv.shape(0) != positions.shape(0) || uv.shape(1) != 2)
            throw std::invalid_argument("positions and uv must have shapes [vertex_count, 3] and [vertex_count, 2]");
        validate_indices(indices);
        const auto view_count = static_cast<py::ssize_t>(images.size());
        if (view_count <= 0 || world_to_clip.ndim() != 3 || world_to_clip.shape(0) != view_count ||
            world_to_clip.shape(1) != 4 || world_to_clip.shape(2) != 4)
            throw std::invalid_argument("world_to_clip must have shape [view_count, 4, 4]");
        py::list mask_list, viewport_list;
        if (!visibility_masks.is_none()) {
            mask_list = visibility_masks.cast<py::list>();
            if (static_cast<py::ssize_t>(mask_list.size()) != view_count)
                throw std::invalid_argument("visibility_masks must contain one entry per view");
        }
        if (!viewports.is_none()) {
            viewport_list = viewports.cast<py::list>();
            if (static_cast<py::ssize_t>(viewport_list.size()) != view_count)
                throw std::invalid_argument("viewports must contain one entry per view");
        }
        std::vector<ProjectionView> views(static_cast<std::size_t>(view_count));
        for (py::ssize_t i = 0; i < view_count; ++i) {
            const auto image = images[static_cast<std::size_t>(i)].cast<FloatInput>();
            if (image.ndim() != 3 || image.shape(0) <= 0 || image.shape(1) <= 0 ||
                (image.shape(2) != 3 && image.shape(2) != 4))
                throw std::invalid_argument("each image must have shape [height, width, 3 or 4]");
            auto& view = views[static_cast<std::size_t>(i)];
            view.height = static_cast<std::uint32_t>(image.shape(0));
            view.width = static_cast<std::uint32_t>(image.shape(1));
            view.channel_count = static_cast<std::uint32_t>(image.shape(2));
            view.image.assign(image.data(), image.data() + image.size());
            std::copy_n(world_to_clip.data() + i * 16, 16, view.world_to_clip.begin());
            if (!visibility_masks.is_none() && !mask_list[static_cast<std::size_t>(i)].is_none()) {
                const auto mask = mask_list[static_cast<std::size_t>(i)].cast<FloatInput>();
                if (mask.ndim() != 2 || mask.shape(0) != image.shape(0) || mask.shape(1) != image.shape(1))
                    throw std::invalid_argument("each visibility mask must match its image dimensions");
                view.visibility_mask.assign(mask.data(), mask.data() + mask.size());
            }
            if (!viewports.is_none() && !viewport_list[static_cast<std::size_t>(i)].is_none()) {
                const auto v = viewport_list[static_cast<std::size_t>(i)].cast<std::array<float, 4>>();
                view.viewport = Viewport{v[0], v[1], v[2], v[3]};
            }
        }
        TextureRefineOptions options;
        options.height = static_cast<std::uint32_t>(initial_texture.shape(0));
        options.width = static_cast<std::uint32_t>(initial_texture.shape(1));
        options.steps = steps;
        options.batch_size = batch_size;
        options.learning_rate = learning_rate;
        options.minimum_learning_rate = minimum_learning_rate;
        options.photometric_epsilon = photometric_epsilon;
        options.seam_polish_steps = seam_polish_steps;
        options.seam_learning_rate = seam_learning_rate;
        FloatInput seam_array;
        std::span<const float> seam_span;
        if (!seam_uv_pairs.is_none()) {
            seam_array = seam_uv_pairs.cast<FloatInput>();
            if (seam_array.ndim() != 3 || seam_array.shape(1) != 2 || seam_array.shape(2) != 2)
                throw std::invalid_argument("seam_uv_pairs must have shape [pair_count, 2, 2]");
            seam_span = as_span(seam_array);
        }
        TextureRefineOutput output;
        {
            py::gil_scoped_release release;
            output = refiner_.refine(as_span(initial_texture), as_span(positions), as_span(uv), as_span(indices),
                                     views, seam_span, options);
        }
        auto color = to_array(std::move(output.color), {static_cast<py::ssize_t>(output.height),
                                                        static_cast<py::ssize_t>(output.width), 3});
        auto history = to_array(std::move(output.loss_history), {static_cast<py::ssize_t>(options.steps)});
        const auto seam_size = static_cast<py::ssize_t>(output.seam_loss_history.size());
        auto seam_history = to_array(std::move(output.seam_loss_history), {seam_size});
        return py::make_tuple(color, history, seam_history, output.precompute_seconds, output.optimization_seconds);
    }

    [[nodiscard]] const DeviceInfo& device_info() const noexcept { return context_.device_info(); }
    [[nodiscard]] std::uintptr_t context_handle() const noexcept {
        return reinterpret_cast<std::uintptr_t>(context_.handle());
    }

private:
    Context context_;
    TextureBaker baker_;
    TextureRefiner refiner_;
};

}  // namespace

PYBIND11_MODULE(_photara_drender, m) {
    m.doc() = "Differentiable rasterization and atlas refinement on AMD Instinct (HIP, gfx950)";

    py::class_<DeviceInfo>(m, "DeviceInfo")
        .def_readonly("name", &DeviceInfo::name)
        .def_readonly("vendor_id", &DeviceInfo::vendor_id)
        .def_readonly("device_id", &DeviceInfo::device_id)
        .def_readonly("api_version", &DeviceInfo::api_version)
        .def_readonly("supports_ray_query", &DeviceInfo::supports_ray_query)
        .def("__repr__", [](const DeviceInfo& info) {
            return "DeviceInfo(name='" + info.name + "', vendor_id=" + std::to_string(info.vendor_id) + ")";
        });

    py::class_<PythonRasterizer>(m, "Rasterizer")
        .def(py::init<std::uint32_t, bool>(), py::arg("device_index") = 0, py::arg("enable_validation") = false)
        .def_property_readonly("device_info", &PythonRasterizer::device_info, py::return_value_policy::reference_internal)
        .def_property_readonly("context_handle", &PythonRasterizer::context_handle)
        .def("forward", &PythonRasterizer::forward, py::arg("positions"), py::arg("indices"), py::arg("resolution"),
             py::arg("cull_mode") = "none", py::arg("output_barycentric_derivatives") = true,
             py::arg("viewport") = py::none())
        .def("backward", &PythonRasterizer::backward, py::arg("positions"), py::arg("indices"), py::arg("raster"),
             py::arg("grad_raster"), py::arg("viewport") = py::none())
        .def("interpolate_forward", &PythonRasterizer::interpolate_forward, py::arg("vertex_attributes"),
             py::arg("indices"), py::arg("raster"))
        .def("interpolate_backward", &PythonRasterizer::interpolate_backward, py::arg("vertex_attributes"),
             py::arg("indices"), py::arg("raster"), py::arg("grad_interpolated"))
        .def("texture_forward", &PythonRasterizer::texture_forward, py::arg("texture"), py::arg("uv"),
             py::arg("raster"), py::arg("address_mode") = "clamp")
        .def("texture_backward", &PythonRasterizer::texture_backward, py::arg("texture"), py::arg("uv"),
             py::arg("raster"), py::arg("grad_sampled"), py::arg("address_mode") = "clamp",
             py::arg("compute_uv_gradient") = true);

    py::class_<PythonTextureBaker>(m, "TextureBaker")
        .def(py::init<std::uint32_t, bool>(), py::arg("device_index") = 0, py::arg("enable_validation") = false)
        .def_property_readonly("device_info", &PythonTextureBaker::device_info,
                               py::return_value_policy::reference_internal)
        .def_property_readonly("context_handle", &PythonTextureBaker::context_handle)
        .def("bake", &PythonTextureBaker::bake, py::arg("positions"), py::arg("normals"), py::arg("uv"),
             py::arg("indices"), py::arg("images"), py::arg("world_to_clip"), py::arg("camera_positions"),
             py::arg("visibility_masks") = py::none(), py::arg("viewports") = py::none(),
             py::arg("resolution") = std::pair<std::uint32_t, std::uint32_t>{1024U, 1024U},
             py::arg("blend_mode") = "weighted_average", py::arg("visibility_mode") = "ray_query",
             py::arg("pcf_radius") = 1, py::arg("allow_visibility_fallback") = true, py::arg("softmax_scale") = 0.0f,
             py::arg("mask_floor") = 0.0f)
        .def("refine_texture", &PythonTextureBaker::refine_texture, py::arg("initial_texture"), py::arg("positions"),
             py::arg("uv"), py::arg("indices"), py::arg("images"), py::arg("world_to_clip"),
             py::arg("visibility_masks") = py::none(), py::arg("viewports") = py::none(),
             py::arg("seam_uv_pairs") = py::none(), py::arg("steps") = 1000, py::arg("batch_size") = 4,
             py::arg("learning_rate") = 5e-3F, py::arg("minimum_learning_rate") = 2.5e-4F,
             py::arg("photometric_epsilon") = 1e-3F, py::arg("seam_polish_steps") = 30,
             py::arg("seam_learning_rate") = 1e-3F);

    m.def("enumerate_devices", &Context::enumerate_devices);
    m.def(
        "suggest_softmax_scale",
        [](const FloatInput& positions, const FloatInput& camera_positions) {
            return suggest_softmax_scale(as_span(positions), as_span(camera_positions));
        },
        py::arg("positions"), py::arg("camera_positions"));
    // python/bindings.cpp:786-834 of the reference: (color, filled_mask, gutter, unobserved, remaining)
    m.def(
        "pad_texture_atlas",
        [](const FloatInput& color, const FloatInput& valid_mask, std::uint32_t margin, const py::object& coverage_mask,
           bool fill_unobserved, std::uint32_t maximum_fill_distance) {
            const auto ci = color.request();
            const auto vi = valid_mask.request();
            if (ci.ndim != 3 || ci.shape[0] <= 0 || ci.shape[1] <= 0 || (ci.shape[2] != 3 && ci.shape[2] != 4))
                throw std::invalid_argument("color must have shape [height, width, 3 or 4]");
            if (vi.ndim != 2 || vi.shape[0] != ci.shape[0] || vi.shape[1] != ci.shape[1])
                throw std::invalid_argument("valid_mask must match the color dimensions");
            TextureBakeOutput output;
            output.width = static_cast<std::uint32_t>(ci.shape[1]);
            output.height = static_cast<std::uint32_t>(ci.shape[0]);
            output.color.assign(color.data(), color.data() + color.size());
            output.valid_mask.assign(valid_mask.data(), valid_mask.data() + valid_mask.size());
            if (!coverage_mask.is_none()) {
                const auto coverage = py::cast<FloatInput>(coverage_mask);
                const auto cov = coverage.request();
                if (cov.ndim != 2 || cov.shape[0] != ci.shape[0] || cov.shape[1] != ci.shape[1])
                    throw std::invalid_argument("coverage_mask must match the color dimensions");
                output.coverage_mask.assign(coverage.data(), coverage.data() + coverage.size());
            }
            TexturePaddingOptions options;
            options.margin = margin;
            options.fill_unobserved = fill_unobserved;
            options.maximum_fill_distance = maximum_fill_distance;
            const auto stats = pad_texture_atlas(output, options);
            py::array_t<float> out_color({ci.shape[0], ci.shape[1], ci.shape[2]});
            std::memcpy(out_color.mutable_data(), output.color.data(), output.color.size() * sizeof(float));
            py::array_t<float> out_filled({ci.shape[0], ci.shape[1]});
            std::memcpy(out_filled.mutable_data(), output.filled_mask.data(), output.filled_mask.size() * sizeof(float));
            return py::make_tuple(out_color, out_filled, stats.gutter_texels, stats.unobserved_texels,
                                  stats.remaining_unobserved);
        },
        py::arg("color"), py::arg("valid_mask"), py::arg("margin") = 4, py::arg("coverage_mask") = py::none(),
        py::arg("fill_unobserved") = true, py::arg("maximum_fill_distance") = 0);
}
