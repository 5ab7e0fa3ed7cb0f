// pdr_bake: TextureBaker::bake (src/texture_baker.cpp:156-361) on the GPU.
//
// Atlas raster of uv * 2 - 1 (no derivatives), world positions and normals
// interpolated per texel, a shadow-map raster of every view, then the
// projection (shaders/project_texture.hlsl:117-247) of up to kBakeChunk views
// per launch and the reference's final normalisation.  Visibility: this
// backend has no ray queries, so it takes the reference's path for a device
// without them (shadow map; ray-query modes need allow_visibility_fallback).
#include "pdr_host.hpp"

using namespace pdrh;

namespace {

constexpr uint32_t kBakeChunk = 32;  // views rasterized and projected per launch

bool finite_viewport(const float vp[4]) {
    return std::isfinite(vp[0]) && std::isfinite(vp[1]) && std::isfinite(vp[2]) && std::isfinite(vp[3]) &&
           vp[2] > 0.0f && vp[3] > 0.0f;
}

// src/texture_padding.cpp:187-222
float softmax_scale_of(const float* pos, uint32_t V, const float* cams, uint32_t nc) {
    const size_t stride = std::max<size_t>(1, V / 10000);
    double best = std::numeric_limits<double>::infinity();
    for (size_t v = 0; v < V; v += stride) {
        const double x = pos[v * 3], y = pos[v * 3 + 1], z = pos[v * 3 + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) continue;
        for (size_t c = 0; c < nc; ++c) {
            const double dx = cams[c * 3] - x, dy = cams[c * 3 + 1] - y, dz = cams[c * 3 + 2] - z;
            best = std::min(best, dx * dx + dy * dy + dz * dz);
        }
    }
    if (!std::isfinite(best) || best <= 0.0) return 0.0f;
    return static_cast<float>(20.0 * best);
}

}  // namespace

extern "C" {

void pdr_bake_options_default(pdr_bake_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // include/photara_drender/types.hpp:154-179
    o->width = 1024;
    o->height = 1024;
    o->blend_mode = 0;
    o->visibility_mode = 2;
    o->pcf_radius = 1;
    o->ray_origin_bias = 1e-4f;
    o->allow_visibility_fallback = 1;
    o->mask_floor = 0.0f;
    o->softmax_scale = 0.0f;
}

void pdr_bake_view_default(pdr_bake_view* v) {
    if (!v) return;
    std::memset(v, 0, sizeof(*v));
    // include/photara_drender/types.hpp:135-152
    v->weight = 1.0f;
    v->depth_bias = 1e-3f;
    v->min_view_cosine = 0.05f;
}

float pdr_suggest_softmax_scale(const float* pos, uint32_t V, const float* cams, uint32_t nc) {
    if (!pos || !cams || V == 0 || nc == 0) return 0.0f;
    return softmax_scale_of(pos, V, cams, nc);
}

}  // extern "C"

namespace {

// Outputs of one bake run: the blend outputs of TextureBaker::bake, or (with
// features set) only the warm-start feature lists [P, 40].
struct BakeOutputs {
    float* color;
    float* confidence;
    uint32_t* source_view;
    float* valid;
    float* coverage;
    uint32_t* face_ids;
    float* features;
};

void run_bake(pdr_context ctx, const float* positions, const float* normals, const float* uv, uint32_t V,
              const uint32_t* idx, uint32_t T, const pdr_bake_view* views, uint32_t nviews,
              const pdr_bake_options* options, const BakeOutputs& out) {
    require_ctx(ctx);
    pdr_bake_options o;
    if (options)
        o = *options;
    else
        pdr_bake_options_default(&o);
    // src/texture_baker.cpp:40-62 validate_mesh
    if (!positions || V == 0) raise(PDR_INVALID_ARGUMENT, "positions must have shape [vertex_count, 3]");
    if (!normals) raise(PDR_INVALID_ARGUMENT, "normals must have shape [vertex_count, 3]");
    if (!uv) raise(PDR_INVALID_ARGUMENT, "uv must have shape [vertex_count, 2]");
    if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices must have shape [triangle_count, 3]");
    const bool idx_device = is_device_pointer(idx);
    if (!idx_device)
        check_indices_host(idx, static_cast<size_t>(T) * 3, V, PDR_OUT_OF_RANGE,
                           "triangle_indices contains an invalid vertex index");
    // :163-178
    if (!views || nviews == 0) raise(PDR_INVALID_ARGUMENT, "at least one projection view is required");
    if (o.width == 0 || o.height == 0 || o.pcf_radius > 4 || !std::isfinite(o.ray_origin_bias) ||
        o.ray_origin_bias <= 0.0f || !std::isfinite(o.mask_floor) || o.mask_floor < 0.0f || o.mask_floor > 1.0f ||
        o.blend_mode > 2 || o.visibility_mode > 2)
        raise(PDR_INVALID_ARGUMENT, "atlas dimensions are invalid or pcf_radius exceeds 4");
    if (o.visibility_mode != 0 && !o.allow_visibility_fallback)
        raise(PDR_RUNTIME, "ray-query visibility was requested but the selected Vulkan device does not support it");
    // :64-80 validate_view
    for (uint32_t k = 0; k < nviews; ++k) {
        const pdr_bake_view& v = views[k];
        if (v.width == 0 || v.height == 0 || (v.channels != 3 && v.channels != 4))
            raise(PDR_INVALID_ARGUMENT, "projection image must have nonzero dimensions and 3 or 4 channels");
        if (!v.image) raise(PDR_INVALID_ARGUMENT, "projection image data size does not match its descriptor");
        if (!std::isfinite(v.weight) || v.weight < 0.0f || !std::isfinite(v.depth_bias) || v.depth_bias < 0.0f ||
            !std::isfinite(v.min_view_cosine) || v.min_view_cosine < -1.0f || v.min_view_cosine > 1.0f)
            raise(PDR_INVALID_ARGUMENT, "projection confidence parameters are invalid");
        if (v.has_viewport && !finite_viewport(v.viewport))
            raise(PDR_INVALID_ARGUMENT, "projection viewport must contain finite values and positive dimensions");
    }
    const uint64_t P64 = static_cast<uint64_t>(o.width) * o.height;
    if (P64 > 0xFFFFFFFFull) raise(PDR_OVERFLOW, "Rasterization output size overflow");
    const uint32_t P = static_cast<uint32_t>(P64);
    const bool features = out.features != nullptr;
    if (features && P64 * pdr::kMlpFeatures > 0xFFFFFFFFull)
        raise(PDR_OVERFLOW, "feature output exceeds backend element limits");
    if (features && reinterpret_cast<uintptr_t>(out.features) % 16 != 0)
        raise(PDR_INVALID_ARGUMENT, "features_out must be 16-byte aligned");

    Lock lock(ctx);
    const hipStream_t st = ctx->use_stream(false, nullptr);
    if (idx_device)
        check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                             "triangle_indices contains an invalid vertex index");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");  // device inputs may be in flight elsewhere
    DeviceBuffers bufs(ctx, features ? "bake_features" : "bake");
    const float* dpos = device_view(bufs, positions, static_cast<size_t>(V) * 3, st);
    const float* dnrm = device_view(bufs, normals, static_cast<size_t>(V) * 3, st);
    const float* duv = device_view(bufs, uv, static_cast<size_t>(V) * 2, st);
    const uint32_t* didx = device_view(bufs, idx, static_cast<size_t>(T) * 3, st);

    // atlas raster of uv * 2 - 1 (src/texture_baker.cpp:85-94, 183-192), positions and normals per texel
    float* aclip = bufs.alloc<float>(static_cast<size_t>(V) * 4);
    hip_check(pdr::atlas_clip_positions(duv, V, aclip, st), "atlas_clip_positions");
    float* araster = bufs.alloc<float>(static_cast<size_t>(P) * 4);
    {
        const float avp[4] = {0.0f, 0.0f, static_cast<float>(o.width), static_cast<float>(o.height)};
        pdr::ViewTable vt{};
        vt.n = 1;
        vt.v[0] = make_view(aclip, o.width, o.height, avp);
        vt.v[0].raster = araster;
        finalize_table(vt);
        run_raster_forward(ctx, vt, didx, T, V, 0, 0, st);
    }
    float* apos = bufs.alloc<float>(static_cast<size_t>(P) * 3);
    float* anrm = bufs.alloc<float>(static_cast<size_t>(P) * 3);
    hip_check(pdr::interp_forward(pixel_table_one(araster, nullptr, nullptr, apos, nullptr, o.width, o.height), dpos, V,
                                  3, didx, T, st),
              "interp_forward");
    hip_check(pdr::interp_forward(pixel_table_one(araster, nullptr, nullptr, anrm, nullptr, o.width, o.height), dnrm, V,
                                  3, didx, T, st),
              "interp_forward");

    // softmax sharpness (src/texture_baker.cpp:206-222), on host copies of the positions
    float softmax_scale = o.softmax_scale;
    if (!features && o.blend_mode == 2 && !(softmax_scale > 0.0f && std::isfinite(softmax_scale))) {
        std::vector<float> hpos(static_cast<size_t>(V) * 3), cams(static_cast<size_t>(nviews) * 3);
        hip_check(hipMemcpyAsync(hpos.data(), dpos, hpos.size() * sizeof(float), hipMemcpyDefault, st),
                  "hipMemcpyAsync");
        sync(st);
        for (uint32_t k = 0; k < nviews; ++k)
            for (int c = 0; c < 3; ++c) cams[k * 3 + c] = views[k].camera_position[c];
        softmax_scale = softmax_scale_of(hpos.data(), V, cams.data(), nviews);
        if (!(softmax_scale > 0.0f) || !std::isfinite(softmax_scale)) softmax_scale = 1.0f;
    }

    float *acc3 = nullptr, *accw = nullptr, *best = nullptr, *feats = nullptr;
    uint32_t* src = nullptr;
    if (features) {
        feats = is_device_pointer(out.features) ? out.features
                                                : bufs.alloc<float>(static_cast<size_t>(P) * pdr::kMlpFeatures);
        hip_check(pdr::fill_f32(feats, 0.0f, static_cast<size_t>(P) * pdr::kMlpFeatures, st), "fill_f32");
    } else {
        acc3 = bufs.alloc<float>(static_cast<size_t>(P) * 3);
        accw = bufs.alloc<float>(P);
        best = bufs.alloc<float>(P);
        src = bufs.alloc<uint32_t>(P);
        hip_check(pdr::bake_init(P, acc3, accw, best, src, st), "bake_init");
    }
    float* clips = bufs.alloc<float>(static_cast<size_t>(std::min(nviews, kBakeChunk)) * V * 4);
    pdr::BakeView* dviews = bufs.alloc<pdr::BakeView>(kBakeChunk);
    // every view's shadow-map keys are kept in the context with the mesh they
    // came from, for a refinement of the same views (pdr_refine)
    auto& cache = ctx->view_keys;
    cache = {};
    cache.T = T;
    cache.V = V;
    size_t cache_count = 0;
    for (uint32_t k = 0; k < nviews; ++k) cache_count += pdr::view_key_count(views[k].width, views[k].height);
    unsigned long long* cache_keys = ctx->get_t<unsigned long long>("view_keys", cache_count);
    hip_check(hipMemcpyAsync(ctx->get_t<float>("view_keys_pos", static_cast<size_t>(V) * 3), dpos,
                             static_cast<size_t>(V) * 3 * sizeof(float), hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(ctx->get_t<uint32_t>("view_keys_idx", static_cast<size_t>(T) * 3), didx,
                             static_cast<size_t>(T) * 3 * sizeof(uint32_t), hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync");
    size_t cache_off = 0;
    for (uint32_t k0 = 0; k0 < nviews; k0 += kBakeChunk) {
        const uint32_t n = std::min(nviews - k0, kBakeChunk);
        DeviceBuffers chunk_bufs(ctx, features ? "bake_features_chunk" : "bake_chunk");
        std::vector<pdr::BakeView> hv(n);
        pdr::ViewTable vt{};
        vt.n = n;
        for (uint32_t j = 0; j < n; ++j) {
            const pdr_bake_view& v = views[k0 + j];
            float* clip = clips + static_cast<size_t>(j) * V * 4;
            hip_check(pdr::transform_positions(dpos, V, v.world_to_clip, clip, st), "transform_positions");
            float vp[4];
            resolve_viewport(v.width, v.height, v.has_viewport ? v.viewport : nullptr, vp);
            vt.v[j] = make_view(clip, v.width, v.height, vp);
            const size_t pix = static_cast<size_t>(v.width) * v.height;
            pdr::BakeView& b = hv[j];
            b.photo = device_view(chunk_bufs, v.image, pix * v.channels, st);
            b.mask = device_view(chunk_bufs, v.visibility_mask, pix, st);
            std::memcpy(b.m, v.world_to_clip, sizeof(b.m));
            std::memcpy(b.cam, v.camera_position, sizeof(b.cam));
            b.weight = !features && o.blend_mode == 2 ? v.weight * softmax_scale : v.weight;  // :307-312
            b.depth_bias = v.depth_bias;
            b.min_cos = v.min_view_cosine;
            b.mask_floor = o.mask_floor;
            std::memcpy(b.vp, vp, sizeof(b.vp));
            b.width = v.width;
            b.height = v.height;
            b.channels = v.channels;
            b.index = k0 + j;
        }
        finalize_table(vt);
        // shadow maps (:266-276): each view's raster keys, read as (depth, covered) by the projection
        const unsigned long long* keys = run_raster_keys(ctx, vt, didx, T, V, 0, st);
        for (uint32_t j = 0; j < n; ++j) {
            hv[j].shadow = keys + vt.v[j].key_base;
            hv[j].key_tx = vt.v[j].key_tx;
            const size_t nk = pdr::view_key_count(vt.v[j].width, vt.v[j].height);
            hip_check(hipMemcpyAsync(cache_keys + cache_off, keys + vt.v[j].key_base, nk * 8,
                                     hipMemcpyDeviceToDevice, st),
                      "hipMemcpyAsync");
            cache.width.push_back(vt.v[j].width);
            cache.height.push_back(vt.v[j].height);
            cache.viewport.push_back({vt.v[j].vx, vt.v[j].vy, vt.v[j].vw, vt.v[j].vh});
            std::array<float, 16> m;
            std::memcpy(m.data(), views[k0 + j].world_to_clip, sizeof(m));
            cache.matrix.push_back(m);
            cache.offset.push_back(cache_off);
            cache_off += nk;
        }
        hip_check(hipMemcpyAsync(dviews, hv.data(), n * sizeof(pdr::BakeView), hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
        if (features)
            ctx->timed("bake_features", st, [&] {
                hip_check(pdr::bake_features(araster, apos, anrm, P, dviews, n, o.pcf_radius, feats, st),
                          "bake_features");
            });
        else
            ctx->timed("bake_project", st, [&] {
                hip_check(pdr::bake_project(araster, apos, anrm, P, dviews, n, o.blend_mode, o.pcf_radius, acc3, accw,
                                            best, src, st),
                          "bake_project");
            });
        sync(st);  // the next chunk stages its images into the same slots
    }

    if (features) {
        copy_any(out.features, feats, static_cast<size_t>(P) * pdr::kMlpFeatures * 4, st);
        sync(st);
        return;
    }
    // outputs (src/texture_baker.cpp:193-204, 331-357)
    const bool dev_out = is_device_pointer(out.color);
    float* dcolor = dev_out ? out.color : bufs.alloc<float>(static_cast<size_t>(P) * 4);
    float* dconf = is_device_pointer(out.confidence) ? out.confidence : bufs.alloc<float>(P);
    float* dvalid = is_device_pointer(out.valid) ? out.valid : bufs.alloc<float>(P);
    float* dcov = is_device_pointer(out.coverage) ? out.coverage : bufs.alloc<float>(P);
    uint32_t* dface =
        out.face_ids ? (is_device_pointer(out.face_ids) ? out.face_ids : bufs.alloc<uint32_t>(P)) : nullptr;
    hip_check(pdr::bake_finish(P, o.blend_mode, araster, acc3, accw, dcolor, dconf, dvalid, dcov, dface, st),
              "bake_finish");
    copy_any(out.color, dcolor, static_cast<size_t>(P) * 16, st);
    copy_any(out.confidence, dconf, static_cast<size_t>(P) * 4, st);
    copy_any(out.source_view, src, static_cast<size_t>(P) * 4, st);
    copy_any(out.valid, dvalid, static_cast<size_t>(P) * 4, st);
    copy_any(out.coverage, dcov, static_cast<size_t>(P) * 4, st);
    if (out.face_ids) copy_any(out.face_ids, dface, static_cast<size_t>(P) * 4, st);
    sync(st);
}

}  // namespace

extern "C" {

pdr_status pdr_bake(pdr_context ctx, const float* positions, const float* normals, const float* uv, uint32_t V,
                    const uint32_t* idx, uint32_t T, const pdr_bake_view* views, uint32_t nviews,
                    const pdr_bake_options* options, float* color_out, float* confidence_out,
                    uint32_t* source_view_out, float* valid_out, float* coverage_out, uint32_t* face_ids_out,
                    int* used_ray_query_out) {
    return guarded([&] {
        run_bake(ctx, positions, normals, uv, V, idx, T, views, nviews, options,
                 BakeOutputs{color_out, confidence_out, source_view_out, valid_out, coverage_out, face_ids_out,
                             nullptr});
        if (used_ray_query_out) *used_ray_query_out = 0;
    });
}

pdr_status pdr_bake_features(pdr_context ctx, const float* positions, const float* normals, const float* uv,
                             uint32_t V, const uint32_t* idx, uint32_t T, const pdr_bake_view* views, uint32_t nviews,
                             const pdr_bake_options* options, float* features_out) {
    return guarded([&] {
        require_ptr(features_out, "features_out");
        run_bake(ctx, positions, normals, uv, V, idx, T, views, nviews, options,
                 BakeOutputs{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, features_out});
    });
}

}  // extern "C"
