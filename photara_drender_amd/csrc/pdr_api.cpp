// libpdr C ABI: contexts, device scratch, host staging, validation and the
// kernel sequences behind every entry point of include/pdr.h.
//
// The reference builds tile bins and adjacency tables on the CPU and submits
// one synchronous Vulkan dispatch per kernel (src/rasterizer.cpp,
// src/context.cpp:455-527).  Here everything after validation runs on the GPU:
// tile binning, gradient reductions and the whole refinement loop (captured as
// a hipGraph).  Host arrays are staged through per-context device scratch;
// device arrays are used in place and the work is enqueued on the caller's
// stream.
#include "pdr_host.hpp"

using namespace pdrh;

thread_local std::string pdrh::g_last_error;

extern "C" {

const char* pdr_last_error(void) { return g_last_error.c_str(); }

pdr_status pdr_context_create(uint32_t device_index, int enable_validation, pdr_context* out) {
    (void)enable_validation;
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
            (void)hipGetLastError();
            raise(PDR_RUNTIME, "No HIP device is available");
        }
        if (device_index >= static_cast<uint32_t>(count)) raise(PDR_OUT_OF_RANGE, "device_index is out of range");
        auto ctx = std::make_unique<pdr_context_s>();
        ctx->device = static_cast<int>(device_index);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        hipMemPool_t pool = nullptr;  // keep freed temporaries of bake / refine calls for reuse
        if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess && pool) {
            uint64_t keep = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        (void)hipGetLastError();
        hip_check(hipEventCreateWithFlags(&ctx->handoff, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&ctx->pinned), 4096, hipHostMallocDefault), "hipHostMalloc");
        pdr::noalloc_scratch_init(ctx->device);  // the device's no-allocation scratch (device current)
        *out = ctx.release();
    });
}

void pdr_context_destroy(pdr_context ctx) { delete ctx; }

pdr_status pdr_context_set_flags(pdr_context ctx, uint32_t flags) {
    return guarded([&] {
        require_ctx(ctx);
        std::lock_guard<std::mutex> g(ctx->mutex);
        ctx->flags = flags;
    });
}

uint32_t pdr_context_flags(pdr_context ctx) { return ctx ? ctx->flags : 0u; }

static void fill_info(int dev, pdr_device_info* info) {
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%s (%s)", prop.name, prop.gcnArchName);
    info->vendor_id = 0x1002u;
    info->device_id = static_cast<uint32_t>(prop.pciDeviceID);
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    info->api_version = static_cast<uint32_t>(rt);
    info->supports_ray_query = 0;
}

pdr_status pdr_context_device_info(pdr_context ctx, pdr_device_info* info) {
    return guarded([&] {
        require_ctx(ctx);
        require_ptr(info, "info");
        fill_info(ctx->device, info);
    });
}

pdr_status pdr_enumerate_devices(pdr_device_info* infos, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        require_ptr(count, "count");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = static_cast<uint32_t>(n);
        for (int i = 0; i < n && infos && static_cast<uint32_t>(i) < capacity; ++i) fill_info(i, &infos[i]);
    });
}

pdr_status pdr_context_set_timing(pdr_context ctx, int enable) {
    return guarded([&] {
        require_ctx(ctx);
        std::lock_guard<std::mutex> g(ctx->mutex);
        ctx->resolve_timing();
        ctx->timing = enable != 0;
        ctx->totals.clear();
    });
}

pdr_status pdr_context_kernel_time(pdr_context ctx, const char* name, double* total_ms, uint64_t* launches) {
    return guarded([&] {
        require_ctx(ctx);
        require_ptr(name, "name");
        std::lock_guard<std::mutex> g(ctx->mutex);
        ctx->resolve_timing();
        const auto it = ctx->totals.find(name);
        if (total_ms) *total_ms = it == ctx->totals.end() ? 0.0 : it->second.first;
        if (launches) *launches = it == ctx->totals.end() ? 0 : it->second.second;
    });
}

pdr_status pdr_autotune_export(pdr_context ctx, char* buffer, size_t capacity, size_t* length) {
    return guarded([&] {
        require_ctx(ctx);
        std::string text;
        {
            Lock lock(ctx);
            for (const auto& kv : ctx->tuned) {
                char line[256];
                std::snprintf(line, sizeof(line), "%s %.9g %u %.9g %u %u %u %u\n", kv.first.c_str(),
                              static_cast<double>(kv.second.wave_work), kv.second.fwd_waves,
                              static_cast<double>(kv.second.flat_work), kv.second.bwd_waves, kv.second.tiled,
                              kv.second.bin_tiles, kv.second.order);
                text += line;
            }
        }
        if (length) *length = text.size() + 1;
        if (buffer && capacity) {
            const size_t n = std::min(capacity - 1, text.size());
            std::memcpy(buffer, text.data(), n);
            buffer[n] = '\0';
        }
    });
}

pdr_status pdr_autotune_import(pdr_context ctx, const char* text) {
    return guarded([&] {
        require_ctx(ctx);
        require_ptr(text, "text");
        std::map<std::string, pdr::TuneParams> parsed;
        const char* p = text;
        while (*p) {
            const char* e = std::strchr(p, '\n');
            const std::string line(p, e ? static_cast<size_t>(e - p) : std::strlen(p));
            p = e ? e + 1 : p + line.size();
            if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
            char kind[32];
            unsigned views = 0, lt = 0, lp = 0, fw = 0, bw = 0, tiled = 0, depth = 0, order = 0;
            double ww = 0.0, flat = 0.0;
            char tail = 0;
            // ten fields, or eleven with the traversal order (cache lines written before it existed omit it)
            const int got = std::sscanf(line.c_str(), "%31s %u %u %u %lf %u %lf %u %u %u %u %c", kind, &views, &lt,
                                        &lp, &ww, &fw, &flat, &bw, &tiled, &depth, &order, &tail);
            if ((got != 10 && got != 11) ||
                (std::strcmp(kind, "forward") != 0 && std::strcmp(kind, "forward_out") != 0 &&
                 std::strcmp(kind, "backward") != 0) ||
                !(ww > 0.0) || !(flat > 0.0) || fw == 0 || bw == 0 || tiled > 1 || depth == 0 || depth > 8 ||
                order > 1)
                raise(PDR_INVALID_ARGUMENT, "autotune cache line is malformed: " + line);
            pdr::TuneParams t;
            t.wave_work = static_cast<float>(ww);
            t.fwd_waves = fw;
            t.flat_work = static_cast<float>(flat);
            t.bwd_waves = bw;
            t.tiled = tiled;
            t.bin_tiles = depth;
            t.order = order;
            parsed[std::string(kind) + " " + std::to_string(views) + " " + std::to_string(lt) + " " +
                   std::to_string(lp)] = t;
        }
        Lock lock(ctx);
        for (const auto& kv : parsed) ctx->tuned[kv.first] = kv.second;
    });
}

pdr_status pdr_autotune_clear(pdr_context ctx) {
    return guarded([&] {
        require_ctx(ctx);
        Lock lock(ctx);
        ctx->tuned.clear();
    });
}

pdr_status pdr_synchronize(pdr_context ctx, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        sync(static_cast<hipStream_t>(stream));
        sync(ctx->stream);
    });
}

pdr_status pdr_rasterize_forward(pdr_context ctx, const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                                 uint32_t W, uint32_t H, uint32_t cull, int want_deriv, const float* viewport4,
                                 float* raster_out, float* deriv_out, float* viewport_out4, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        validate_mesh_counts(pos, V, idx, T);
        require_ptr(raster_out, "raster_out");
        if (!want_deriv) deriv_out = nullptr;
        const bool device = classify({pos, idx, raster_out, deriv_out});
        if (!device)
            check_indices_host(idx, static_cast<size_t>(T) * 3, V, PDR_OUT_OF_RANGE,
                               "triangle_indices contains an invalid vertex index");
        if (W == 0 || H == 0) raise(PDR_INVALID_ARGUMENT, "Rasterization width and height must be positive");
        if (cull > 2) raise(PDR_INVALID_ARGUMENT, "cull_mode must be 'none', 'back', or 'front'");
        float vp[4];
        resolve_viewport(W, H, viewport4, vp);
        if (viewport_out4) std::memcpy(viewport_out4, vp, sizeof(vp));
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const size_t P = static_cast<size_t>(W) * H;
        const float* dpos = pos;
        const uint32_t* didx = idx;
        float* draster = raster_out;
        float* dderiv = deriv_out;
        if (device) {
            check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                                 "triangle_indices contains an invalid vertex index");
        } else {
            dpos = stage_in(ctx, "h_pos", pos, static_cast<size_t>(V) * 4, st);
            didx = stage_in(ctx, "h_idx", idx, static_cast<size_t>(T) * 3, st);
            draster = ctx->get_t<float>("h_raster", P * 4);
            dderiv = deriv_out ? ctx->get_t<float>("h_deriv", P * 4) : nullptr;
        }
        pdr::ViewTable vt{};
        vt.n = 1;
        vt.v[0] = make_view(dpos, W, H, vp);
        vt.v[0].raster = draster;
        vt.v[0].deriv = dderiv;
        finalize_table(vt);
        run_raster_forward(ctx, vt, didx, T, V, cull, want_deriv, st);
        if (!device) {
            copy_out(raster_out, draster, P * 16, st);
            copy_out(deriv_out, dderiv, P * 16, st);
            sync(st);
        }
    });
}

pdr_status pdr_rasterize_forward_batched(pdr_context ctx, const pdr_raster_view* views, uint32_t nviews, uint32_t V,
                                         const uint32_t* idx, uint32_t T, uint32_t cull, int want_deriv,
                                         void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        if (!views || nviews == 0) raise(PDR_INVALID_ARGUMENT, "views must contain at least one view");
        if (V == 0) raise(PDR_INVALID_ARGUMENT, "clip_positions must have shape [vertex_count, 4]");
        if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices must have shape [triangle_count, 3]");
        if (cull > 2) raise(PDR_INVALID_ARGUMENT, "cull_mode must be 'none', 'back', or 'front'");
        for (uint32_t k = 0; k < nviews; ++k) {
            if (!views[k].positions || !views[k].raster)
                raise(PDR_INVALID_ARGUMENT, "each view needs positions and a raster output");
            if (views[k].width == 0 || views[k].height == 0)
                raise(PDR_INVALID_ARGUMENT, "Rasterization width and height must be positive");
            if (!classify({views[k].positions, views[k].raster, idx, want_deriv ? views[k].derivatives : nullptr}))
                raise(PDR_INVALID_ARGUMENT, "batched rasterization requires device memory");
        }
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(true, stream);
        check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                             "triangle_indices contains an invalid vertex index");
        for (uint32_t base = 0; base < nviews; base += pdr::kMaxViews) {
            pdr::ViewTable vt{};
            vt.n = std::min<uint32_t>(pdr::kMaxViews, nviews - base);
            for (uint32_t k = 0; k < vt.n; ++k) {
                const pdr_raster_view& in = views[base + k];
                float vp[4];
                resolve_viewport(in.width, in.height, in.viewport[2] > 0.0f ? in.viewport : nullptr, vp);
                vt.v[k] = make_view(in.positions, in.width, in.height, vp);
                vt.v[k].raster = in.raster;
                vt.v[k].deriv = want_deriv ? in.derivatives : nullptr;
            }
            finalize_table(vt);
            run_raster_forward(ctx, vt, idx, T, V, cull, want_deriv, st);
        }
    });
}

pdr_status pdr_rasterize_backward(pdr_context ctx, const float* pos, uint32_t V, const uint32_t* idx, uint32_t T,
                                  const float* raster, uint32_t W, uint32_t H, const float* grad_raster,
                                  const float* viewport4, float* grad_pos_out, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        validate_mesh_counts(pos, V, idx, T);
        const bool device = classify({pos, idx, raster, grad_raster, grad_pos_out});
        if (!device)
            check_indices_host(idx, static_cast<size_t>(T) * 3, V, PDR_OUT_OF_RANGE,
                               "triangle_indices contains an invalid vertex index");
        if (W == 0 || H == 0 || !raster) raise(PDR_INVALID_ARGUMENT, "forward_output has inconsistent dimensions");
        if (!grad_raster) raise(PDR_INVALID_ARGUMENT, "grad_raster must have shape [height, width, 4]");
        require_ptr(grad_pos_out, "grad_pos_out");
        float vp[4];
        resolve_viewport(W, H, viewport4, vp);
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const size_t P = static_cast<size_t>(W) * H;
        const float *dpos = pos, *draster = raster, *dgrad = grad_raster;
        const uint32_t* didx = idx;
        float* dout = grad_pos_out;
        if (device) {
            check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                                 "triangle_indices contains an invalid vertex index");
        } else {
            dpos = stage_in(ctx, "h_pos", pos, static_cast<size_t>(V) * 4, st);
            didx = stage_in(ctx, "h_idx", idx, static_cast<size_t>(T) * 3, st);
            draster = stage_in(ctx, "h_raster", raster, P * 4, st);
            dgrad = stage_in(ctx, "h_grad_raster", grad_raster, P * 4, st);
            dout = ctx->get_t<float>("h_grad_pos", static_cast<size_t>(V) * 4);
        }
        pdr::ViewTable vt{};
        vt.n = 1;
        vt.v[0] = make_view(dpos, W, H, vp);
        vt.v[0].raster = const_cast<float*>(draster);
        vt.v[0].grad_raster = dgrad;
        vt.v[0].grad_pos = dout;
        finalize_table(vt);
        run_raster_backward(ctx, vt, didx, T, V, st);
        if (!device) {
            copy_out(grad_pos_out, dout, static_cast<size_t>(V) * 16, st);
            sync(st);
        }
    });
}

pdr_status pdr_rasterize_backward_batched(pdr_context ctx, const pdr_raster_view* views, uint32_t nviews, uint32_t V,
                                          const uint32_t* idx, uint32_t T, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        if (!views || nviews == 0) raise(PDR_INVALID_ARGUMENT, "views must contain at least one view");
        if (V == 0) raise(PDR_INVALID_ARGUMENT, "clip_positions must have shape [vertex_count, 4]");
        if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices must have shape [triangle_count, 3]");
        for (uint32_t k = 0; k < nviews; ++k) {
            const pdr_raster_view& in = views[k];
            if (!in.positions || !in.raster || !in.grad_raster || !in.grad_positions)
                raise(PDR_INVALID_ARGUMENT, "each view needs positions, raster, grad_raster and grad_positions");
            if (in.width == 0 || in.height == 0)
                raise(PDR_INVALID_ARGUMENT, "forward_output has inconsistent dimensions");
            if (!classify({in.positions, in.raster, in.grad_raster, in.grad_positions, idx}))
                raise(PDR_INVALID_ARGUMENT, "batched rasterization requires device memory");
        }
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(true, stream);
        check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                             "triangle_indices contains an invalid vertex index");
        for (uint32_t base = 0; base < nviews; base += pdr::kMaxViews) {
            pdr::ViewTable vt{};
            vt.n = std::min<uint32_t>(pdr::kMaxViews, nviews - base);
            for (uint32_t k = 0; k < vt.n; ++k) {
                const pdr_raster_view& in = views[base + k];
                float vp[4];
                resolve_viewport(in.width, in.height, in.viewport[2] > 0.0f ? in.viewport : nullptr, vp);
                vt.v[k] = make_view(in.positions, in.width, in.height, vp);
                vt.v[k].raster = in.raster;
                vt.v[k].grad_raster = in.grad_raster;
                vt.v[k].grad_pos = in.grad_positions;
            }
            finalize_table(vt);
            run_raster_backward(ctx, vt, idx, T, V, st);
        }
    });
}

pdr_status pdr_interpolate_forward(pdr_context ctx, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                   uint32_t T, const float* raster, uint32_t W, uint32_t H, float* out,
                                   void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        // src/rasterizer.cpp:544-560
        if (A == 0 || V == 0 || !attr)
            raise(PDR_INVALID_ARGUMENT, "vertex_attributes must have shape [vertex_count, attribute_count]");
        if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices must have shape [triangle_count, 3]");
        const bool device = classify({attr, idx, raster, out});
        if (!device)
            check_indices_host(idx, static_cast<size_t>(T) * 3, V, PDR_OUT_OF_RANGE,
                               "triangle_indices contains an invalid attribute vertex index");
        if (W == 0 || H == 0 || !raster) raise(PDR_INVALID_ARGUMENT, "raster_output has inconsistent dimensions");
        const uint64_t P64 = static_cast<uint64_t>(W) * H;
        if (P64 > 0xFFFFFFFFull) raise(PDR_OVERFLOW, "Interpolation pixel count exceeds backend limits");
        require_ptr(out, "out");
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const size_t P = static_cast<size_t>(P64);
        const float *dattr = attr, *draster = raster;
        const uint32_t* didx = idx;
        float* dout = out;
        if (device) {
            check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                                 "triangle_indices contains an invalid attribute vertex index");
        } else {
            dattr = stage_in(ctx, "h_attr", attr, static_cast<size_t>(V) * A, st);
            didx = stage_in(ctx, "h_idx", idx, static_cast<size_t>(T) * 3, st);
            draster = stage_in(ctx, "h_raster", raster, P * 4, st);
            dout = ctx->get_t<float>("h_out", P * A);
        }
        const pdr::PixelTable pt = pixel_table_one(draster, nullptr, nullptr, dout, nullptr, W, H);
        ctx->timed("interp_fwd", st,
                   [&] { hip_check(pdr::interp_forward(pt, dattr, V, A, didx, T, st), "interp_forward"); });
        if (!device) {
            copy_out(out, dout, P * A * 4, st);
            sync(st);
        }
    });
}

// Scratch of the ordered (PDR_FLAG_DETERMINISTIC) gradient gathers: `per` entries
// per pixel, keys 0 .. nkeys.
static pdr::DeterministicScratch det_scratch(pdr_context ctx, uint64_t pixels, uint32_t per, uint32_t nkeys) {
    if (pixels * 4ull >= 0xFFFFFFFFull)
        raise(PDR_OVERFLOW, "deterministic gradients support at most 2^30 pixels per call");
    const uint32_t E = static_cast<uint32_t>(pixels * per);
    pdr::DeterministicScratch s;
    s.keys_in = ctx->get_t<uint32_t>("det_keys_in", E ? E : 1);
    s.keys_out = ctx->get_t<uint32_t>("det_keys_out", E ? E : 1);
    s.vals_in = ctx->get_t<uint32_t>("det_vals_in", E ? E : 1);
    s.vals_out = ctx->get_t<uint32_t>("det_vals_out", E ? E : 1);
    s.starts = ctx->get_t<uint32_t>("det_starts", static_cast<size_t>(nkeys) + 1);
    s.temp_bytes = pdr::deterministic_temp_bytes(E ? E : 1);
    s.temp = ctx->get_t<unsigned char>("det_temp", s.temp_bytes ? s.temp_bytes : 1);
    return s;
}

pdr_status pdr_interpolate_backward(pdr_context ctx, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                    uint32_t T, const float* raster, uint32_t W, uint32_t H, const float* grad,
                                    float* grad_attr, float* grad_raster, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        // src/rasterizer.cpp:609-627
        if (A == 0 || V == 0 || !attr)
            raise(PDR_INVALID_ARGUMENT, "vertex_attributes must have shape [vertex_count, attribute_count]");
        if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices is incompatible with vertex_attributes");
        const bool device = classify({attr, idx, raster, grad, grad_attr, grad_raster});
        if (!device)
            check_indices_host(idx, static_cast<size_t>(T) * 3, V, PDR_INVALID_ARGUMENT,
                               "triangle_indices is incompatible with vertex_attributes");
        if (W == 0 || H == 0 || !raster) raise(PDR_INVALID_ARGUMENT, "raster_output has inconsistent dimensions");
        if (!grad) raise(PDR_INVALID_ARGUMENT, "grad_interpolated must have shape [height, width, attribute_count]");
        const uint64_t P64 = static_cast<uint64_t>(W) * H;
        if (P64 > 0xFFFFFFFFull) raise(PDR_OVERFLOW, "Interpolation pixel count exceeds backend limits");
        require_ptr(grad_attr, "grad_attr");
        require_ptr(grad_raster, "grad_raster");
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const size_t P = static_cast<size_t>(P64);
        const float *dattr = attr, *draster = raster, *dgrad = grad;
        const uint32_t* didx = idx;
        float *dga = grad_attr, *dgr = grad_raster;
        if (device) {
            check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_INVALID_ARGUMENT,
                                 "triangle_indices is incompatible with vertex_attributes");
        } else {
            dattr = stage_in(ctx, "h_attr", attr, static_cast<size_t>(V) * A, st);
            didx = stage_in(ctx, "h_idx", idx, static_cast<size_t>(T) * 3, st);
            draster = stage_in(ctx, "h_raster", raster, P * 4, st);
            dgrad = stage_in(ctx, "h_grad", grad, P * A, st);
            dga = ctx->get_t<float>("h_grad_attr", static_cast<size_t>(V) * A);
            dgr = ctx->get_t<float>("h_grad_raster", P * 4);
        }
        const pdr::PixelTable pt = pixel_table_one(draster, dgrad, nullptr, nullptr, dgr, W, H);
        hip_check(pdr::fill_f32(dga, 0.0f, static_cast<size_t>(V) * A, st), "fill");
        ctx->timed("interp_bwd", st,
                   [&] { hip_check(pdr::interp_backward(pt, dattr, V, A, didx, T, dga, st), "interp_backward"); });
        if (ctx->flags & PDR_FLAG_DETERMINISTIC) {
            const pdr::DeterministicScratch ds = det_scratch(ctx, P, 3, V);
            ctx->timed("interp_bwd_det", st, [&] {
                hip_check(pdr::interp_backward_deterministic(pt, V, A, didx, T, dga, ds, st),
                          "interp_backward_deterministic");
            });
        }
        if (!device) {
            copy_out(grad_attr, dga, static_cast<size_t>(V) * A * 4, st);
            copy_out(grad_raster, dgr, P * 16, st);
            sync(st);
        }
    });
}

static pdr::PixelTable pixel_table_many(const pdr_interp_view* views, uint32_t n, bool backward) {
    pdr::PixelTable t{};
    t.n = n;
    uint32_t tiles = 0, pixels = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const pdr_interp_view& in = views[k];
        pdr::PixelView& v = t.v[k];
        v.raster = in.raster;
        v.in0 = backward ? in.grad : nullptr;
        v.out0 = backward ? nullptr : in.out;
        v.out1 = backward ? in.grad_raster : nullptr;
        v.width = in.width;
        v.height = in.height;
        v.tiles_x = (in.width + pdr::kGradTile - 1) / pdr::kGradTile;
        v.tile_base = tiles;
        v.pixel_base = pixels;
        const uint64_t nt = tiles + static_cast<uint64_t>(v.tiles_x) * ((in.height + pdr::kGradTile - 1) / pdr::kGradTile);
        const uint64_t np = pixels + static_cast<uint64_t>(in.width) * in.height;
        if (nt > 0x7FFFFFFFu || np > 0xFFFFFFFFu) raise(PDR_OVERFLOW, "Interpolation pixel count exceeds backend limits");
        tiles = static_cast<uint32_t>(nt);
        pixels = static_cast<uint32_t>(np);
    }
    t.total_tiles = tiles;
    t.total_pixels = pixels;
    return t;
}

pdr_status pdr_interpolate_forward_batched(pdr_context ctx, const pdr_interp_view* views, uint32_t nviews,
                                           const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                           uint32_t T, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        if (A == 0 || V == 0 || !attr)
            raise(PDR_INVALID_ARGUMENT, "vertex_attributes must have shape [vertex_count, attribute_count]");
        if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices must have shape [triangle_count, 3]");
        if (!views || nviews == 0) raise(PDR_INVALID_ARGUMENT, "views must contain at least one view");
        for (uint32_t k = 0; k < nviews; ++k) {
            if (!views[k].raster || !views[k].out || views[k].width == 0 || views[k].height == 0)
                raise(PDR_INVALID_ARGUMENT, "raster_output has inconsistent dimensions");
            if (!classify({attr, idx, views[k].raster, views[k].out}))
                raise(PDR_INVALID_ARGUMENT, "batched interpolation requires device memory");
        }
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(true, stream);
        check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_OUT_OF_RANGE,
                             "triangle_indices contains an invalid attribute vertex index");
        for (uint32_t base = 0; base < nviews; base += pdr::kMaxViews) {
            const uint32_t n = std::min<uint32_t>(pdr::kMaxViews, nviews - base);
            const pdr::PixelTable pt = pixel_table_many(views + base, n, false);
            ctx->timed("interp_fwd", st,
                       [&] { hip_check(pdr::interp_forward(pt, attr, V, A, idx, T, st), "interp_forward"); });
        }
    });
}

pdr_status pdr_interpolate_backward_batched(pdr_context ctx, const pdr_interp_view* views, uint32_t nviews,
                                            const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                                            uint32_t T, float* grad_attr, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        if (A == 0 || V == 0 || !attr)
            raise(PDR_INVALID_ARGUMENT, "vertex_attributes must have shape [vertex_count, attribute_count]");
        if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices is incompatible with vertex_attributes");
        if (!views || nviews == 0) raise(PDR_INVALID_ARGUMENT, "views must contain at least one view");
        require_ptr(grad_attr, "grad_attr");
        for (uint32_t k = 0; k < nviews; ++k) {
            if (!views[k].raster || !views[k].grad || !views[k].grad_raster || views[k].width == 0 ||
                views[k].height == 0)
                raise(PDR_INVALID_ARGUMENT, "raster_output has inconsistent dimensions");
            if (!classify({attr, idx, views[k].raster, views[k].grad, views[k].grad_raster, grad_attr}))
                raise(PDR_INVALID_ARGUMENT, "batched interpolation requires device memory");
        }
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(true, stream);
        check_indices_device(ctx, idx, static_cast<size_t>(T) * 3, V, st, PDR_INVALID_ARGUMENT,
                             "triangle_indices is incompatible with vertex_attributes");
        if (nviews > pdr::kMaxViews)
            raise(PDR_INVALID_ARGUMENT, "batched interpolation backward takes at most 32 views per call");
        const pdr::PixelTable pt = pixel_table_many(views, nviews, true);
        hip_check(pdr::fill_f32(grad_attr, 0.0f, static_cast<size_t>(V) * A, st), "fill");
        ctx->timed("interp_bwd", st, [&] {
            hip_check(pdr::interp_backward(pt, attr, V, A, idx, T, grad_attr, st), "interp_backward");
        });
        if (ctx->flags & PDR_FLAG_DETERMINISTIC) {
            const pdr::DeterministicScratch ds = det_scratch(ctx, pt.total_pixels, 3, V);
            ctx->timed("interp_bwd_det", st, [&] {
                hip_check(pdr::interp_backward_deterministic(pt, V, A, idx, T, grad_attr, ds, st),
                          "interp_backward_deterministic");
            });
        }
    });
}

pdr_status pdr_texture_forward(pdr_context ctx, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                               uint32_t mode, const float* uv, const float* raster, uint32_t W, uint32_t H,
                               float* out, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        const uint32_t P = validate_texture(tw, th, C, W, H);
        if (static_cast<uint64_t>(P) * C > 0xFFFFFFFFull)
            raise(PDR_OVERFLOW, "Sampled texture output exceeds backend element limits");
        if (mode > 2) raise(PDR_INVALID_ARGUMENT, "address_mode must be 'clamp', 'wrap', or 'mirror'");
        require_ptr(tex, "texture");
        require_ptr(uv, "uv");
        require_ptr(raster, "raster");
        require_ptr(out, "out");
        const bool device = classify({tex, uv, raster, out});
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const float *dtex = tex, *duv = uv, *draster = raster;
        float* dout = out;
        if (!device) {
            dtex = stage_in(ctx, "h_tex", tex, static_cast<size_t>(tw) * th * C, st);
            duv = stage_in(ctx, "h_uv", uv, static_cast<size_t>(P) * 2, st);
            draster = stage_in(ctx, "h_raster", raster, static_cast<size_t>(P) * 4, st);
            dout = ctx->get_t<float>("h_out", static_cast<size_t>(P) * C);
        }
        const pdr::PixelTable pt = pixel_table_one(draster, duv, nullptr, dout, nullptr, W, H);
        ctx->timed("texture_fwd", st,
                   [&] { hip_check(pdr::texture_forward(pt, dtex, tw, th, C, mode, st), "texture_forward"); });
        if (!device) {
            copy_out(out, dout, static_cast<size_t>(P) * C * 4, st);
            sync(st);
        }
    });
}

pdr_status pdr_texture_backward(pdr_context ctx, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                                uint32_t mode, const float* uv, const float* raster, uint32_t W, uint32_t H,
                                const float* grad, int compute_uv, float* grad_tex, float* grad_uv,
                                void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        const uint32_t P = validate_texture(tw, th, C, W, H);
        if (!grad)
            raise(PDR_INVALID_ARGUMENT, "grad_sampled must have shape [image_height, image_width, channel_count]");
        if (static_cast<uint64_t>(P) * C > 0xFFFFFFFFull || static_cast<uint64_t>(tw) * th * C > 0xFFFFFFFFull ||
            P > 0xFFFFFFFFu / 4)
            raise(PDR_OVERFLOW, "Texture gradient operation exceeds backend element limits");
        if (mode > 2) raise(PDR_INVALID_ARGUMENT, "address_mode must be 'clamp', 'wrap', or 'mirror'");
        require_ptr(tex, "texture");
        require_ptr(uv, "uv");
        require_ptr(raster, "raster");
        require_ptr(grad_tex, "grad_tex");
        if (!compute_uv) grad_uv = nullptr;
        if (compute_uv) require_ptr(grad_uv, "grad_uv");
        const bool device = classify({tex, uv, raster, grad, grad_tex, grad_uv});
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const float *dtex = tex, *duv = uv, *draster = raster, *dgrad = grad;
        float *dgt = grad_tex, *dgu = grad_uv;
        if (!device) {
            dtex = stage_in(ctx, "h_tex", tex, static_cast<size_t>(tw) * th * C, st);
            duv = stage_in(ctx, "h_uv", uv, static_cast<size_t>(P) * 2, st);
            draster = stage_in(ctx, "h_raster", raster, static_cast<size_t>(P) * 4, st);
            dgrad = stage_in(ctx, "h_grad", grad, static_cast<size_t>(P) * C, st);
            dgt = ctx->get_t<float>("h_grad_tex", static_cast<size_t>(tw) * th * C);
            dgu = grad_uv ? ctx->get_t<float>("h_grad_uv", static_cast<size_t>(P) * 2) : nullptr;
        }
        const pdr::PixelTable pt = pixel_table_one(draster, duv, dgrad, nullptr, dgu, W, H);
        if (!device || (ctx->flags & PDR_FLAG_VALIDATE_DEVICE_INPUTS)) {
            // src/rasterizer.cpp:198-211, first offending covered pixel decides
            uint32_t* flags = ctx->get_t<uint32_t>("uv_flags", 2);
            hip_check(pdr::fill_u32(flags, 0xFFFFFFFFu, 2, st), "fill");
            hip_check(pdr::texture_check_uv(pt, tw, th, flags, st), "texture_check_uv");
            hip_check(hipMemcpyAsync(ctx->pinned, flags, 8, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            sync(st);
            const uint32_t bad_finite = ctx->pinned[0], bad_range = ctx->pinned[1];
            if (bad_finite != 0xFFFFFFFFu || bad_range != 0xFFFFFFFFu) {
                if (bad_finite <= bad_range)
                    raise(PDR_INVALID_ARGUMENT, "uv contains a non-finite value in a covered pixel");
                raise(PDR_OUT_OF_RANGE, "uv is too large for portable texture addressing");
            }
        }
        hip_check(pdr::fill_f32(dgt, 0.0f, static_cast<size_t>(tw) * th * C, st), "fill");
        ctx->timed("texture_bwd", st, [&] {
            hip_check(pdr::texture_backward(pt, dtex, tw, th, C, mode, compute_uv, dgt, st), "texture_backward");
        });
        if (ctx->flags & PDR_FLAG_DETERMINISTIC) {
            const pdr::DeterministicScratch ds = det_scratch(ctx, static_cast<uint64_t>(W) * H, 4, tw * th);
            ctx->timed("texture_bwd_det", st, [&] {
                hip_check(pdr::texture_backward_deterministic(pt, tw, th, C, mode, dgt, ds, st),
                          "texture_backward_deterministic");
            });
        }
        if (!device) {
            copy_out(grad_tex, dgt, static_cast<size_t>(tw) * th * C * 4, st);
            copy_out(grad_uv, dgu, static_cast<size_t>(P) * 8, st);
            sync(st);
        }
    });
}

}  // extern "C"
