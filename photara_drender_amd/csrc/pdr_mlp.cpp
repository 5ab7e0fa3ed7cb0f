// Warm-start MLP entry points: inference, training gradient and a fused
// gradient + Adam step over a batch of texels (kernels in mlp.hip).
#include "pdr_host.hpp"

using namespace pdrh;

namespace {

uint32_t checked_count(uint64_t n) {
    if (n > 0x7FFFFFFFull) raise(PDR_OVERFLOW, "texel count exceeds backend element limits");
    return static_cast<uint32_t>(n);
}

void require_aligned(const float* p, const char* name) {
    if (reinterpret_cast<uintptr_t>(p) % 16 != 0) raise(PDR_INVALID_ARGUMENT, std::string(name) + " must be 16-byte aligned");
}

// Gradient partials of one batch; returns the number of waves written.
uint32_t run_gradient(pdr_context ctx, const float* x, const float* target, const float* valid, uint32_t n,
                      const float* w, float** part_out, hipStream_t st) {
    const uint32_t waves = pdr::mlp_grad_waves(n);
    float* part = ctx->get_t<float>("mlp_partials", static_cast<size_t>(waves) * pdr::kMlpPartStride);
    ctx->timed("mlp_grad", st, [&] { hip_check(pdr::mlp_grad(x, target, valid, n, w, part, waves, st), "mlp_grad"); });
    *part_out = part;
    return waves;
}

}  // namespace

extern "C" {

pdr_status pdr_mlp_infer(pdr_context ctx, const float* features, uint64_t count, const float* params, float* out_rgb,
                         void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        const uint32_t n = checked_count(count);
        require_ptr(params, "params");
        if (n == 0) return;
        require_ptr(features, "features");
        require_ptr(out_rgb, "out");
        const bool device = classify({features, params, out_rgb});
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const float *x = features, *w = params;
        float* out = out_rgb;
        if (!device) {
            x = stage_in(ctx, "mlp_x", features, static_cast<size_t>(n) * pdr::kMlpFeatures, st);
            w = stage_in(ctx, "mlp_w", params, pdr::kMlpParams, st);
            out = ctx->get_t<float>("mlp_out", static_cast<size_t>(n) * 3);
        }
        require_aligned(x, "features");
        ctx->timed("mlp_infer", st, [&] { hip_check(pdr::mlp_infer(x, n, w, out, st), "mlp_infer"); });
        if (!device) {
            copy_out(out_rgb, out, static_cast<size_t>(n) * 12, st);
            sync(st);
        }
    });
}

pdr_status pdr_mlp_gradient(pdr_context ctx, const float* features, const float* target, const float* valid,
                            uint64_t count, const float* params, float* grad_out, float* loss_count_out,
                            void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        const uint32_t n = checked_count(count);
        require_ptr(params, "params");
        require_ptr(grad_out, "grad");
        if (n) {
            require_ptr(features, "features");
            require_ptr(target, "target");
        }
        const bool device = classify({features, target, valid, params, grad_out, loss_count_out});
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const float *x = features, *t = target, *v = valid, *w = params;
        float *g = grad_out, *lc = loss_count_out;
        if (!device) {
            x = stage_in(ctx, "mlp_x", features, static_cast<size_t>(n) * pdr::kMlpFeatures, st);
            t = stage_in(ctx, "mlp_target", target, static_cast<size_t>(n) * 3, st);
            v = stage_in(ctx, "mlp_valid", valid, n, st);
            w = stage_in(ctx, "mlp_w", params, pdr::kMlpParams, st);
            g = ctx->get_t<float>("mlp_grad", pdr::kMlpParams);
            lc = loss_count_out ? ctx->get_t<float>("mlp_loss", 2) : nullptr;
        }
        if (n) require_aligned(x, "features");
        float* part = nullptr;
        const uint32_t waves = run_gradient(ctx, x, t, v, n, w, &part, st);
        ctx->timed("mlp_reduce", st, [&] {
            hip_check(pdr::mlp_reduce(part, waves, g, lc, nullptr, nullptr, nullptr, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f,
                                      st),
                      "mlp_reduce");
        });
        if (!device) {
            copy_out(grad_out, g, pdr::kMlpParams * 4, st);
            if (loss_count_out) copy_out(loss_count_out, lc, 8, st);
            sync(st);
        }
    });
}

pdr_status pdr_mlp_adam_step(pdr_context ctx, const float* features, const float* target, const float* valid,
                             uint64_t count, float* params, float* adam_m, float* adam_v, uint32_t step,
                             float learning_rate, float beta1, float beta2, float epsilon, float* loss_count_out,
                             void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        const uint32_t n = checked_count(count);
        require_ptr(params, "params");
        require_ptr(adam_m, "adam_m");
        require_ptr(adam_v, "adam_v");
        if (n) {
            require_ptr(features, "features");
            require_ptr(target, "target");
        }
        if (step == 0) raise(PDR_INVALID_ARGUMENT, "step counts from 1");
        if (!(learning_rate >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) ||
            !(epsilon > 0.0f))
            raise(PDR_INVALID_ARGUMENT, "Adam parameters are invalid");
        const bool device = classify({features, target, valid, params, adam_m, adam_v, loss_count_out});
        Lock lock(ctx);
        const hipStream_t st = ctx->use_stream(device, stream);
        const float *x = features, *t = target, *v = valid;
        float *w = params, *m1 = adam_m, *m2 = adam_v, *lc = loss_count_out;
        if (!device) {
            x = stage_in(ctx, "mlp_x", features, static_cast<size_t>(n) * pdr::kMlpFeatures, st);
            t = stage_in(ctx, "mlp_target", target, static_cast<size_t>(n) * 3, st);
            v = stage_in(ctx, "mlp_valid", valid, n, st);
            w = const_cast<float*>(stage_in(ctx, "mlp_w", static_cast<const float*>(params), pdr::kMlpParams, st));
            m1 = const_cast<float*>(stage_in(ctx, "mlp_m", static_cast<const float*>(adam_m), pdr::kMlpParams, st));
            m2 = const_cast<float*>(stage_in(ctx, "mlp_v", static_cast<const float*>(adam_v), pdr::kMlpParams, st));
            lc = loss_count_out ? ctx->get_t<float>("mlp_loss", 2) : nullptr;
        }
        if (n) require_aligned(x, "features");
        float* part = nullptr;
        const uint32_t waves = run_gradient(ctx, x, t, v, n, w, &part, st);
        const float b1c = static_cast<float>(1.0 - std::pow(static_cast<double>(beta1), step));
        const float b2c = static_cast<float>(1.0 - std::pow(static_cast<double>(beta2), step));
        ctx->timed("mlp_reduce", st, [&] {
            hip_check(pdr::mlp_reduce(part, waves, nullptr, lc, w, m1, m2, learning_rate, beta1, beta2, epsilon, b1c,
                                      b2c, st),
                      "mlp_reduce");
        });
        if (!device) {
            copy_out(params, w, pdr::kMlpParams * 4, st);
            copy_out(adam_m, m1, pdr::kMlpParams * 4, st);
            copy_out(adam_v, m2, pdr::kMlpParams * 4, st);
            if (loss_count_out) copy_out(loss_count_out, lc, 8, st);
            sync(st);
        }
    });
}

}  // extern "C"
