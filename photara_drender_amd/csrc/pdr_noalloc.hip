// pdr_rasterize_forward_noalloc: the single-view rasterize forward for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It runs the same kernels as
// pdr_rasterize_forward (raster.hip: k_raster_small / k_raster_large,
// k_raster_resolve) on device arrays the caller owns, with their scratch -
// the per-pixel (depth, id) keys, the large-triangle queue and its counters -
// carved out of one
// file-scope device array of kScratchBytes per GPU instead of the context's
// grow-only pool.  It calls no hipMalloc / hipFree and no allocating library
// routine.
//
// Every call on one GPU shares that GPU's array (it keeps the last call's
// contents; the keys are cleared to all-ones and the counters to zero by fill
// kernels at the start of each call), so calls that overlap on one GPU - on
// different streams, or from different contexts of that device - race.
#include <array>
#include <cstring>
#include <mutex>

#include "pdr_host.hpp"

using namespace pdrh;

namespace {
constexpr size_t kScratchBytes = size_t{32} << 20;  // 32 MiB per GPU, held until the process ends
constexpr int kMaxDevices = 64;
}  // namespace

// The scratch array itself.  In host code its name is a host-side stand-in, not
// a device address: only the address from hipGetSymbolAddress (g_scratch) is
// handed to kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mutex;
std::array<unsigned char*, kMaxDevices> g_scratch{};  // device address of `scratch` on each GPU

inline size_t up256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Scratch layout of one call: keys [view_key_count(W, H)] u64 (16 per 4 x 4
// pixel tile, rounded up to 64), large queue [T] uint4, counters [4] u32, each
// 256-byte aligned.
struct Layout {
    size_t keys = 0, queue = 0, counters = 0, total = 0;
};
inline Layout layout_of(uint32_t W, uint32_t H, uint32_t T) {
    Layout l;
    size_t off = 0;
    l.keys = off;
    off += up256(pdr::view_key_count(W, H) * sizeof(unsigned long long));
    l.queue = off;
    off += up256(static_cast<size_t>(T) * sizeof(uint4));
    l.counters = off;
    off += up256(4 * sizeof(uint32_t));
    l.total = off;
    return l;
}

}  // namespace

namespace pdr {

// Device address of `scratch` on `device`, taken once per device with that
// device current (from pdr_context_create).
void noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) raise(PDR_OUT_OF_RANGE, "device_index is out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch[device]) return;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    g_scratch[device] = static_cast<unsigned char*>(p);
}

}  // namespace pdr

extern "C" {

pdr_status pdr_rasterize_forward_noalloc(pdr_context ctx, const float* pos, uint32_t V, const uint32_t* idx,
                                         uint32_t T, uint32_t W, uint32_t H, uint32_t cull, int want_deriv,
                                         const float* viewport4, float* raster_out, float* deriv_out,
                                         float* viewport_out4, void* stream) {
    return guarded([&] {
        require_ctx(ctx);
        validate_mesh_counts(pos, V, idx, T);
        require_ptr(raster_out, "raster_out");
        if (!want_deriv) deriv_out = nullptr;
        if (!classify({pos, idx, raster_out, deriv_out}))
            raise(PDR_INVALID_ARGUMENT, "the no-allocation rasterize forward takes device memory only");
        if (W == 0 || H == 0) raise(PDR_INVALID_ARGUMENT, "Rasterization width and height must be positive");
        if (cull > 2) raise(PDR_INVALID_ARGUMENT, "cull_mode must be 'none', 'back', or 'front'");
        const Layout l = layout_of(W, H, T);
        if (l.total > kScratchBytes)
            raise(PDR_OVERFLOW, "the no-allocation rasterize forward needs " + std::to_string(l.total) +
                                    " bytes of scratch (8 per pixel of the size rounded up to multiples of 4, "
                                    "+ 16 per triangle); its limit is " +
                                    std::to_string(kScratchBytes));
        float vp[4];
        resolve_viewport(W, H, viewport4, vp);
        if (viewport_out4) std::memcpy(viewport_out4, vp, sizeof(vp));
        Lock lock(ctx);
        unsigned char* base = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_scratch_mutex);
            base = g_scratch[ctx->device];
        }
        if (!base) raise(PDR_RUNTIME, "the no-allocation scratch of this device was not set up");
        const hipStream_t st = static_cast<hipStream_t>(stream);
        pdr::ViewTable vt{};
        vt.n = 1;
        vt.v[0] = make_view(pos, W, H, vp);
        vt.v[0].raster = raster_out;
        vt.v[0].deriv = deriv_out;
        finalize_table(vt);
        pdr::RasterScratch s;
        s.keys = reinterpret_cast<unsigned long long*>(base + l.keys);
        s.queue = reinterpret_cast<uint4*>(base + l.queue);
        s.counters = reinterpret_cast<uint32_t*>(base + l.counters);
        // Indices are not checked with a readback here: the kernels skip a
        // triangle whose index is out of range, as the shader does
        // (rasterize_forward.hlsl:75-80).  Launch parameters are the defaults
        // (every choice gives the same raster).
        hip_check(pdr::fill_keys(s.keys, pdr::view_key_count(W, H), st), "fill_keys");
        hip_check(pdr::fill_u32(s.counters, 0u, 4, st), "fill");  // the queue's consumer clears it again
        const pdr::TuneParams tp{};
        hip_check(pdr::raster_candidates(vt, idx, T, V, cull, s, tp, st), "raster_candidates");
        hip_check(pdr::raster_resolve(vt, idx, want_deriv, s, st), "raster_resolve");
    });
}

}  // extern "C"
