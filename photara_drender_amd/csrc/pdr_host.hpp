// Host-side runtime shared by the libpdr C ABI translation units: the context
// (stream, device scratch, pinned readback area), error plumbing, host staging
// and argument validation with the reference's messages.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "pdr.h"
#include "pdr_internal.hpp"

namespace pdrh {

extern thread_local std::string g_last_error;

struct PdrError {
    pdr_status status;
    std::string message;
};

[[noreturn]] inline void raise(pdr_status s, const std::string& m) { throw PdrError{s, m}; }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) raise(PDR_RUNTIME, std::string(what) + " failed: " + hipGetErrorString(e));
}

template <class F>
pdr_status guarded(F&& f) {
    try {
        f();
        return PDR_OK;
    } catch (const PdrError& e) {
        g_last_error = e.message;
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return PDR_RUNTIME;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PDR_RUNTIME;
    }
}

inline bool is_device_pointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    const hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// All non-null arrays of one call must live on the same side.
inline bool classify(std::initializer_list<const void*> ptrs) {
    int device = -1;
    for (const void* p : ptrs) {
        if (!p) continue;
        const int d = is_device_pointer(p) ? 1 : 0;
        if (device < 0) device = d;
        else if (device != d)
            raise(PDR_INVALID_ARGUMENT, "all arrays of one call must be host memory or all device memory");
    }
    return device == 1;
}

}  // namespace pdrh

struct pdr_context_s {
    int device = 0;
    uint32_t flags = PDR_FLAG_VALIDATE_DEVICE_INPUTS;
    hipStream_t stream = nullptr;
    hipStream_t last_stream = nullptr;
    bool have_last = false;
    hipEvent_t handoff = nullptr;
    std::mutex mutex;
    uint32_t* pinned = nullptr;  // small host-visible readback area
    struct Slot {
        void* ptr = nullptr;
        size_t bytes = 0;
    };
    std::unordered_map<std::string, Slot> scratch;
    bool keys_clean = false;      // raster key buffer is all-ones
    bool counters_clean = false;  // large-work queue counters are zero (the kernel after their consumer clears them)
    bool bwd_clean = false;       // backward won marks and view flags are zero (the vertex reduce clears them)
    std::map<std::string, pdr::TuneParams> tuned;  // autotuner: chosen launch parameters per workload class

    // The last bake's shadow-map keys (each view's raster candidate pass: the
    // winning (depth, id) word per pixel) in the "view_keys" slot, with what
    // they were computed from: the mesh in "view_keys_pos" / "view_keys_idx"
    // and each view's size, resolved viewport and matrix.  A refinement whose
    // view k matches view k here resolves these keys instead of rasterising.
    struct ViewKeys {
        uint32_t T = 0, V = 0;
        std::vector<uint32_t> width, height;
        std::vector<std::array<float, 4>> viewport;
        std::vector<std::array<float, 16>> matrix;
        std::vector<size_t> offset;  // first key of each view in "view_keys"
    } view_keys;

    // Optional per-kernel HIP-event timing (pdr_context_set_timing).
    bool timing = false;
    struct Region {
        std::string name;
        hipEvent_t start, stop;
    };
    std::vector<Region> pending;
    std::vector<hipEvent_t> event_pool;
    std::unordered_map<std::string, std::pair<double, uint64_t>> totals;

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        pdrh::hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    // Runs f (which enqueues kernels on st) between two events when timing is on.
    template <class F>
    void timed(const char* name, hipStream_t st, F&& f) {
        if (!timing) {
            f();
            return;
        }
        Region r{name, take_event(), take_event()};
        pdrh::hip_check(hipEventRecord(r.start, st), "hipEventRecord");
        f();
        pdrh::hip_check(hipEventRecord(r.stop, st), "hipEventRecord");
        pending.push_back(r);
    }

    void resolve_timing() {
        for (Region& r : pending) {
            pdrh::hip_check(hipEventSynchronize(r.stop), "hipEventSynchronize");
            float ms = 0.0f;
            pdrh::hip_check(hipEventElapsedTime(&ms, r.start, r.stop), "hipEventElapsedTime");
            auto& t = totals[r.name];
            t.first += ms;
            t.second += 1;
            event_pool.push_back(r.start);
            event_pool.push_back(r.stop);
        }
        pending.clear();
    }

    // Stream of a call: device-memory calls run on the caller's stream (NULL is
    // the HIP default stream); host-memory calls stage through the context's
    // own stream and synchronise it.  The chosen stream is ordered after the
    // previous owner of the context's scratch.
    hipStream_t use_stream(bool device, void* user) {
        hipStream_t st = device ? static_cast<hipStream_t>(user) : stream;
        if (have_last && last_stream != st) {
            pdrh::hip_check(hipEventRecord(handoff, last_stream), "hipEventRecord");
            pdrh::hip_check(hipStreamWaitEvent(st, handoff, 0), "hipStreamWaitEvent");
        }
        last_stream = st;
        have_last = true;
        return st;
    }

    // Grow-only scratch buffer `key` of at least `bytes`; *fresh (when given)
    // tells whether it was (re)allocated, its old contents lost.
    void* get(const std::string& key, size_t bytes, bool* fresh = nullptr) {
        Slot& s = scratch[key];
        if (bytes == 0) bytes = 16;
        if (fresh) *fresh = s.bytes < bytes;
        if (s.bytes < bytes) {
            if (s.ptr) {
                pdrh::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
                pdrh::hip_check(hipFree(s.ptr), "hipFree");
                s.ptr = nullptr;
                s.bytes = 0;
            }
            const size_t grown = std::max(bytes, s.bytes + s.bytes / 2);
            pdrh::hip_check(hipMalloc(&s.ptr, grown), "hipMalloc");
            s.bytes = grown;
        }
        return s.ptr;
    }
    template <class T>
    T* get_t(const std::string& key, size_t count, bool* fresh = nullptr) {
        return static_cast<T*>(get(key, count * sizeof(T), fresh));
    }

    ~pdr_context_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& kv : scratch)
            if (kv.second.ptr) (void)hipFree(kv.second.ptr);
        for (auto& r : pending) {
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.stop);
        }
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        if (pinned) (void)hipHostFree(pinned);
        if (handoff) (void)hipEventDestroy(handoff);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace pdrh {

struct Lock {
    std::unique_lock<std::mutex> lock;
    explicit Lock(pdr_context ctx) : lock(ctx->mutex) { hip_check(hipSetDevice(ctx->device), "hipSetDevice"); }
};

inline void require_ctx(pdr_context ctx) {
    if (!ctx) raise(PDR_INVALID_ARGUMENT, "context is null");
}

inline void require_ptr(const void* p, const char* name) {
    if (!p) raise(PDR_INVALID_ARGUMENT, std::string(name) + " is null");
}

// src/rasterizer.cpp:48-59
inline void resolve_viewport(uint32_t W, uint32_t H, const float* vp, float out[4]) {
    if (vp) {
        std::memcpy(out, vp, 4 * sizeof(float));
    } else {
        out[0] = 0.0f;
        out[1] = 0.0f;
        out[2] = static_cast<float>(W);
        out[3] = static_cast<float>(H);
    }
    if (!std::isfinite(out[0]) || !std::isfinite(out[1]) || !std::isfinite(out[2]) || !std::isfinite(out[3]) ||
        out[2] <= 0.0f || out[3] <= 0.0f)
        raise(PDR_INVALID_ARGUMENT, "viewport must contain finite values and positive dimensions");
}

// Copies `bytes` from host `src` into scratch `key` and returns the device copy.
template <class T>
const T* stage_in(pdr_context ctx, const std::string& key, const T* src, size_t count, hipStream_t st) {
    if (!src || count == 0) return nullptr;
    T* d = ctx->get_t<T>(key, count);
    hip_check(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    return d;
}

inline void copy_out(void* host, const void* dev, size_t bytes, hipStream_t st) {
    if (!host || !bytes) return;
    hip_check(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
}

inline void sync(hipStream_t st) { hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"); }

inline void check_indices_host(const uint32_t* idx, size_t n, uint32_t V, pdr_status s, const char* msg) {
    for (size_t i = 0; i < n; ++i)
        if (idx[i] >= V) raise(s, msg);
}

inline void check_indices_device(pdr_context ctx, const uint32_t* idx, size_t n, uint32_t V, hipStream_t st, pdr_status s,
                          const char* msg) {
    if (!(ctx->flags & PDR_FLAG_VALIDATE_DEVICE_INPUTS)) return;
    uint32_t* flag = ctx->get_t<uint32_t>("idx_flag", 1);
    hip_check(pdr::fill_u32(flag, 0u, 1, st), "fill");
    hip_check(pdr::check_indices(idx, n, V, flag, st), "check_indices");
    hip_check(hipMemcpyAsync(ctx->pinned, flag, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    sync(st);
    if (ctx->pinned[0]) raise(s, msg);
}

inline pdr::ViewDesc make_view(const float* pos, uint32_t W, uint32_t H, const float vp[4]) {
    pdr::ViewDesc v{};
    v.pos = pos;
    v.width = W;
    v.height = H;
    v.tiles_x = (W + pdr::kTile - 1) / pdr::kTile;
    v.tiles_y = (H + pdr::kTile - 1) / pdr::kTile;
    v.vx = vp[0];
    v.vy = vp[1];
    v.vw = vp[2];
    v.vh = vp[3];
    return v;
}

inline void finalize_table(pdr::ViewTable& t) {
    uint32_t tiles = 0, pixels = 0, rtiles = 0, keys = 0;
    for (uint32_t k = 0; k < t.n; ++k) {
        t.v[k].key_tx = pdr::key_tiles(t.v[k].width);
        t.v[k].key_base = keys;
        const uint64_t nk = static_cast<uint64_t>(keys) + pdr::view_key_count(t.v[k].width, t.v[k].height);
        if (nk > 0xFFFFFFFFu) raise(PDR_OVERFLOW, "Rasterization output size overflow");
        keys = static_cast<uint32_t>(nk);
        t.v[k].rtiles_x = (t.v[k].width + pdr::kRasterTile - 1) / pdr::kRasterTile;
        t.v[k].rtiles_y = (t.v[k].height + pdr::kRasterTile - 1) / pdr::kRasterTile;
        t.v[k].rtile_base = rtiles;
        rtiles += t.v[k].rtiles_x * t.v[k].rtiles_y;
        t.v[k].tile_base = tiles;
        t.v[k].pixel_base = pixels;
        const uint64_t nt = static_cast<uint64_t>(tiles) + static_cast<uint64_t>(t.v[k].tiles_x) * t.v[k].tiles_y;
        const uint64_t np = static_cast<uint64_t>(pixels) + static_cast<uint64_t>(t.v[k].width) * t.v[k].height;
        if (nt > 0x7FFFFFFFu || np > 0xFFFFFFFFu) raise(PDR_OVERFLOW, "Rasterization output size overflow");
        tiles = static_cast<uint32_t>(nt);
        pixels = static_cast<uint32_t>(np);
    }
    t.total_tiles = tiles;
    t.total_pixels = pixels;
    t.total_rtiles = rtiles;
    t.total_keys = keys;
}

inline pdr::PixelTable pixel_table_one(const float* raster, const float* in0, const float* in1, float* out0, float* out1,
                                uint32_t W, uint32_t H) {
    pdr::PixelTable t{};
    t.n = 1;
    pdr::PixelView& v = t.v[0];
    v.raster = raster;
    v.in0 = in0;
    v.in1 = in1;
    v.out0 = out0;
    v.out1 = out1;
    v.width = W;
    v.height = H;
    v.tiles_x = (W + pdr::kGradTile - 1) / pdr::kGradTile;
    v.tile_base = 0;
    v.pixel_base = 0;
    t.total_tiles = v.tiles_x * ((H + pdr::kGradTile - 1) / pdr::kGradTile);
    t.total_pixels = W * H;
    return t;
}

// Scratch of the triangle-parallel raster for one table of views: key buffer
// guaranteed all-ones (the resolve pass restores it).
inline pdr::RasterScratch raster_scratch(pdr_context ctx, const pdr::ViewTable& vt, uint32_t T, hipStream_t st,
                                         bool with_keys) {
    pdr::RasterScratch s;
    s.queue = ctx->get_t<uint4>("large_queue", static_cast<size_t>(vt.n) * T);
    bool fresh = false;
    s.counters = ctx->get_t<uint32_t>("large_counters", 4, &fresh);
    // zero unless the last raster call on this context ran to the end (its queue consumers cleared them)
    if (fresh || !ctx->counters_clean) hip_check(pdr::fill_u32(s.counters, 0u, 4, st), "fill");
    ctx->counters_clean = false;  // set again by the caller once its kernels are enqueued
    if (with_keys) {
        const size_t before = ctx->scratch["raster_keys"].bytes;
        s.keys = ctx->get_t<unsigned long long>("raster_keys", vt.total_keys);
        const size_t after = ctx->scratch["raster_keys"].bytes;
        if (after != before || !ctx->keys_clean) hip_check(pdr::fill_keys(s.keys, after / 8, st), "fill_keys");
    }
    return s;
}

// ---------------------------------------------------------------- autotuner
// Workload class of a raster call: its kind, view count and the log2 buckets
// of its triangle and pixel counts.
inline std::string tune_key(const char* kind, const pdr::ViewTable& vt, uint32_t T) {
    auto lg = [](uint64_t x) {
        int b = 0;
        while (b < 63 && (1ull << b) < x) ++b;
        return b;
    };
    return std::string(kind) + " " + std::to_string(vt.n) + " " + std::to_string(lg(T)) + " " +
           std::to_string(lg(vt.total_pixels));
}

inline bool stream_capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return cs != hipStreamCaptureStatusNone;
}

// Launch parameters of a call: the cached choice of its class; else, with
// PDR_FLAG_AUTOTUNE set and the stream not being captured, every candidate is
// run (after one warm-up run) and timed three times with HIP events, and the
// fastest (best of its three) is cached for the class unless the defaults
// (the first candidate) are within 3 % of it; else the defaults.  `run`
// enqueues the tuned work on st (each candidate produces the same results).
template <class Run>
pdr::TuneParams tuned_params(pdr_context ctx, const std::string& key, const std::vector<pdr::TuneParams>& cands,
                             hipStream_t st, Run&& run) {
    const auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) return it->second;
    if (!(ctx->flags & PDR_FLAG_AUTOTUNE) || stream_capturing(st)) return pdr::TuneParams{};
    hipEvent_t a = ctx->take_event(), b = ctx->take_event();
    run(cands[0]);
    float best = 0.0f, first = 0.0f;
    pdr::TuneParams choice = cands[0];
    for (size_t c = 0; c < cands.size(); ++c) {
        float fastest = 0.0f;
        for (int rep = 0; rep < 3; ++rep) {
            hip_check(hipEventRecord(a, st), "hipEventRecord");
            run(cands[c]);
            hip_check(hipEventRecord(b, st), "hipEventRecord");
            hip_check(hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            fastest = rep == 0 ? ms : std::min(fastest, ms);
        }
        if (c == 0) first = fastest;
        if (c == 0 || fastest < best) {
            best = fastest;
            choice = cands[c];
        }
    }
    if (first <= best * 1.03f) choice = cands[0];
    ctx->event_pool.push_back(a);
    ctx->event_pool.push_back(b);
    ctx->tuned[key] = choice;
    return choice;
}

// Candidates; the first is the defaults.
// Forward candidates.  With outputs (tiled_choice): the tiled path at bin
// depths of 2, 4 and 8 raster tiles per triangle, and the triangle-parallel
// path with its walk limit, wave count and traversal order (index order, or
// each view's nearer end first); the keys-only pass (the bake's shadow maps)
// has the triangle-parallel candidates alone.
inline std::vector<pdr::TuneParams> forward_candidates(bool tiled_choice = false) {
    std::vector<pdr::TuneParams> out;
    for (uint32_t order : {0u, 1u})
        for (float work : {16384.0f, 4096.0f, 65536.0f})
            for (uint32_t waves : {4096u, 2048u, 8192u}) {
                pdr::TuneParams p;
                p.tiled = 0;
                p.wave_work = work;
                p.fwd_waves = waves;
                p.order = order;
                out.push_back(p);
            }
    if (tiled_choice)
        for (uint32_t depth : {4u, 2u, 8u}) {
            pdr::TuneParams p;
            p.tiled = 1;
            p.bin_tiles = depth;
            out.push_back(p);
        }
    return out;
}

// The bins of the tiled forward (grow-only context scratch): capacity
// bin_tiles * views * T triangle entries.
inline void tiled_scratch(pdr_context ctx, const pdr::ViewTable& vt, uint32_t T, uint32_t bin_tiles,
                          pdr::RasterScratch& s) {
    const size_t tiles = static_cast<size_t>(vt.total_rtiles) + 1;
    s.bin_counts = ctx->get_t<uint32_t>("bin_counts", tiles);
    s.bin_offsets = ctx->get_t<uint32_t>("bin_offsets", tiles);
    s.bin_seeded = reinterpret_cast<uint8_t*>(ctx->get_t<uint32_t>("bin_seeded", (tiles + 3) / 4));
    const uint64_t cap = static_cast<uint64_t>(bin_tiles) * vt.n * T;
    if (cap > 0xFFFFFFFFull) raise(PDR_OVERFLOW, "tiled rasterization supports at most 2^32 binned triangles per call");
    s.bin_list = ctx->get_t<uint32_t>("bin_list", static_cast<size_t>(cap ? cap : 1));
    s.medium_list = ctx->get_t<uint32_t>("medium_list", static_cast<size_t>(vt.n) * T);
    s.medium_count = ctx->get_t<uint32_t>("medium_count", pdr::kMaxViews);
    s.tri_records = ctx->get("tri_records", static_cast<size_t>(vt.n) * T * pdr::tri_record_bytes());
    s.tri_spans = ctx->get_t<unsigned long long>("tri_spans", static_cast<size_t>(vt.n) * T);
    s.bin_temp_bytes = pdr::bin_scan_temp_bytes(static_cast<uint32_t>(tiles));
    s.bin_temp = ctx->get("bin_temp", s.bin_temp_bytes);
}

inline std::vector<pdr::TuneParams> backward_candidates() {
    std::vector<pdr::TuneParams> out{pdr::TuneParams{}};
    for (float work : {256.0f, 512.0f, 1024.0f, 2048.0f})
        for (uint32_t waves : {2048u, 4096u, 8192u, 16384u}) {
            if (work == out[0].flat_work && waves == out[0].bwd_waves) continue;
            pdr::TuneParams p;
            p.flat_work = work;
            p.bwd_waves = waves;
            out.push_back(p);
        }
    return out;
}

// Raster forward of one table of views (device pointers), on stream st: the
// triangle-parallel path (global keys, then a resolve pass) or the tiled one
// (bins, LDS keys per raster tile), as the tuned parameters of the workload
// class say (the triangle-parallel path by default).
inline void run_raster_forward(pdr_context ctx, const pdr::ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V,
                               uint32_t cull, int want_deriv, hipStream_t st) {
    if (!vt.total_pixels) return;
    pdr::RasterScratch s = raster_scratch(ctx, vt, T, st, true);
    ctx->keys_clean = false;
    auto run = [&](const pdr::TuneParams& c, bool timed) {
        if (c.tiled) {
            tiled_scratch(ctx, vt, T, c.bin_tiles, s);
            auto f = [&] { hip_check(pdr::raster_tiled(vt, idx, T, V, cull, want_deriv, s, c, st), "raster_tiled"); };
            if (timed)
                ctx->timed("raster_tiled", st, f);
            else
                f();
            return;
        }
        auto cand = [&] { hip_check(pdr::raster_candidates(vt, idx, T, V, cull, s, c, st), "raster_candidates"); };
        auto res = [&] { hip_check(pdr::raster_resolve(vt, idx, want_deriv, s, st), "raster_resolve"); };
        if (timed) {
            ctx->timed("raster_candidates", st, cand);
            ctx->timed("raster_resolve", st, res);  // resets the keys
        } else {
            cand();
            res();
        }
    };
    const pdr::TuneParams tp = tuned_params(ctx, tune_key("forward_out", vt, T), forward_candidates(true), st,
                                            [&](const pdr::TuneParams& c) { run(c, false); });
    run(tp, true);
    ctx->keys_clean = true;
    ctx->counters_clean = true;
}

// The forward's resolve pass over keys taken from `src_keys` (view j's keys at
// src_keys + offsets[j]), e.g. a bake's cached shadow-map keys of the same
// views: the same raster as run_raster_forward (the winner is the same
// lexicographic minimum), without the candidate pass.
inline void run_raster_resolve_keys(pdr_context ctx, const pdr::ViewTable& vt, const uint32_t* idx, uint32_t T,
                                    int want_deriv, const unsigned long long* src_keys,
                                    const std::vector<size_t>& offsets, hipStream_t st) {
    if (!vt.total_pixels) return;
    pdr::RasterScratch s = raster_scratch(ctx, vt, T, st, true);
    ctx->keys_clean = false;
    for (uint32_t j = 0; j < vt.n; ++j)
        hip_check(hipMemcpyAsync(s.keys + vt.v[j].key_base, src_keys + offsets[j],
                                 pdr::view_key_count(vt.v[j].width, vt.v[j].height) * sizeof(unsigned long long),
                                 hipMemcpyDeviceToDevice, st),
                  "hipMemcpyAsync");
    ctx->timed("raster_resolve", st, [&] {
        hip_check(pdr::raster_resolve(vt, idx, want_deriv, s, st), "raster_resolve");  // resets the keys
    });
    ctx->keys_clean = true;
    ctx->counters_clean = true;
}

// The forward's candidate pass alone: each pixel's winning (depth key, id)
// word stays in the raster scratch keys (index key_base + key_index), ~0 where
// no triangle covers the pixel.  They stay marked as used, so the next raster
// call on this context clears them first.
inline const unsigned long long* run_raster_keys(pdr_context ctx, const pdr::ViewTable& vt, const uint32_t* idx,
                                                 uint32_t T, uint32_t V, uint32_t cull, hipStream_t st) {
    if (!vt.total_pixels) return nullptr;
    const pdr::RasterScratch s = raster_scratch(ctx, vt, T, st, true);
    ctx->keys_clean = false;
    const pdr::TuneParams tp =
        tuned_params(ctx, tune_key("forward", vt, T), forward_candidates(), st, [&](const pdr::TuneParams& c) {
            hip_check(pdr::raster_candidates(vt, idx, T, V, cull, s, c, st), "raster_candidates");
            // the views may have no raster output here: clear the keys (and the queue counters, which a
            // resolve pass would clear) for the next candidate
            hip_check(pdr::fill_keys(s.keys, vt.total_keys, st), "fill_keys");
            hip_check(pdr::fill_u32(s.counters, 0u, 4, st), "fill");
        });
    ctx->timed("raster_candidates", st,
               [&] { hip_check(pdr::raster_candidates(vt, idx, T, V, cull, s, tp, st), "raster_candidates"); });
    return s.keys;  // no resolve pass cleared the queue counters: the next raster call fills them
}

// Per-triangle corner sums + ordered vertex reduce for one table of views.
inline void run_raster_backward(pdr_context ctx, const pdr::ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V,
                                hipStream_t st) {
    const pdr::RasterScratch s = raster_scratch(ctx, vt, T, st, false);
    // The adjacency persists in the context: a reallocated part invalidates it,
    // and newly allocated counts start at zero.
    pdr::VertexCsr csr;
    bool f[6];
    csr.counts = ctx->get_t<uint32_t>("csr_counts", static_cast<size_t>(V) + 1, &f[0]);
    csr.offsets = ctx->get_t<uint32_t>("csr_offsets", static_cast<size_t>(V) + 1, &f[1]);
    csr.corners = ctx->get_t<uint32_t>("csr_corners", static_cast<size_t>(T) * 3, &f[2]);
    csr.slot_of = ctx->get_t<uint32_t>("csr_slot_of", static_cast<size_t>(T) * 3, &f[3]);
    csr.saved = ctx->get_t<uint32_t>("csr_saved", static_cast<size_t>(T) * 3, &f[4]);
    csr.state = ctx->get_t<uint32_t>("csr_state", 8, &f[5]);
    csr.totals = ctx->get_t<uint32_t>("csr_totals", pdr::csr_total_words(V));
    if (f[0]) hip_check(pdr::fill_u32(csr.counts, 0u, ctx->scratch["csr_counts"].bytes / 4, st), "fill");
    if (f[1] || f[2] || f[3] || f[4] || f[5]) hip_check(pdr::fill_u32(csr.state, 0u, 8, st), "fill");
    float4* corner = ctx->get_t<float4>("corner_grads", static_cast<size_t>(vt.n) * T * 3);
    pdr::RasterScratch sb = s;
    sb.cid = ctx->get_t<uint32_t>("backward_ids", vt.total_keys);
    sb.cgrad = ctx->get_t<float2>("backward_grads", vt.total_keys);
    bool fresh_flags = false, fresh_won = false;
    sb.view_flags = ctx->get_t<uint32_t>("backward_view_flags", pdr::kMaxViews, &fresh_flags);
    sb.won = reinterpret_cast<uint8_t*>(
        ctx->get_t<uint32_t>("backward_won", (static_cast<size_t>(vt.n) * T + 3) / 4, &fresh_won));
    // zero unless the last backward on this context ran to the end (its vertex reduce cleared them)
    if (fresh_flags || !ctx->bwd_clean)
        hip_check(pdr::fill_u32(sb.view_flags, 0u, ctx->scratch["backward_view_flags"].bytes / 4, st), "fill");
    if (fresh_won || !ctx->bwd_clean)
        hip_check(pdr::fill_u32(reinterpret_cast<uint32_t*>(sb.won), 0u, ctx->scratch["backward_won"].bytes / 4, st),
                  "fill");
    ctx->bwd_clean = false;
    sb.slot_of = csr.slot_of;
    // boxes over 65,536 pixels: the reference's row-major order under PDR_FLAG_DETERMINISTIC
    const bool ordered = (ctx->flags & PDR_FLAG_DETERMINISTIC) != 0;
    ctx->timed("vertex_csr", st, [&] { hip_check(pdr::build_vertex_csr(idx, T, V, csr, st), "build_vertex_csr"); });
    ctx->timed("bwd_prepass", st,
               [&] { hip_check(pdr::backward_prepass(vt, idx, T, V, sb, st), "backward_prepass"); });
    const pdr::TuneParams tp =
        tuned_params(ctx, tune_key("backward", vt, T), backward_candidates(), st, [&](const pdr::TuneParams& c) {
            hip_check(pdr::corner_gradients(vt, idx, T, V, sb, corner, c, ordered, st), "corner_gradients");
            // the vertex reduce clears the queue counters after the tuned run; clear them between candidates
            hip_check(pdr::fill_u32(sb.counters, 0u, 4, st), "fill");
        });
    ctx->timed("corner_grads", st,
               [&] { hip_check(pdr::corner_gradients(vt, idx, T, V, sb, corner, tp, ordered, st), "corner_gradients"); });
    ctx->timed("vertex_reduce", st,
               [&] { hip_check(pdr::vertex_reduce(vt, T, V, csr, corner, sb.won, sb.view_flags, sb.counters, st), "vertex_reduce"); });
    ctx->bwd_clean = true;
    ctx->counters_clean = true;
}

inline void validate_mesh_counts(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T) {
    if (!pos || V == 0) raise(PDR_INVALID_ARGUMENT, "clip_positions must have shape [vertex_count, 4]");
    if (!idx || T == 0) raise(PDR_INVALID_ARGUMENT, "triangle_indices must have shape [triangle_count, 3]");
}

inline uint32_t validate_texture(uint32_t tw, uint32_t th, uint32_t C, uint32_t W, uint32_t H) {
    // src/rasterizer.cpp:260-289
    if (tw == 0 || th == 0 || C == 0)
        raise(PDR_INVALID_ARGUMENT, "Texture width, height, and channel_count must be positive");
    if (tw > static_cast<uint32_t>(std::numeric_limits<int32_t>::max() / 2) ||
        th > static_cast<uint32_t>(std::numeric_limits<int32_t>::max() / 2))
        raise(PDR_OUT_OF_RANGE, "Texture dimensions exceed portable addressing limits");
    if (W == 0 || H == 0) raise(PDR_INVALID_ARGUMENT, "raster_output has inconsistent dimensions");
    const uint64_t P = static_cast<uint64_t>(W) * H;
    if (P > 0xFFFFFFFFull || static_cast<uint64_t>(tw) * th > 0xFFFFFFFFull)
        raise(PDR_OVERFLOW, "Texture operation exceeds backend element limits");
    return static_cast<uint32_t>(P);
}

// A call's temporary device arrays, taken from the context's grow-only
// scratch slots "<tag>#<n>" (the n-th array this call asks for): repeated
// calls with the same sequence of arrays reuse the same memory, with no
// allocation or free on the host path once the slots are large enough.  Calls
// on one context are serialised (its lock, and each call's stream ordered
// after the previous one's), so a call's slots are never in use by another.
struct DeviceBuffers {
    DeviceBuffers(pdr_context c, const char* t) : ctx(c), tag(t) {}
    DeviceBuffers(const DeviceBuffers&) = delete;
    DeviceBuffers& operator=(const DeviceBuffers&) = delete;
    pdr_context ctx;
    std::string tag;
    uint32_t count = 0;
    template <class T>
    T* alloc(size_t n) {
        return ctx->get_t<T>(tag + "#" + std::to_string(count++), std::max<size_t>(n, 1));
    }
};

// Device copy of a host array (device arrays are used in place).
template <class T>
const T* device_view(DeviceBuffers& bufs, const T* p, size_t count, hipStream_t st) {
    if (!p) return nullptr;
    if (is_device_pointer(p)) return p;
    T* d = bufs.alloc<T>(count);
    hip_check(hipMemcpyAsync(d, p, count * sizeof(T), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    return d;
}

inline void copy_any(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (!dst || !bytes || dst == src) return;
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st), "hipMemcpyAsync");
}

}  // namespace pdrh
