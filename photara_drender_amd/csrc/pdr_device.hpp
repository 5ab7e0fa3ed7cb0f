// Device-side helpers shared by every libpdr kernel (gfx950, wave64).
//
// Arithmetic helpers reproduce the reference shaders' float semantics exactly:
// the library is compiled with -ffp-contract=off and correctly rounded
// division/sqrt, so each expression rounds as written.  Out-of-range
// float->int conversions saturate (NaN -> 0) and clamps are
// min(max(x, lo), hi) with std::min/std::max semantics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdr {

constexpr uint32_t kTile = 16;         // src/rasterizer.cpp:19 RASTER_TILE_SIZE
constexpr uint32_t kTilePixels = kTile * kTile;
constexpr uint32_t kMaxViews = 32;     // views per batched launch (kernel-argument table)
constexpr uint32_t kRasterTile = 32;   // tiled forward: pixels per raster tile side (LDS depth keys)
constexpr uint32_t kGradTile = 64;     // pixels per side of the interpolation-backward tiles

__host__ __device__ __forceinline__ int32_t f2i_sat(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return static_cast<int32_t>(f);
}

__host__ __device__ __forceinline__ int32_t d2i_sat(double f) {
    if (f != f) return 0;
    if (f >= 2147483648.0) return 2147483647;
    if (f <= -2147483648.0) return (-2147483647 - 1);
    return static_cast<int32_t>(f);
}

// (uint)round(id) with HLSL round = half-to-even; negative/NaN ids decode as 0.
__host__ __device__ __forceinline__ uint32_t decode_id(float f) {
    const float r = __builtin_rintf(f);
    if (!(r > 0.0f)) return 0u;
    if (r >= 4294967296.0f) return 0xFFFFFFFFu;
    return static_cast<uint32_t>(r);
}

__host__ __device__ __forceinline__ float cross2(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

__host__ __device__ __forceinline__ float maxf_std(float a, float b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ float minf_std(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ float clampf(float x, float lo, float hi) { return minf_std(maxf_std(x, lo), hi); }

__host__ __device__ __forceinline__ int32_t add1(int32_t c) { return static_cast<int32_t>(static_cast<uint32_t>(c) + 1u); }

// texture_forward.hlsl:16-35 address_coordinate
__host__ __device__ __forceinline__ int32_t address_coordinate(int32_t c, uint32_t size, uint32_t mode) {
    const int32_t n = static_cast<int32_t>(size);
    if (mode == 0) return c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
    if (mode == 1) {
        const int32_t r = c % n;
        return r < 0 ? r + n : r;
    }
    const int32_t period = n * 2;
    int32_t m = c % period;
    if (m < 0) m += period;
    return m < n ? m : period - 1 - m;
}

// XCD-aware block order: consecutive hardware block ids are dealt round-robin to
// the 8 XCDs; map them so each XCD walks a contiguous range of work items
// (neighbouring tiles share triangles and texels in that XCD's L2).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    constexpr uint32_t kXcd = 8;
    if (n < kXcd * 4) return b;
    const uint32_t per = n / kXcd, rem = n % kXcd;
    const uint32_t xcd = b % kXcd, local = b / kXcd;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Sum over the 64 lanes of a wave with the groupshared tree of the atlas shaders
// (stride 32, 16, ..., 1: lane i adds lane i+stride).  Lane 0 holds the result,
// bit-identical to the shader's group_loss[0].
__device__ __forceinline__ float wave_tree_sum(float v) {
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_down(v, s, 64);
    return v;
}

// View (index into table.v, ordered by pixel_base) of the first pixel of the
// calling wave, looked up once per wave with scalar loads; `split` is set when
// the wave's 64 pixels may reach into a later view (the callers then look the
// view up per lane).
template <class Table>
__device__ __forceinline__ uint32_t wave_view(const Table& t, uint32_t gp, bool& split) {
    const uint32_t first = __builtin_amdgcn_readfirstlane(gp);
    uint32_t k = 0;
    while (k + 1 < t.n && first >= t.v[k + 1].pixel_base) ++k;
    split = k + 1 < t.n && first + 63u >= t.v[k + 1].pixel_base;
    return k;
}

}  // namespace pdr
