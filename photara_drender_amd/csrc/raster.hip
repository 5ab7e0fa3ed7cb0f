// Tile binning, rasterize forward and rasterize backward for gfx950.
//
// Forward (shaders/rasterize_forward.hlsl:45-175 over the bins of
// src/rasterizer.cpp:61-138):
//   1. k_bin_count   one thread per (view, triangle): the reference's pixel
//                    bounds, atomically counted per 16x16 tile;
//   2. scan          inclusive prefix sum of the counts (hipcub);
//   3. k_bin_fill    the same bounds, each triangle appended to its tiles;
//   4. k_raster_fwd  one 256-thread workgroup per tile (one pixel per lane);
//                    the tile's triangles are set up once per chunk into LDS
//                    (screen positions, orientation sign, clip positions) and
//                    every lane runs the reference's coverage / depth test
//                    against each of them.
// Bins are filled in arbitrary order: the winner is the lexicographic
// (depth, triangle index) minimum (rasterize_forward.hlsl:136-141), which does
// not depend on visit order.  Bins use the reference bounds exactly, so every
// pixel sees the same candidate set as the reference.
//
// Backward (shaders/rasterize_backward.hlsl:36-158 + reduce_vertex_gradients):
// one workgroup per tile; each pixel evaluates its own triangle's corner terms
// (the reference's per-pixel expressions, gated by the same bounding box), the
// terms are summed per triangle in an LDS hash table and flushed with one
// global float atomic per (triangle corner, component).
#include <hipcub/hipcub.hpp>

#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr uint32_t kBlock = 256;

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

uint32_t grid_for(size_t n, uint32_t block = kBlock) {
    const size_t g = (n + block - 1) / block;
    return static_cast<uint32_t>(g < 65536 ? (g ? g : 1) : 65536);
}

struct Rect {
    int32_t x0, x1, y0, y1;
};

// src/rasterizer.cpp:73-116: NDC bounds of a valid triangle -> inclusive pixel rect.
__device__ __forceinline__ bool bin_rect(const ViewDesc& vd, const uint32_t* idx, uint32_t t, uint32_t V, Rect* r) {
    float min_x = __builtin_inff(), min_y = __builtin_inff();
    float max_x = -__builtin_inff(), max_y = -__builtin_inff();
    const uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    if (i0 >= V || i1 >= V || i2 >= V) return false;
    const uint32_t vi[3] = {i0, i1, i2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float4 p = reinterpret_cast<const float4*>(vd.pos)[vi[c]];
        if (!(p.w > 1e-8f) || !__builtin_isfinite(p.w)) return false;
        const float x = p.x / p.w, y = p.y / p.w;
        if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) return false;
        min_x = minf_std(min_x, x);
        min_y = minf_std(min_y, y);
        max_x = maxf_std(max_x, x);
        max_y = maxf_std(max_y, y);
    }
    if (max_x < -1.0f || min_x > 1.0f || max_y < -1.0f || min_y > 1.0f) return false;
    const int64_t wm = static_cast<int64_t>(vd.width) - 1, hm = static_cast<int64_t>(vd.height) - 1;
    auto cl = [](int64_t v, int64_t hi) { return static_cast<int32_t>(v < 0 ? 0 : (v > hi ? hi : v)); };
    r->x0 = cl(f2i_sat(floorf((min_x * 0.5f + 0.5f) * vd.vw + vd.vx)), wm);
    r->x1 = cl(static_cast<int64_t>(f2i_sat(ceilf((max_x * 0.5f + 0.5f) * vd.vw + vd.vx))) - 1, wm);
    r->y0 = cl(f2i_sat(floorf((min_y * 0.5f + 0.5f) * vd.vh + vd.vy)), hm);
    r->y1 = cl(static_cast<int64_t>(f2i_sat(ceilf((max_y * 0.5f + 0.5f) * vd.vh + vd.vy))) - 1, hm);
    return true;
}

__global__ __launch_bounds__(kBlock) void k_bin_count(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                      uint32_t V, uint32_t* __restrict__ counts) {
    const ViewDesc& vd = vt.v[blockIdx.y];
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < T; t += gridDim.x * kBlock) {
        Rect r;
        if (!bin_rect(vd, idx, t, V, &r)) continue;
        for (uint32_t ty = r.y0 / kTile; ty <= static_cast<uint32_t>(r.y1) / kTile; ++ty)
            for (uint32_t tx = r.x0 / kTile; tx <= static_cast<uint32_t>(r.x1) / kTile; ++tx)
                atomicAdd(&counts[vd.tile_base + ty * vd.tiles_x + tx], 1u);
    }
}

__global__ __launch_bounds__(kBlock) void k_bin_fill(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                     uint32_t V, const uint32_t* __restrict__ offsets,
                                                     uint32_t* __restrict__ cursors, uint32_t* __restrict__ list) {
    const ViewDesc& vd = vt.v[blockIdx.y];
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < T; t += gridDim.x * kBlock) {
        Rect r;
        if (!bin_rect(vd, idx, t, V, &r)) continue;
        for (uint32_t ty = r.y0 / kTile; ty <= static_cast<uint32_t>(r.y1) / kTile; ++ty)
            for (uint32_t tx = r.x0 / kTile; tx <= static_cast<uint32_t>(r.x1) / kTile; ++tx) {
                const uint32_t tile = vd.tile_base + ty * vd.tiles_x + tx;
                list[offsets[tile] + atomicAdd(&cursors[tile], 1u)] = t;
            }
    }
}

__device__ __forceinline__ uint32_t find_view(const ViewTable& vt, uint32_t g) {
    uint32_t k = 0;
    while (k + 1 < vt.n && g >= vt.v[k + 1].tile_base) ++k;
    return k;
}

// Per-triangle record staged in LDS: screen positions + orientation, then the
// clip positions needed once a lane passes the coverage test.
struct __align__(16) TriRec {
    float4 s01;   // s0.x s0.y s1.x s1.y
    float4 s2sg;  // s2.x s2.y sign(screen_area) triangle
    float4 p0, p1, p2;
};

__global__ __launch_bounds__(kBlock) void k_raster_fwd(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                       uint32_t V, uint32_t cull, int want_deriv,
                                                       const uint32_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ list) {
    __shared__ TriRec tris[kBlock];
    __shared__ uint32_t n_valid;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const ViewDesc& vd = vt.v[find_view(vt, g)];
    const uint32_t local = g - vd.tile_base;
    const uint32_t px = (local % vd.tiles_x) * kTile + (threadIdx.x & 15u);
    const uint32_t py = (local / vd.tiles_x) * kTile + (threadIdx.x >> 4);
    const float sx = (static_cast<float>(px) - vd.vx + 0.5f) * 2.0f / vd.vw - 1.0f;
    const float sy = (static_cast<float>(py) - vd.vy + 0.5f) * 2.0f / vd.vh - 1.0f;
    float best_depth = 3.402823466e+38f;
    uint32_t best_tri = 0;
    float best_b0 = 0.0f, best_b1 = 0.0f;
    const uint32_t begin = offsets[g], end = offsets[g + 1];
    for (uint32_t base = begin; base < end; base += kBlock) {
        if (threadIdx.x == 0) n_valid = 0;
        __syncthreads();
        const uint32_t e = base + threadIdx.x;
        if (e < end) {
            const uint32_t t = list[e];
            const uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
            if (i0 < V && i1 < V && i2 < V) {
                const float4* P = reinterpret_cast<const float4*>(vd.pos);
                const float4 p0 = P[i0], p1 = P[i1], p2 = P[i2];
                if (!(p0.w <= 1e-8f || p1.w <= 1e-8f || p2.w <= 1e-8f)) {
                    const float s0x = p0.x / p0.w, s0y = p0.y / p0.w;
                    const float s1x = p1.x / p1.w, s1y = p1.y / p1.w;
                    const float s2x = p2.x / p2.w, s2y = p2.y / p2.w;
                    const float sa = cross2(s1x - s0x, s1y - s0y, s2x - s0x, s2y - s0y);
                    const bool culled = (cull == 1 && sa <= 0.0f) || (cull == 2 && sa >= 0.0f);
                    if (!(fabsf(sa) <= 1e-12f) && !culled) {
                        const uint32_t slot = atomicAdd(&n_valid, 1u);
                        TriRec& r = tris[slot];
                        r.s01 = make_float4(s0x, s0y, s1x, s1y);
                        r.s2sg = make_float4(s2x, s2y, sa > 0.0f ? 1.0f : -1.0f, __uint_as_float(t));
                        r.p0 = p0;
                        r.p1 = p1;
                        r.p2 = p2;
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t n = n_valid;
        for (uint32_t j = 0; j < n; ++j) {
            const float4 s01 = tris[j].s01;
            const float4 s2sg = tris[j].s2sg;
            const float e0 = cross2(s01.z - sx, s01.w - sy, s2sg.x - sx, s2sg.y - sy);
            const float e1 = cross2(s2sg.x - sx, s2sg.y - sy, s01.x - sx, s01.y - sy);
            const float e2 = cross2(s01.x - sx, s01.y - sy, s01.z - sx, s01.w - sy);
            const float sgn = s2sg.z;
            if (e0 * sgn < -1e-7f || e1 * sgn < -1e-7f || e2 * sgn < -1e-7f) continue;
            const float4 p0 = tris[j].p0, p1 = tris[j].p1, p2 = tris[j].p2;
            const float q0x = p0.x - sx * p0.w, q0y = p0.y - sy * p0.w;
            const float q1x = p1.x - sx * p1.w, q1y = p1.y - sy * p1.w;
            const float q2x = p2.x - sx * p2.w, q2y = p2.y - sy * p2.w;
            const float a0 = cross2(q1x, q1y, q2x, q2y);
            const float a1 = cross2(q2x, q2y, q0x, q0y);
            const float a2 = cross2(q0x, q0y, q1x, q1y);
            const float area = a0 + a1 + a2;
            if (fabsf(area) <= 1e-12f) continue;
            const float inv = 1.0f / area;
            const float iw = p0.w * a0 + p1.w * a1 + p2.w * a2;
            if (fabsf(iw) <= 1e-12f) continue;
            const float iz = p0.z * a0 + p1.z * a1 + p2.z * a2;
            const float depth = clampf(iz / iw, -1.0f, 1.0f);
            const uint32_t t1 = __float_as_uint(s2sg.w) + 1u;
            if (depth > best_depth || (depth == best_depth && best_tri != 0 && t1 >= best_tri)) continue;
            best_depth = depth;
            best_tri = t1;
            best_b0 = clampf(a0 * inv, 0.0f, 1.0f);
            best_b1 = clampf(a1 * inv, 0.0f, 1.0f);
        }
        __syncthreads();
    }
    if (px >= vd.width || py >= vd.height) return;
    const size_t pixel = static_cast<size_t>(py) * vd.width + px;
    float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (best_tri != 0) {
        r = make_float4(best_b0, best_b1, best_depth, static_cast<float>(best_tri));
        if (want_deriv) {
            // rasterize_forward.hlsl:145-162 for the winning triangle
            const uint32_t t = best_tri - 1;
            const float4* P = reinterpret_cast<const float4*>(vd.pos);
            const float4 p0 = P[idx[t * 3 + 0]], p1 = P[idx[t * 3 + 1]], p2 = P[idx[t * 3 + 2]];
            const float q0x = p0.x - sx * p0.w, q0y = p0.y - sy * p0.w;
            const float q1x = p1.x - sx * p1.w, q1y = p1.y - sy * p1.w;
            const float q2x = p2.x - sx * p2.w, q2y = p2.y - sy * p2.w;
            const float a0 = cross2(q1x, q1y, q2x, q2y);
            const float a1 = cross2(q2x, q2y, q0x, q0y);
            const float a2 = cross2(q0x, q0y, q1x, q1y);
            const float inv = 1.0f / (a0 + a1 + a2);
            const float da0_dx = -p1.w * p2.y + p2.w * p1.y;
            const float da0_dy = -p1.x * p2.w + p1.w * p2.x;
            const float da1_dx = -p2.w * p0.y + p2.y * p0.w;
            const float da1_dy = -p2.x * p0.w + p2.w * p0.x;
            const float da2_dx = -p0.w * p1.y + p0.y * p1.w;
            const float da2_dy = -p0.x * p1.w + p0.w * p1.x;
            const float dA_dx = da0_dx + da1_dx + da2_dx;
            const float dA_dy = da0_dy + da1_dy + da2_dy;
            d.x = (da0_dx - best_b0 * dA_dx) * inv * 2.0f / vd.vw;
            d.z = (da1_dx - best_b1 * dA_dx) * inv * 2.0f / vd.vw;
            d.y = (da0_dy - best_b0 * dA_dy) * inv * 2.0f / vd.vh;
            d.w = (da1_dy - best_b1 * dA_dy) * inv * 2.0f / vd.vh;
        }
    }
    reinterpret_cast<float4*>(vd.raster)[pixel] = r;
    if (vd.deriv) reinterpret_cast<float4*>(vd.deriv)[pixel] = d;
}

constexpr uint32_t kHashSlots = 512;

__device__ __forceinline__ uint32_t hash_slot(uint32_t* keys, uint32_t key) {
    uint32_t h = (key * 2654435761u) >> (32 - 9);
    while (true) {
        const uint32_t old = atomicCAS(&keys[h], 0u, key);
        if (old == 0u || old == key) return h;
        h = (h + 1) & (kHashSlots - 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_raster_bwd(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                       uint32_t V) {
    __shared__ uint32_t keys[kHashSlots];
    __shared__ float acc[kHashSlots * 9];
    for (uint32_t s = threadIdx.x; s < kHashSlots; s += kBlock) keys[s] = 0u;
    for (uint32_t s = threadIdx.x; s < kHashSlots * 9; s += kBlock) acc[s] = 0.0f;
    __syncthreads();
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const ViewDesc& vd = vt.v[find_view(vt, g)];
    const uint32_t local = g - vd.tile_base;
    const uint32_t px = (local % vd.tiles_x) * kTile + (threadIdx.x & 15u);
    const uint32_t py = (local / vd.tiles_x) * kTile + (threadIdx.x >> 4);
    if (px < vd.width && py < vd.height) {
        const size_t pixel = static_cast<size_t>(py) * vd.width + px;
        const uint32_t id = decode_id(vd.raster[pixel * 4 + 3]);
        const float2 lg = reinterpret_cast<const float2*>(vd.grad_raster)[pixel * 2];
        if (id != 0 && id <= T && !(lg.x == 0.0f && lg.y == 0.0f)) {
            const uint32_t t = id - 1;
            const uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
            const float4* P = reinterpret_cast<const float4*>(vd.pos);
            if (i0 < V && i1 < V && i2 < V) {
                const float4 p0 = P[i0], p1 = P[i1], p2 = P[i2];
                if (!(p0.w <= 1e-8f || p1.w <= 1e-8f || p2.w <= 1e-8f)) {
                    // rasterize_backward.hlsl:71-87: the reference visits only the
                    // triangle's screen bounding box.
                    const float s0x = p0.x / p0.w, s0y = p0.y / p0.w;
                    const float s1x = p1.x / p1.w, s1y = p1.y / p1.w;
                    const float s2x = p2.x / p2.w, s2y = p2.y / p2.w;
                    const float mnx = minf_std(s0x, minf_std(s1x, s2x)), mny = minf_std(s0y, minf_std(s1y, s2y));
                    const float mxx = maxf_std(s0x, maxf_std(s1x, s2x)), mxy = maxf_std(s0y, maxf_std(s1y, s2y));
                    auto cl = [](int32_t v, int32_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
                    const int32_t W = static_cast<int32_t>(vd.width), H = static_cast<int32_t>(vd.height);
                    const int32_t x0 = cl(f2i_sat(floorf((mnx * 0.5f + 0.5f) * vd.vw + vd.vx)), W);
                    const int32_t x1 = cl(f2i_sat(ceilf((mxx * 0.5f + 0.5f) * vd.vw + vd.vx)), W);
                    const int32_t y0 = cl(f2i_sat(floorf((mny * 0.5f + 0.5f) * vd.vh + vd.vy)), H);
                    const int32_t y1 = cl(f2i_sat(ceilf((mxy * 0.5f + 0.5f) * vd.vh + vd.vy)), H);
                    const int32_t ix = static_cast<int32_t>(px), iy = static_cast<int32_t>(py);
                    if (ix >= x0 && ix < x1 && iy >= y0 && iy < y1) {
                        const float sx = (static_cast<float>(px) - vd.vx + 0.5f) * 2.0f / vd.vw - 1.0f;
                        const float sy = (static_cast<float>(py) - vd.vy + 0.5f) * 2.0f / vd.vh - 1.0f;
                        const float q0x = p0.x - sx * p0.w, q0y = p0.y - sy * p0.w;
                        const float q1x = p1.x - sx * p1.w, q1y = p1.y - sy * p1.w;
                        const float q2x = p2.x - sx * p2.w, q2y = p2.y - sy * p2.w;
                        const float area =
                            cross2(q1x, q1y, q2x, q2y) + cross2(q2x, q2y, q0x, q0y) + cross2(q0x, q0y, q1x, q1y);
                        if (!(fabsf(area) <= 1e-12f)) {
                            const float eps = area >= 0.0f ? 1e-6f : -1e-6f;
                            const float inv = 1.0f / (area + eps);
                            const float a0 = cross2(q1x, q1y, q2x, q2y);
                            const float a1 = cross2(q2x, q2y, q0x, q0y);
                            const float b0 = a0 * inv, b1 = a1 * inv;
                            const float gb0 = lg.x * inv, gb1 = lg.y * inv;
                            const float gc = gb0 * b0 + gb1 * b1;
                            float c[9];
                            c[0] = gc * (q2y - q1y) - gb1 * q2y;
                            c[1] = gc * (q1x - q2x) + gb1 * q2x;
                            c[3] = gc * (q0y - q2y) + gb0 * q2y;
                            c[4] = gc * (q2x - q0x) - gb0 * q2x;
                            c[6] = gc * (q1y - q0y) - gb0 * q1y + gb1 * q0y;
                            c[7] = gc * (q0x - q1x) + gb0 * q1x - gb1 * q0x;
                            c[2] = -sx * c[0] - sy * c[1];
                            c[5] = -sx * c[3] - sy * c[4];
                            c[8] = -sx * c[6] - sy * c[7];
                            const uint32_t slot = hash_slot(keys, id);
#pragma unroll
                            for (int k = 0; k < 9; ++k) atomicAdd(&acc[slot * 9 + k], c[k]);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < kHashSlots; s += kBlock) {
        const uint32_t key = keys[s];
        if (!key) continue;
        const uint32_t t = key - 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float* gp = vd.grad_pos + static_cast<size_t>(idx[t * 3 + c]) * 4;
            unsafeAtomicAdd(gp + 0, acc[s * 9 + c * 3 + 0]);
            unsafeAtomicAdd(gp + 1, acc[s * 9 + c * 3 + 1]);
            unsafeAtomicAdd(gp + 3, acc[s * 9 + c * 3 + 2]);
        }
    }
}

__global__ void k_check_indices(const uint32_t* idx, size_t n, uint32_t V, uint32_t* flags) {
    bool bad = false;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        bad |= idx[i] >= V;
    if (__any(bad) && (threadIdx.x & 63u) == 0) atomicOr(flags, 1u);
}

}  // namespace

hipError_t fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_u32<<<grid_for(n), kBlock, 0, st>>>(p, value, n);
    return hipGetLastError();
}

hipError_t fill_f32(float* p, float value, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_f32<<<grid_for(n), kBlock, 0, st>>>(p, value, n);
    return hipGetLastError();
}

hipError_t check_indices(const uint32_t* idx, size_t n, uint32_t V, uint32_t* flags, hipStream_t st) {
    k_check_indices<<<grid_for(n), kBlock, 0, st>>>(idx, n, V, flags);
    return hipGetLastError();
}

size_t bin_scan_temp_bytes(uint32_t total_tiles) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, bytes, static_cast<const uint32_t*>(nullptr),
                                     static_cast<uint32_t*>(nullptr), static_cast<int>(total_tiles));
    return bytes;
}

hipError_t bin_count(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, const BinScratch& s,
                     hipStream_t st) {
    hipError_t e = fill_u32(s.counts, 0u, vt.total_tiles, st);
    if (e != hipSuccess) return e;
    e = fill_u32(s.offsets, 0u, 1, st);
    if (e != hipSuccess) return e;
    const dim3 grid((T + kBlock - 1) / kBlock < 4096 ? (T + kBlock - 1) / kBlock : 4096, vt.n);
    k_bin_count<<<grid, kBlock, 0, st>>>(vt, idx, T, V, s.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t bytes = s.scan_temp_bytes;
    return hipcub::DeviceScan::InclusiveSum(s.scan_temp, bytes, s.counts, s.offsets + 1,
                                            static_cast<int>(vt.total_tiles), st);
}

hipError_t bin_fill(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, const BinScratch& s,
                    hipStream_t st) {
    hipError_t e = fill_u32(s.counts, 0u, vt.total_tiles, st);
    if (e != hipSuccess) return e;
    const dim3 grid((T + kBlock - 1) / kBlock < 4096 ? (T + kBlock - 1) / kBlock : 4096, vt.n);
    k_bin_fill<<<grid, kBlock, 0, st>>>(vt, idx, T, V, s.offsets, s.counts, s.list);
    return hipGetLastError();
}

hipError_t raster_forward(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, uint32_t cull,
                          int want_deriv, const BinScratch& s, hipStream_t st) {
    k_raster_fwd<<<vt.total_tiles, kBlock, 0, st>>>(vt, idx, T, V, cull, want_deriv, s.offsets, s.list);
    return hipGetLastError();
}

hipError_t raster_backward(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, hipStream_t st) {
    k_raster_bwd<<<vt.total_tiles, kBlock, 0, st>>>(vt, idx, T, V);
    return hipGetLastError();
}

}  // namespace pdr
