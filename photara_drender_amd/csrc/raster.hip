// Rasterize forward and backward for gfx950, triangle-parallel.
//
// Forward (shaders/rasterize_forward.hlsl:45-175; candidates from the tile bins
// of src/rasterizer.cpp:61-138):
//   k_sample_tables   pixel-centre NDC of every column / row of every view
//                     (rasterize_forward.hlsl:55-57, computed once per call);
//   k_raster_small    one lane per (view, triangle): the triangle's candidate
//                     pixels are those of its 16x16 bins (the reference's pixel
//                     bounds, tile-expanded) intersected with a conservative
//                     bound of every pixel that can pass the reference's
//                     epsilon-inclusive edge test.  Each candidate runs the
//                     reference's coverage / depth expressions and, if
//                     covered, does a 64-bit atomicMin of (depth, id) into the
//                     pixel's key.  Triangles with large candidate rectangles
//                     are queued;
//   k_raster_large    persistent workgroups walk the queued rectangles, one
//                     pixel per lane;
//   k_raster_resolve  one lane per pixel: decodes the winning triangle,
//                     recomputes its barycentrics / depth / derivatives with
//                     the reference's expressions, writes raster and
//                     derivatives, and resets the key.
// The reference keeps the lexicographic minimum of (depth, triangle index)
// (rasterize_forward.hlsl:136-141), which is independent of visit order, so
// atomicMin over an order-preserving key selects the same winner; -0.0 depths
// are canonicalised to +0.0 in the key (they compare equal in the reference).
//
// Backward (shaders/rasterize_backward.hlsl:36-158 + reduce_vertex_gradients.hlsl):
//   k_corner_small    one lane per (view, triangle) loops over the reference's
//                     screen bounding box in the reference's row-major order
//                     and sums the corner gradients in registers (bit-exact
//                     per-triangle sums); larger boxes are queued;
//   k_corner_large    queued boxes, one workgroup each, fixed-order reduction;
//   k_vertex_reduce   per vertex, corner gradients summed in ascending corner
//                     order over a vertex -> corner adjacency built on the GPU
//                     (counting + stable radix sort), as the reference's CSR.
// No float atomics: the backward is deterministic.
#include <hipcub/hipcub.hpp>

#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kInlinePixels = 4;    // candidate rectangle tested by its own lane
constexpr uint32_t kWavePixels = 16384;  // candidate rectangle tested by one wave
constexpr uint32_t kInlineBox = 8;       // backward box summed by its own lane
constexpr uint32_t kWaveBox = 65536;     // backward box summed by one wave (in pixel order)
constexpr uint32_t kLargeBlocks = 1024;  // persistent workgroups for queued work

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

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

uint32_t grid_for(size_t n, uint32_t block = kBlock) {
    const size_t g = (n + block - 1) / block;
    return static_cast<uint32_t>(g < 65536 ? (g ? g : 1) : 65536);
}

__device__ __forceinline__ uint32_t find_view(const ViewTable& vt, uint32_t g) {
    uint32_t k = 0;
    while (k + 1 < vt.n && g >= vt.v[k + 1].pixel_base) ++k;
    return k;
}

// rasterize_forward.hlsl:55-57 for every column and row of every view.
__global__ void k_sample_tables(ViewTable vt, float* __restrict__ sx, float* __restrict__ sy) {
    const ViewDesc& vd = vt.v[blockIdx.y];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < vd.width || i < vd.height; i += gridDim.x * kBlock) {
        if (i < vd.width) sx[vd.col_base + i] = (static_cast<float>(i) - vd.vx + 0.5f) * 2.0f / vd.vw - 1.0f;
        if (i < vd.height) sy[vd.row_base + i] = (static_cast<float>(i) - vd.vy + 0.5f) * 2.0f / vd.vh - 1.0f;
    }
}

struct Tri {
    float4 p0, p1, p2;
    float s0x, s0y, s1x, s1y, s2x, s2y, sgn;
    int32_t x0, x1, y0, y1;  // inclusive candidate rectangle
};

// Pixel range whose centres can lie in { x : sgn * E_i(x) >= -eps for all i },
// E_i the exact edge functions of the screen triangle, eps the reference's
// 1e-7 plus a bound on the float evaluation error of the edge test.  Returns
// false when the region is not usefully bounded (near-degenerate triangles);
// the caller then keeps the whole bin rectangle.
__device__ bool halo_rect(const Tri& r, const ViewDesc& vd, int32_t cx0, int32_t cx1, int32_t cy0, int32_t cy1,
                          int32_t* hx0, int32_t* hx1, int32_t* hy0, int32_t* hy1) {
    const double S[3][2] = {{r.s0x, r.s0y}, {r.s1x, r.s1y}, {r.s2x, r.s2y}};
    const double vx = vd.vx, vy = vd.vy, vw = vd.vw, vh = vd.vh;
    auto ndc_x = [&](double px) { return ((px - vx) + 0.5) * 2.0 / vw - 1.0; };
    auto ndc_y = [&](double py) { return ((py - vy) + 0.5) * 2.0 / vh - 1.0; };
    const double nx0 = ndc_x(cx0), nx1 = ndc_x(cx1), ny0 = ndc_y(cy0), ny1 = ndc_y(cy1);
    double M = 0.0, cmax = 0.0;
    for (int j = 0; j < 3; ++j) {
        M = fmax(M, fmax(fmax(fabs(S[j][0] - nx0), fabs(S[j][0] - nx1)), fmax(fabs(S[j][1] - ny0), fabs(S[j][1] - ny1))));
        cmax = fmax(cmax, fmax(fabs(S[j][0]), fabs(S[j][1])));
    }
    // float edge test error <= ~4u (|ax||by| + |ay||bx|) <= 8u M^2; take 32u M^2
    const double eps = 1.0000001e-7 + 32.0 * 5.9604644775390625e-8 * M * M + 1e-15 * cmax * cmax;
    double a[3], b[3], c[3];
    const int jk[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    for (int i = 0; i < 3; ++i) {
        const double* sj = S[jk[i][0]];
        const double* sk = S[jk[i][1]];
        // E_i(x, y) = (sj.x sk.y - sj.y sk.x) + x (sj.y - sk.y) + y (sk.x - sj.x)
        a[i] = r.sgn * (sj[1] - sk[1]);
        b[i] = r.sgn * (sk[0] - sj[0]);
        c[i] = -r.sgn * (sj[0] * sk[1] - sj[1] * sk[0]) - eps;  // region: a x + b y >= c
    }
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        const double det = a[i] * b[j] - a[j] * b[i];
        const double scale = (fabs(a[i]) + fabs(b[i])) * (fabs(a[j]) + fabs(b[j]));
        if (!(fabs(det) > 1e-9 * scale)) return false;
        const double x = (c[i] * b[j] - c[j] * b[i]) / det;
        const double y = (a[i] * c[j] - a[j] * c[i]) / det;
        if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) return false;
        xmin = fmin(xmin, x);
        xmax = fmax(xmax, x);
        ymin = fmin(ymin, y);
        ymax = fmax(ymax, y);
    }
    // pixel-centre rounding of the float sample position + double rounding above
    const double pad = 1e-5 * (1.0 + fmax(fabs(xmin), fmax(fabs(xmax), fmax(fabs(ymin), fabs(ymax)))));
    auto to_px = [](double ndc, double v0, double vsize) { return (ndc + 1.0) * vsize * 0.5 + v0 - 0.5; };
    const double fx0 = ceil(to_px(xmin - pad, vx, vw) - 1e-6), fx1 = floor(to_px(xmax + pad, vx, vw) + 1e-6);
    const double fy0 = ceil(to_px(ymin - pad, vy, vh) - 1e-6), fy1 = floor(to_px(ymax + pad, vy, vh) + 1e-6);
    auto cl = [](double v) { return static_cast<int32_t>(v < -2.0e9 ? -2.0e9 : (v > 2.0e9 ? 2.0e9 : v)); };
    *hx0 = cl(fx0);
    *hx1 = cl(fx1);
    *hy0 = cl(fy0);
    *hy1 = cl(fy1);
    return true;
}

// Binning validity (src/rasterizer.cpp:73-100), shader validity
// (rasterize_forward.hlsl:75-101) and the candidate rectangle.
__device__ bool setup_forward(const ViewDesc& vd, const uint32_t* __restrict__ idx, uint32_t t, uint32_t V,
                              uint32_t cull, Tri& r) {
    const uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    if (i0 >= V || i1 >= V || i2 >= V) return false;
    const float4* P = reinterpret_cast<const float4*>(vd.pos);
    r.p0 = P[i0];
    r.p1 = P[i1];
    r.p2 = P[i2];
    const float4* ps[3] = {&r.p0, &r.p1, &r.p2};
    float s[3][2];
    float min_x = __builtin_inff(), min_y = __builtin_inff();
    float max_x = -__builtin_inff(), max_y = -__builtin_inff();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float w = ps[c]->w;
        if (!(w > 1e-8f) || !__builtin_isfinite(w)) return false;
        s[c][0] = ps[c]->x / w;
        s[c][1] = ps[c]->y / w;
        if (!__builtin_isfinite(s[c][0]) || !__builtin_isfinite(s[c][1])) return false;
        min_x = minf_std(min_x, s[c][0]);
        min_y = minf_std(min_y, s[c][1]);
        max_x = maxf_std(max_x, s[c][0]);
        max_y = maxf_std(max_y, s[c][1]);
    }
    if (max_x < -1.0f || min_x > 1.0f || max_y < -1.0f || min_y > 1.0f) return false;
    r.s0x = s[0][0];
    r.s0y = s[0][1];
    r.s1x = s[1][0];
    r.s1y = s[1][1];
    r.s2x = s[2][0];
    r.s2y = s[2][1];
    const float sa = cross2(r.s1x - r.s0x, r.s1y - r.s0y, r.s2x - r.s0x, r.s2y - r.s0y);
    if (fabsf(sa) <= 1e-12f) return false;
    if ((cull == 1 && sa <= 0.0f) || (cull == 2 && sa >= 0.0f)) return false;
    r.sgn = sa > 0.0f ? 1.0f : -1.0f;
    // src/rasterizer.cpp:101-116 bounds, expanded to whole 16x16 tiles
    const int64_t wm = static_cast<int64_t>(vd.width) - 1, hm = static_cast<int64_t>(vd.height) - 1;
    auto cl = [](int64_t v, int64_t hi) { return static_cast<int32_t>(v < 0 ? 0 : (v > hi ? hi : v)); };
    const int32_t bx0 = cl(f2i_sat(floorf((min_x * 0.5f + 0.5f) * vd.vw + vd.vx)), wm);
    const int32_t bx1 = cl(static_cast<int64_t>(f2i_sat(ceilf((max_x * 0.5f + 0.5f) * vd.vw + vd.vx))) - 1, wm);
    const int32_t by0 = cl(f2i_sat(floorf((min_y * 0.5f + 0.5f) * vd.vh + vd.vy)), hm);
    const int32_t by1 = cl(static_cast<int64_t>(f2i_sat(ceilf((max_y * 0.5f + 0.5f) * vd.vh + vd.vy))) - 1, hm);
    int32_t cx0 = (bx0 / static_cast<int32_t>(kTile)) * kTile;
    int32_t cx1 = static_cast<int32_t>(min<int64_t>((bx1 / static_cast<int32_t>(kTile)) * kTile + kTile - 1, wm));
    int32_t cy0 = (by0 / static_cast<int32_t>(kTile)) * kTile;
    int32_t cy1 = static_cast<int32_t>(min<int64_t>((by1 / static_cast<int32_t>(kTile)) * kTile + kTile - 1, hm));
    int32_t hx0, hx1, hy0, hy1;
    if (halo_rect(r, vd, cx0, cx1, cy0, cy1, &hx0, &hx1, &hy0, &hy1)) {
        cx0 = max(cx0, hx0);
        cx1 = min(cx1, hx1);
        cy0 = max(cy0, hy0);
        cy1 = min(cy1, hy1);
    }
    r.x0 = cx0;
    r.x1 = cx1;
    r.y0 = cy0;
    r.y1 = cy1;
    return cx0 <= cx1 && cy0 <= cy1;
}

__device__ __forceinline__ uint32_t order_key(float depth) {
    const uint32_t u = __float_as_uint(depth + 0.0f);  // -0 -> +0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// rasterize_forward.hlsl:103-135 for one candidate pixel; returns the key or ~0.
__device__ __forceinline__ unsigned long long test_pixel(const Tri& r, uint32_t t, float sx, float sy) {
    const float e0 = cross2(r.s1x - sx, r.s1y - sy, r.s2x - sx, r.s2y - sy);
    const float e1 = cross2(r.s2x - sx, r.s2y - sy, r.s0x - sx, r.s0y - sy);
    const float e2 = cross2(r.s0x - sx, r.s0y - sy, r.s1x - sx, r.s1y - sy);
    if (e0 * r.sgn < -1e-7f || e1 * r.sgn < -1e-7f || e2 * r.sgn < -1e-7f) return ~0ull;
    const float q0x = r.p0.x - sx * r.p0.w, q0y = r.p0.y - sy * r.p0.w;
    const float q1x = r.p1.x - sx * r.p1.w, q1y = r.p1.y - sy * r.p1.w;
    const float q2x = r.p2.x - sx * r.p2.w, q2y = r.p2.y - sy * r.p2.w;
    const float a0 = cross2(q1x, q1y, q2x, q2y);
    const float a1 = cross2(q2x, q2y, q0x, q0y);
    const float a2 = cross2(q0x, q0y, q1x, q1y);
    if (fabsf(a0 + a1 + a2) <= 1e-12f) return ~0ull;
    const float iw = r.p0.w * a0 + r.p1.w * a1 + r.p2.w * a2;
    if (fabsf(iw) <= 1e-12f) return ~0ull;
    const float iz = r.p0.z * a0 + r.p1.z * a1 + r.p2.z * a2;
    const float depth = clampf(iz / iw, -1.0f, 1.0f);
    return (static_cast<unsigned long long>(order_key(depth)) << 32) | static_cast<unsigned long long>(t + 1u);
}

// Intra-wave hand-off of LDS records written by other lanes of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Keys only decrease, so a stale read can only make the filter let an atomic through.
__device__ __forceinline__ void key_min(unsigned long long* p, unsigned long long key) {
    if (key < __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMin(p, key);
}

struct TriRec {
    float4 p0, p1, p2;
    float s0x, s0y, s1x, s1y, s2x, s2y, sgn;
    int32_t x0, y0;
    uint32_t w, h, t;
};

// One 64-triangle batch per wave: each lane sets up one triangle; rectangles of
// at most kInlinePixels are tested by their own lane, larger ones (up to
// kWavePixels) by the whole wave, one pixel per lane, and larger still are
// queued for k_raster_large.
__global__ __launch_bounds__(kBlock) void k_raster_small(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, uint32_t cull, const float* __restrict__ sxt,
                                                         const float* __restrict__ syt,
                                                         unsigned long long* __restrict__ keys,
                                                         uint2* __restrict__ queue, uint32_t* __restrict__ counters) {
    __shared__ TriRec recs[kBlock];
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    TriRec* wrecs = recs + wave * 64;
    unsigned long long* kbase = keys + vd.pixel_base;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t batch = (blockIdx.x * 4 + wave) * 64; batch < T; batch += stride) {
        const uint32_t t = batch + lane;
        Tri r;
        const bool ok = t < T && setup_forward(vd, idx, t, V, cull, r);
        const uint32_t w = ok ? static_cast<uint32_t>(r.x1 - r.x0 + 1) : 0u;
        const uint32_t h = ok ? static_cast<uint32_t>(r.y1 - r.y0 + 1) : 0u;
        const uint32_t area = w * h;
        bool deferred = false;
        if (ok && area <= kInlinePixels) {
            for (int32_t y = r.y0; y <= r.y1; ++y) {
                const float sy = syt[vd.row_base + y];
                for (int32_t x = r.x0; x <= r.x1; ++x) {
                    const unsigned long long key = test_pixel(r, t, sxt[vd.col_base + x], sy);
                    if (key != ~0ull) key_min(&kbase[static_cast<size_t>(y) * vd.width + x], key);
                }
            }
        } else if (ok && area > kWavePixels) {
            queue[atomicAdd(&counters[0], 1u)] = make_uint2(view, t);
        } else if (ok) {
            deferred = true;
            TriRec& q = wrecs[lane];
            q.p0 = r.p0;
            q.p1 = r.p1;
            q.p2 = r.p2;
            q.s0x = r.s0x; q.s0y = r.s0y; q.s1x = r.s1x; q.s1y = r.s1y; q.s2x = r.s2x; q.s2y = r.s2y;
            q.sgn = r.sgn;
            q.x0 = r.x0;
            q.y0 = r.y0;
            q.w = w;
            q.h = h;
            q.t = t;
        }
        wave_lds_sync();
        unsigned long long pending = __ballot(deferred);
        while (pending) {
            const uint32_t j = __ffsll(pending) - 1;
            pending &= pending - 1;
            const TriRec& q = wrecs[j];
            Tri rr;
            rr.p0 = q.p0; rr.p1 = q.p1; rr.p2 = q.p2;
            rr.s0x = q.s0x; rr.s0y = q.s0y; rr.s1x = q.s1x; rr.s1y = q.s1y; rr.s2x = q.s2x; rr.s2y = q.s2y;
            rr.sgn = q.sgn;
            const uint32_t qw = q.w, qa = q.w * q.h, qt = q.t;
            for (uint32_t c = 0; c < qa; c += 64) {
                const uint32_t i = c + lane;
                if (i < qa) {
                    const int32_t x = q.x0 + static_cast<int32_t>(i % qw), y = q.y0 + static_cast<int32_t>(i / qw);
                    const unsigned long long key = test_pixel(rr, qt, sxt[vd.col_base + x], syt[vd.row_base + y]);
                    if (key != ~0ull) key_min(&kbase[static_cast<size_t>(y) * vd.width + x], key);
                }
            }
        }
        wave_lds_sync();
    }
}

__global__ __launch_bounds__(kBlock) void k_raster_large(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, uint32_t cull, const float* __restrict__ sxt,
                                                         const float* __restrict__ syt,
                                                         unsigned long long* __restrict__ keys,
                                                         const uint2* __restrict__ queue,
                                                         uint32_t* __restrict__ counters) {
    __shared__ uint32_t entry;
    const uint32_t count = counters[0];
    while (true) {
        if (threadIdx.x == 0) entry = atomicAdd(&counters[1], 1u);
        __syncthreads();
        const uint32_t e = entry;
        __syncthreads();
        if (e >= count) return;
        const uint2 q = queue[e];
        const ViewDesc& vd = vt.v[q.x];
        Tri r;
        if (!setup_forward(vd, idx, q.y, V, cull, r)) continue;
        const uint32_t w = static_cast<uint32_t>(r.x1 - r.x0 + 1), h = static_cast<uint32_t>(r.y1 - r.y0 + 1);
        unsigned long long* base = keys + vd.pixel_base;
        for (uint32_t i = threadIdx.x; i < w * h; i += kBlock) {
            const int32_t x = r.x0 + static_cast<int32_t>(i % w), y = r.y0 + static_cast<int32_t>(i / w);
            const unsigned long long key = test_pixel(r, q.y, sxt[vd.col_base + x], syt[vd.row_base + y]);
            if (key != ~0ull) key_min(&base[static_cast<size_t>(y) * vd.width + x], key);
        }
    }
}

// Winner's outputs (rasterize_forward.hlsl:115-162, 165-174); resets the key.
__global__ __launch_bounds__(kBlock) void k_raster_resolve(ViewTable vt, const uint32_t* __restrict__ idx,
                                                           int want_deriv, const float* __restrict__ sxt,
                                                           const float* __restrict__ syt,
                                                           unsigned long long* __restrict__ keys) {
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= vt.total_pixels) return;
    const ViewDesc& vd = vt.v[find_view(vt, gp)];
    const uint32_t p = gp - vd.pixel_base;
    const unsigned long long key = keys[gp];
    keys[gp] = ~0ull;
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (key != ~0ull) {
        const uint32_t t = static_cast<uint32_t>(key & 0xFFFFFFFFull) - 1u;
        const float sx = sxt[vd.col_base + p % vd.width], sy = syt[vd.row_base + p / vd.width];
        const float4* P = reinterpret_cast<const float4*>(vd.pos);
        const float4 p0 = P[idx[t * 3 + 0]], p1 = P[idx[t * 3 + 1]], p2 = P[idx[t * 3 + 2]];
        const float q0x = p0.x - sx * p0.w, q0y = p0.y - sy * p0.w;
        const float q1x = p1.x - sx * p1.w, q1y = p1.y - sy * p1.w;
        const float q2x = p2.x - sx * p2.w, q2y = p2.y - sy * p2.w;
        const float a0 = cross2(q1x, q1y, q2x, q2y);
        const float a1 = cross2(q2x, q2y, q0x, q0y);
        const float a2 = cross2(q0x, q0y, q1x, q1y);
        const float inv = 1.0f / (a0 + a1 + a2);
        const float b0 = clampf(a0 * inv, 0.0f, 1.0f);
        const float b1 = clampf(a1 * inv, 0.0f, 1.0f);
        const float iz = p0.z * a0 + p1.z * a1 + p2.z * a2;
        const float iw = p0.w * a0 + p1.w * a1 + p2.w * a2;
        r = make_float4(b0, b1, clampf(iz / iw, -1.0f, 1.0f), static_cast<float>(t + 1u));
        if (want_deriv) {
            const float da0_dx = -p1.w * p2.y + p2.w * p1.y;
            const float da0_dy = -p1.x * p2.w + p1.w * p2.x;
            const float da1_dx = -p2.w * p0.y + p2.y * p0.w;
            const float da1_dy = -p2.x * p0.w + p2.w * p0.x;
            const float da2_dx = -p0.w * p1.y + p0.y * p1.w;
            const float da2_dy = -p0.x * p1.w + p0.w * p1.x;
            const float dA_dx = da0_dx + da1_dx + da2_dx;
            const float dA_dy = da0_dy + da1_dy + da2_dy;
            d.x = (da0_dx - b0 * dA_dx) * inv * 2.0f / vd.vw;
            d.z = (da1_dx - b1 * dA_dx) * inv * 2.0f / vd.vw;
            d.y = (da0_dy - b0 * dA_dy) * inv * 2.0f / vd.vh;
            d.w = (da1_dy - b1 * dA_dy) * inv * 2.0f / vd.vh;
        }
    }
    reinterpret_cast<float4*>(vd.raster)[p] = r;
    if (vd.deriv) reinterpret_cast<float4*>(vd.deriv)[p] = d;
}

// ----------------------------------------------------------------- backward

struct BwdTri {
    float4 p[3];
    int32_t x0, x1, y0, y1;  // [x0, x1) x [y0, y1), rasterize_backward.hlsl:75-87
};

__device__ bool setup_backward(const ViewDesc& vd, const uint32_t* __restrict__ idx, uint32_t t, uint32_t V,
                               BwdTri& b) {
    const uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    if (i0 >= V || i1 >= V || i2 >= V) return false;
    const float4* P = reinterpret_cast<const float4*>(vd.pos);
    b.p[0] = P[i0];
    b.p[1] = P[i1];
    b.p[2] = P[i2];
    if (b.p[0].w <= 1e-8f || b.p[1].w <= 1e-8f || b.p[2].w <= 1e-8f) return false;
    const float s0x = b.p[0].x / b.p[0].w, s0y = b.p[0].y / b.p[0].w;
    const float s1x = b.p[1].x / b.p[1].w, s1y = b.p[1].y / b.p[1].w;
    const float s2x = b.p[2].x / b.p[2].w, s2y = b.p[2].y / b.p[2].w;
    const float mnx = minf_std(s0x, minf_std(s1x, s2x)), mny = minf_std(s0y, minf_std(s1y, s2y));
    const float mxx = maxf_std(s0x, maxf_std(s1x, s2x)), mxy = maxf_std(s0y, maxf_std(s1y, s2y));
    auto cl = [](int32_t v, int32_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    const int32_t W = static_cast<int32_t>(vd.width), H = static_cast<int32_t>(vd.height);
    b.x0 = cl(f2i_sat(floorf((mnx * 0.5f + 0.5f) * vd.vw + vd.vx)), W);
    b.x1 = cl(f2i_sat(ceilf((mxx * 0.5f + 0.5f) * vd.vw + vd.vx)), W);
    b.y0 = cl(f2i_sat(floorf((mny * 0.5f + 0.5f) * vd.vh + vd.vy)), H);
    b.y1 = cl(f2i_sat(ceilf((mxy * 0.5f + 0.5f) * vd.vh + vd.vy)), H);
    return true;
}

// rasterize_backward.hlsl:96-151 for one pixel of triangle t; adds into g (x, y, w per corner).
__device__ __forceinline__ void corner_terms(const BwdTri& b, const ViewDesc& vd, uint32_t t, int32_t x, int32_t y,
                                             const float* __restrict__ sxt, const float* __restrict__ syt,
                                             float g[3][3]) {
    const size_t pix = static_cast<size_t>(y) * vd.width + static_cast<size_t>(x);
    if (decode_id(vd.raster[pix * 4 + 3]) != t + 1u) return;
    const float2 lg = reinterpret_cast<const float2*>(vd.grad_raster)[pix * 2];
    if (lg.x == 0.0f && lg.y == 0.0f) return;
    const float sx = sxt[vd.col_base + x], sy = syt[vd.row_base + y];
    const float q0x = b.p[0].x - sx * b.p[0].w, q0y = b.p[0].y - sy * b.p[0].w;
    const float q1x = b.p[1].x - sx * b.p[1].w, q1y = b.p[1].y - sy * b.p[1].w;
    const float q2x = b.p[2].x - sx * b.p[2].w, q2y = b.p[2].y - sy * b.p[2].w;
    const float area = cross2(q1x, q1y, q2x, q2y) + cross2(q2x, q2y, q0x, q0y) + cross2(q0x, q0y, q1x, q1y);
    if (fabsf(area) <= 1e-12f) return;
    const float eps = area >= 0.0f ? 1e-6f : -1e-6f;
    const float inv = 1.0f / (area + eps);
    const float a0 = cross2(q1x, q1y, q2x, q2y);
    const float a1 = cross2(q2x, q2y, q0x, q0y);
    const float b0 = a0 * inv, b1 = a1 * inv;
    const float gb0 = lg.x * inv, gb1 = lg.y * inv;
    const float gc = gb0 * b0 + gb1 * b1;
    const float g0x = gc * (q2y - q1y) - gb1 * q2y;
    const float g1x = gc * (q0y - q2y) + gb0 * q2y;
    const float g2x = gc * (q1y - q0y) - gb0 * q1y + gb1 * q0y;
    const float g0y = gc * (q1x - q2x) + gb1 * q2x;
    const float g1y = gc * (q2x - q0x) - gb0 * q2x;
    const float g2y = gc * (q0x - q1x) + gb0 * q1x - gb1 * q0x;
    g[0][0] += g0x;
    g[0][1] += g0y;
    g[0][2] += -sx * g0x - sy * g0y;
    g[1][0] += g1x;
    g[1][1] += g1y;
    g[1][2] += -sx * g1x - sy * g1y;
    g[2][0] += g2x;
    g[2][1] += g2y;
    g[2][2] += -sx * g2x - sy * g2y;
}

// Terms of one pixel (rasterize_backward.hlsl:94-151): returns false when the
// pixel does not contribute; t9 = (x, y, w) for corners 0..2.
__device__ __forceinline__ bool pixel_terms(const float4* p, const ViewDesc& vd, uint32_t t, int32_t x, int32_t y,
                                            const float* __restrict__ sxt, const float* __restrict__ syt,
                                            float t9[9]) {
    const size_t pix = static_cast<size_t>(y) * vd.width + static_cast<size_t>(x);
    if (decode_id(vd.raster[pix * 4 + 3]) != t + 1u) return false;
    const float2 lg = reinterpret_cast<const float2*>(vd.grad_raster)[pix * 2];
    if (lg.x == 0.0f && lg.y == 0.0f) return false;
    const float sx = sxt[vd.col_base + x], sy = syt[vd.row_base + y];
    const float q0x = p[0].x - sx * p[0].w, q0y = p[0].y - sy * p[0].w;
    const float q1x = p[1].x - sx * p[1].w, q1y = p[1].y - sy * p[1].w;
    const float q2x = p[2].x - sx * p[2].w, q2y = p[2].y - sy * p[2].w;
    const float area = cross2(q1x, q1y, q2x, q2y) + cross2(q2x, q2y, q0x, q0y) + cross2(q0x, q0y, q1x, q1y);
    if (fabsf(area) <= 1e-12f) return false;
    const float eps = area >= 0.0f ? 1e-6f : -1e-6f;
    const float inv = 1.0f / (area + eps);
    const float a0 = cross2(q1x, q1y, q2x, q2y);
    const float a1 = cross2(q2x, q2y, q0x, q0y);
    const float b0 = a0 * inv, b1 = a1 * inv;
    const float gb0 = lg.x * inv, gb1 = lg.y * inv;
    const float gc = gb0 * b0 + gb1 * b1;
    t9[0] = gc * (q2y - q1y) - gb1 * q2y;
    t9[1] = gc * (q1x - q2x) + gb1 * q2x;
    t9[3] = gc * (q0y - q2y) + gb0 * q2y;
    t9[4] = gc * (q2x - q0x) - gb0 * q2x;
    t9[6] = gc * (q1y - q0y) - gb0 * q1y + gb1 * q0y;
    t9[7] = gc * (q0x - q1x) + gb0 * q1x - gb1 * q0x;
    t9[2] = -sx * t9[0] - sy * t9[1];
    t9[5] = -sx * t9[3] - sy * t9[4];
    t9[8] = -sx * t9[6] - sy * t9[7];
    return true;
}

struct BoxRec {
    float4 p[3];
    int32_t x0, y0;
    uint32_t w, h, t;
};

// Per-triangle corner gradients summed over the reference's bounding box in
// its row-major order.  Small boxes: the triangle's own lane loops; medium:
// the wave tests 64 pixels at a time and adds the contributing pixels' terms
// in ascending pixel order (ballot + readlane), which is the reference's order.
__global__ __launch_bounds__(kBlock) void k_corner_small(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, const float* __restrict__ sxt,
                                                         const float* __restrict__ syt, float4* __restrict__ corner,
                                                         uint2* __restrict__ queue, uint32_t* __restrict__ counters) {
    __shared__ BoxRec recs[kBlock];
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    BoxRec* wrecs = recs + wave * 64;
    float4* cview = corner + static_cast<size_t>(view) * T * 3;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t batch = (blockIdx.x * 4 + wave) * 64; batch < T; batch += stride) {
        const uint32_t t = batch + lane;
        BwdTri b{};
        const bool ok = t < T && setup_backward(vd, idx, t, V, b) && b.x1 > b.x0 && b.y1 > b.y0;
        const uint32_t w = ok ? static_cast<uint32_t>(b.x1 - b.x0) : 0u, h = ok ? static_cast<uint32_t>(b.y1 - b.y0) : 0u;
        const uint32_t area = w * h;
        bool deferred = false;
        if (t < T && (!ok || area <= kInlineBox)) {
            float g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int32_t y = b.y0; ok && y < b.y1; ++y)
                for (int32_t x = b.x0; x < b.x1; ++x) {
                    float t9[9];
                    if (pixel_terms(b.p, vd, t, x, y, sxt, syt, t9))
#pragma unroll
                        for (int k = 0; k < 9; ++k) g[k] += t9[k];
                }
#pragma unroll
            for (int c = 0; c < 3; ++c) cview[t * 3 + c] = make_float4(g[c * 3], g[c * 3 + 1], 0.0f, g[c * 3 + 2]);
        } else if (ok && area > kWaveBox) {
            queue[atomicAdd(&counters[0], 1u)] = make_uint2(view, t);
        } else if (ok) {
            deferred = true;
            BoxRec& q = wrecs[lane];
            q.p[0] = b.p[0];
            q.p[1] = b.p[1];
            q.p[2] = b.p[2];
            q.x0 = b.x0;
            q.y0 = b.y0;
            q.w = w;
            q.h = h;
            q.t = t;
        }
        wave_lds_sync();
        unsigned long long pending = __ballot(deferred);
        while (pending) {
            const uint32_t j = __ffsll(pending) - 1;
            pending &= pending - 1;
            const BoxRec& q = wrecs[j];
            const float4 pj[3] = {q.p[0], q.p[1], q.p[2]};
            const uint32_t qw = q.w, qa = q.w * q.h, qt = q.t;
            float g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (uint32_t c = 0; c < qa; c += 64) {
                const uint32_t i = c + lane;
                float t9[9];
                bool hit = false;
                if (i < qa)
                    hit = pixel_terms(pj, vd, qt, q.x0 + static_cast<int32_t>(i % qw), q.y0 + static_cast<int32_t>(i / qw),
                                      sxt, syt, t9);
                unsigned long long m = __ballot(hit);
                while (m) {
                    const int src = __ffsll(m) - 1;
                    m &= m - 1;
#pragma unroll
                    for (int k = 0; k < 9; ++k)
                        g[k] += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t9[k]), src));
                }
            }
            if (lane == 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) cview[qt * 3 + c] = make_float4(g[c * 3], g[c * 3 + 1], 0.0f, g[c * 3 + 2]);
        }
        wave_lds_sync();
    }
}

// Large boxes: each lane sums a strided subset of the box, then a fixed-order
// tree over the workgroup (deterministic; not the reference's summation order).
__global__ __launch_bounds__(kBlock) void k_corner_large(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, const float* __restrict__ sxt,
                                                         const float* __restrict__ syt, float4* __restrict__ corner,
                                                         const uint2* __restrict__ queue,
                                                         uint32_t* __restrict__ counters) {
    __shared__ uint32_t entry;
    __shared__ float red[9][kBlock];
    const uint32_t count = counters[0];
    while (true) {
        if (threadIdx.x == 0) entry = atomicAdd(&counters[1], 1u);
        __syncthreads();
        const uint32_t e = entry;
        __syncthreads();
        if (e >= count) return;
        const uint2 q = queue[e];
        const ViewDesc& vd = vt.v[q.x];
        BwdTri b;
        setup_backward(vd, idx, q.y, V, b);
        const uint32_t w = static_cast<uint32_t>(b.x1 - b.x0), h = static_cast<uint32_t>(b.y1 - b.y0);
        float g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (uint32_t i = threadIdx.x; i < w * h; i += kBlock)
            corner_terms(b, vd, q.y, b.x0 + static_cast<int32_t>(i % w), b.y0 + static_cast<int32_t>(i / w), sxt,
                         syt, g);
#pragma unroll
        for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = g[k / 3][k % 3];
        __syncthreads();
        for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s)
#pragma unroll
                for (int k = 0; k < 9; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            float4* out = corner + (static_cast<size_t>(q.x) * T + q.y) * 3;
            for (int c = 0; c < 3; ++c) out[c] = make_float4(red[c * 3][0], red[c * 3 + 1][0], 0.0f, red[c * 3 + 2][0]);
        }
        __syncthreads();
    }
}

__global__ void k_count_corners(const uint32_t* __restrict__ idx, uint32_t n, uint32_t V, uint32_t* __restrict__ counts) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        if (idx[i] < V) atomicAdd(&counts[idx[i]], 1u);
}

// reduce_vertex_gradients.hlsl:12-29 for every view
__global__ __launch_bounds__(kBlock) void k_vertex_reduce(ViewTable vt, uint32_t T, uint32_t V,
                                                          const uint32_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ corners,
                                                          const float4* __restrict__ corner) {
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    const float4* cg = corner + static_cast<size_t>(view) * T * 3;
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock) {
        float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (uint32_t a = offsets[v]; a < offsets[v + 1]; ++a) {
            const float4 c = cg[corners[a]];
            s.x += c.x;
            s.y += c.y;
            s.z += c.z;
            s.w += c.w;
        }
        reinterpret_cast<float4*>(vd.grad_pos)[v] = s;
    }
}

__global__ void k_check_indices(const uint32_t* idx, size_t n, uint32_t V, uint32_t* flags) {
    bool bad = false;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        bad |= idx[i] >= V;
    if (__any(bad) && (threadIdx.x & 63u) == 0) atomicOr(flags, 1u);
}

uint32_t tri_grid(uint32_t T) { return (T + kBlock - 1) / kBlock < 8192 ? (T + kBlock - 1) / kBlock : 8192; }

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

hipError_t fill_keys(unsigned long long* p, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_u64<<<grid_for(n), kBlock, 0, st>>>(p, ~0ull, n);
    return hipGetLastError();
}

hipError_t check_indices(const uint32_t* idx, size_t n, uint32_t V, uint32_t* flags, hipStream_t st) {
    k_check_indices<<<grid_for(n), kBlock, 0, st>>>(idx, n, V, flags);
    return hipGetLastError();
}

hipError_t sample_tables(const ViewTable& vt, float* sx, float* sy, hipStream_t st) {
    uint32_t longest = 0;
    for (uint32_t k = 0; k < vt.n; ++k) longest = max(longest, max(vt.v[k].width, vt.v[k].height));
    k_sample_tables<<<dim3((longest + kBlock - 1) / kBlock, vt.n), kBlock, 0, st>>>(vt, sx, sy);
    return hipGetLastError();
}

hipError_t raster_candidates(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, uint32_t cull,
                             const RasterScratch& s, hipStream_t st) {
    hipError_t e = fill_u32(s.counters, 0u, 4, st);
    if (e != hipSuccess) return e;
    k_raster_small<<<dim3(tri_grid(T), vt.n), kBlock, 0, st>>>(vt, idx, T, V, cull, s.sx, s.sy, s.keys, s.queue,
                                                               s.counters);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_raster_large<<<kLargeBlocks, kBlock, 0, st>>>(vt, idx, T, V, cull, s.sx, s.sy, s.keys, s.queue, s.counters);
    return hipGetLastError();
}

hipError_t raster_resolve(const ViewTable& vt, const uint32_t* idx, int want_deriv, const RasterScratch& s,
                          hipStream_t st) {
    k_raster_resolve<<<(vt.total_pixels + kBlock - 1) / kBlock, kBlock, 0, st>>>(vt, idx, want_deriv, s.sx, s.sy,
                                                                                s.keys);
    return hipGetLastError();
}

size_t csr_temp_bytes(uint32_t T, uint32_t V) {
    size_t a = 0, b = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), static_cast<int>(V + 1));
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<int>(T * 3));
    return a > b ? a : b;
}

// Vertex -> corner adjacency with corners ascending per vertex
// (src/rasterizer.cpp:145-162): counts + exclusive scan for the offsets, a
// stable radix sort of (vertex, corner) pairs for the lists.
hipError_t build_vertex_csr(const uint32_t* idx, uint32_t T, uint32_t V, const VertexCsr& c, hipStream_t st) {
    const uint32_t n = T * 3;
    hipError_t e = fill_u32(c.counts, 0u, static_cast<size_t>(V) + 1, st);
    if (e != hipSuccess) return e;
    k_count_corners<<<grid_for(n), kBlock, 0, st>>>(idx, n, V, c.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t bytes = c.temp_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(c.temp, bytes, c.counts, c.offsets, static_cast<int>(V + 1), st)) !=
        hipSuccess)
        return e;
    k_iota<<<grid_for(n), kBlock, 0, st>>>(c.iota, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int bits = 1;
    while (bits < 32 && (1ull << bits) < V) ++bits;
    bytes = c.temp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(c.temp, bytes, idx, c.keys_out, c.iota, c.corners, static_cast<int>(n), 0,
                                              bits, st);
}

hipError_t corner_gradients(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, const RasterScratch& s,
                            float4* corner, hipStream_t st) {
    hipError_t e = fill_u32(s.counters, 0u, 4, st);
    if (e != hipSuccess) return e;
    k_corner_small<<<dim3(tri_grid(T), vt.n), kBlock, 0, st>>>(vt, idx, T, V, s.sx, s.sy, corner, s.queue, s.counters);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_corner_large<<<kLargeBlocks, kBlock, 0, st>>>(vt, idx, T, V, s.sx, s.sy, corner, s.queue, s.counters);
    return hipGetLastError();
}

hipError_t vertex_reduce(const ViewTable& vt, uint32_t T, uint32_t V, const VertexCsr& csr, const float4* corner,
                         hipStream_t st) {
    k_vertex_reduce<<<dim3(grid_for(V) < 4096 ? grid_for(V) : 4096, vt.n), kBlock, 0, st>>>(vt, T, V, csr.offsets,
                                                                                         csr.corners, corner);
    return hipGetLastError();
}

}  // namespace pdr
