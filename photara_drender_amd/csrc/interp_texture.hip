// Attribute interpolation and bilinear texture sampling, forward and backward.
//
// Forward kernels run one lane per pixel over all views of a batch
// (shaders/interpolate_forward.hlsl:15-54, texture_forward.hlsl:37-79) and
// produce the reference's values bit for bit.
//
// Backward kernels turn the reference's gathers into per-pixel scatters:
//  * interpolation: grad_raster is per pixel (interpolate_grad_raster.hlsl, exact);
//    the attribute gradient (interpolate_grad_attributes.hlsl loops over every
//    pixel for every (vertex, attribute): O(V*A*P)) becomes 3*A terms per pixel,
//    summed per vertex in an LDS hash table over a 64x64 tile and flushed with
//    one float atomic per (vertex, attribute) of the tile;
//  * texture: grad_uv is per pixel (texture_grad_uv.hlsl, exact); the texel
//    gradient (texture_grad_texture.hlsl over the CPU-built CSR of
//    src/rasterizer.cpp:187-258) becomes 4*C atomic terms per pixel, with the
//    reference's double-precision bilinear fractions.
// Each scattered term is the reference's term exactly; only the order in which
// terms are summed differs.
#include <hipcub/hipcub.hpp>

#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint32_t find_pixel_view(const PixelTable& pt, uint32_t gp) {
    uint32_t k = 0;
    while (k + 1 < pt.n && gp >= pt.v[k + 1].pixel_base) ++k;
    return k;
}

__device__ __forceinline__ uint32_t find_tile_view(const PixelTable& pt, uint32_t g) {
    uint32_t k = 0;
    while (k + 1 < pt.n && g >= pt.v[k + 1].tile_base) ++k;
    return k;
}

template <int kA>
__global__ __launch_bounds__(kBlock) void k_interp_fwd(PixelTable pt, const float* __restrict__ attr, uint32_t V,
                                                       uint32_t A_rt, const uint32_t* __restrict__ idx, uint32_t T) {
    const uint32_t A = kA > 0 ? static_cast<uint32_t>(kA) : A_rt;
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= pt.total_pixels) return;
    bool split;
    const uint32_t kv = wave_view(pt, gp, split);
    auto body = [&](const PixelView& pv) {
        const uint32_t p = gp - pv.pixel_base;
        const float4 r = reinterpret_cast<const float4*>(pv.raster)[p];
        float* out = pv.out0 + static_cast<size_t>(p) * A;
        const uint32_t id = decode_id(r.w);
        uint32_t i0 = 0, i1 = 0, i2 = 0;
        bool ok = id != 0 && id <= T;
        if (ok) {
            i0 = idx[(id - 1) * 3 + 0];
            i1 = idx[(id - 1) * 3 + 1];
            i2 = idx[(id - 1) * 3 + 2];
            ok = i0 < V && i1 < V && i2 < V;
        }
        if (!ok) {
            for (uint32_t a = 0; a < A; ++a) out[a] = 0.0f;
            return;
        }
        const float u = r.x, v = r.y;
        const float w = 1.0f - u - v;
        const float* a0 = attr + static_cast<size_t>(i0) * A;
        const float* a1 = attr + static_cast<size_t>(i1) * A;
        const float* a2 = attr + static_cast<size_t>(i2) * A;
        if constexpr (kA == 2) {
            reinterpret_cast<float2*>(out)[0] = make_float2(u * a0[0] + v * a1[0] + w * a2[0], u * a0[1] + v * a1[1] + w * a2[1]);
        } else {
            for (uint32_t a = 0; a < A; ++a) out[a] = u * a0[a] + v * a1[a] + w * a2[a];
        }
    };
    if (!split)
        body(pt.v[kv]);
    else
        body(pt.v[find_pixel_view(pt, gp)]);
}

// A = 8 (16-byte aligned rows): kN pixels per lane, gp, gp + part, ..., their
// dependent loads (raster -> triangle -> attribute rows) issued side by side so
// kN times as many are in flight; the values are k_interp_fwd's.
constexpr uint32_t kInterpPixels = 4;
template <int kN>
__global__ __launch_bounds__(kBlock) void k_interp_fwd8_multi(PixelTable pt, const float* __restrict__ attr,
                                                             uint32_t V, const uint32_t* __restrict__ idx, uint32_t T,
                                                             uint32_t part) {
    const uint32_t g0 = blockIdx.x * kBlock + threadIdx.x;
    if (g0 >= part) return;
    uint32_t gps[kN];
#pragma unroll
    for (int h = 0; h < kN; ++h) gps[h] = g0 + static_cast<uint32_t>(h) * part;
    bool live[kN];
    const PixelView* pv[kN];
#pragma unroll
    for (int h = 0; h < kN; ++h) {
        live[h] = gps[h] < pt.total_pixels;
        const uint32_t g = live[h] ? gps[h] : pt.total_pixels - 1u;
        bool split;
        const uint32_t kv = wave_view(pt, g, split);
        pv[h] = &pt.v[split ? find_pixel_view(pt, g) : kv];
    }
    float4 r[kN];
    uint32_t p[kN];
#pragma unroll
    for (int h = 0; h < kN; ++h) {
        const uint32_t g = live[h] ? gps[h] : pt.total_pixels - 1u;
        p[h] = g - pv[h]->pixel_base;
        r[h] = reinterpret_cast<const float4*>(pv[h]->raster)[p[h]];
    }
    uint32_t iv[kN][3];
    bool ok[kN];
#pragma unroll
    for (int h = 0; h < kN; ++h) {
        const uint32_t id = decode_id(r[h].w);
        ok[h] = live[h] && id != 0 && id <= T;
        iv[h][0] = iv[h][1] = iv[h][2] = 0;
        if (ok[h]) {
            iv[h][0] = idx[(id - 1) * 3 + 0];
            iv[h][1] = idx[(id - 1) * 3 + 1];
            iv[h][2] = idx[(id - 1) * 3 + 2];
        }
    }
    float4 row[kN][3][2];
#pragma unroll
    for (int h = 0; h < kN; ++h) {
        ok[h] = ok[h] && iv[h][0] < V && iv[h][1] < V && iv[h][2] < V;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            row[h][c][0] = row[h][c][1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (ok[h]) {
                const float4* a = reinterpret_cast<const float4*>(attr + static_cast<size_t>(iv[h][c]) * 8);
                row[h][c][0] = a[0];
                row[h][c][1] = a[1];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < kN; ++h) {
        if (!live[h]) continue;
        // the output is streamed with nontemporal stores
        typedef float nt4 __attribute__((ext_vector_type(4)));
        nt4* ot = reinterpret_cast<nt4*>(pv[h]->out0 + static_cast<size_t>(p[h]) * 8);
        if (!ok[h]) {
            __builtin_nontemporal_store(nt4{0.0f, 0.0f, 0.0f, 0.0f}, ot);
            __builtin_nontemporal_store(nt4{0.0f, 0.0f, 0.0f, 0.0f}, ot + 1);
            continue;
        }
        const float u = r[h].x, v = r[h].y;
        const float w = 1.0f - u - v;
        float a0[8], a1[8], a2[8];
        auto unpack = [](const float4 (&q)[2], float (&o)[8]) {
            o[0] = q[0].x; o[1] = q[0].y; o[2] = q[0].z; o[3] = q[0].w;
            o[4] = q[1].x; o[5] = q[1].y; o[6] = q[1].z; o[7] = q[1].w;
        };
        unpack(row[h][0], a0);
        unpack(row[h][1], a1);
        unpack(row[h][2], a2);
        float o[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) o[a] = u * a0[a] + v * a1[a] + w * a2[a];
        __builtin_nontemporal_store(nt4{o[0], o[1], o[2], o[3]}, ot);
        __builtin_nontemporal_store(nt4{o[4], o[5], o[6], o[7]}, ot + 1);
    }
}

constexpr uint32_t kVertexSlots = 1024;  // vertex accumulators per tile
constexpr uint32_t kMaxHashedAttrs = 8;  // attributes per accumulator

constexpr uint32_t kBucketSlots = 8;  // the table is 128 buckets of 8 slots

// First slot of the bucket of vertex key (index + 1).
__device__ __forceinline__ uint32_t vertex_bucket(uint32_t key) {
    return ((key * 2654435761u) >> 25) * kBucketSlots;
}

// Slot of `key` in its bucket given a snapshot of the bucket's keys (slots
// only ever go from 0 to a key): the slot holding it, else the first empty
// slot it wins (or finds taken by the same key), else kVertexSlots.
__device__ __forceinline__ uint32_t vertex_slot(uint32_t* keys, uint32_t key, uint32_t b, uint4 lo, uint4 hi) {
    uint32_t k[kBucketSlots] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
    for (uint32_t i = 0; i < kBucketSlots; ++i)
        if (k[i] == key) return b + i;
#pragma unroll
    for (uint32_t i = 0; i < kBucketSlots; ++i) {
        if (k[i] != 0u) continue;
        const uint32_t old = atomicCAS(&keys[b + i], 0u, key);
        if (old == 0u || old == key) return b + i;
    }
    return kVertexSlots;
}

// Adds the accumulated terms of up to three vertices (v[e] with on[e]) into
// the tile table, acc[a * kVertexSlots + slot] (or into global memory for a
// vertex whose bucket is full).  The three buckets are read together.
__device__ __forceinline__ void flush_vertices(uint32_t* keys, float* acc, float* __restrict__ grad_attr, uint32_t A,
                                               const uint32_t (&v)[3], const bool (&on)[3],
                                               const float (&val)[3][kMaxHashedAttrs]) {
    uint32_t b[3];
    uint4 lo[3], hi[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        b[e] = vertex_bucket(on[e] ? v[e] + 1u : 1u);
        lo[e] = *reinterpret_cast<const uint4*>(&keys[b[e]]);
        hi[e] = *reinterpret_cast<const uint4*>(&keys[b[e] + 4]);
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        if (!on[e]) continue;
        const uint32_t slot = vertex_slot(keys, v[e] + 1u, b[e], lo[e], hi[e]);
        if (slot < kVertexSlots) {
#pragma unroll
            for (uint32_t a = 0; a < kMaxHashedAttrs; ++a)
                if (a < A) atomicAdd(&acc[a * kVertexSlots + slot], val[e][a]);
        } else {
            float* dst = grad_attr + static_cast<size_t>(v[e]) * A;
#pragma unroll
            for (uint32_t a = 0; a < kMaxHashedAttrs; ++a)
                if (a < A) unsafeAtomicAdd(&dst[a], val[e][a]);
        }
    }
}

// One tile of kGradTile columns x 4 bands per workgroup (A <= kMaxHashedAttrs);
// each wave walks a band of pt.band_rows rows, one column per lane, top to bottom.
// grad_raster is written per pixel (interpolate_grad_raster.hlsl, exact).  The
// attribute-gradient terms b_c * g[p, a] (interpolate_grad_attributes.hlsl:
// 97-110) accumulate in registers per vertex of the lane's current triangle;
// consecutive pixels of a column mostly belong to the same or an adjacent
// triangle, so a vertex's sum leaves the registers only when the walk moves
// to a triangle without it.  It then goes into a per-tile LDS table keyed by
// vertex, and the table into global memory with one float atomic per
// (vertex, attribute) of the tile.  kA = 8: vector loads (16-byte aligned rows).
template <int kA>
__global__ __launch_bounds__(kBlock) void k_interp_bwd(PixelTable pt, const float* __restrict__ attr, uint32_t V,
                                                       uint32_t A_rt, const uint32_t* __restrict__ idx, uint32_t T,
                                                       float* __restrict__ grad_attr) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    constexpr uint32_t kM = kMaxHashedAttrs;
    __shared__ __align__(16) uint32_t keys[kVertexSlots];
    __shared__ float acc[kVertexSlots * kMaxHashedAttrs];
    const uint32_t A = kA > 0 ? static_cast<uint32_t>(kA) : A_rt;
    for (uint32_t s = threadIdx.x; s < kVertexSlots; s += kBlock) keys[s] = 0u;
    for (uint32_t s = threadIdx.x; s < kVertexSlots * A; s += kBlock) acc[s] = 0.0f;
    __syncthreads();
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const PixelView& pv = pt.v[find_tile_view(pt, g)];
    const uint32_t local = g - pv.tile_base;
    const uint32_t rows = pt.band_rows;
    const uint32_t px = (local % pv.tiles_x) * kGradTile + (threadIdx.x & 63u);
    const uint32_t band = (local / pv.tiles_x) * (rows * (kBlock / 64)) + (threadIdx.x >> 6) * rows;
    uint32_t cur = 0;                   // current triangle id (raster encoding), 0: none
    uint32_t ev[3] = {kNone, kNone, kNone};  // vertex of each accumulator = corner of cur
    float ea[3][kM];                    // per-corner accumulated terms
    float av[3][kM];                    // attribute rows of cur's corners (for grad_raster)
    bool in_range = false;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (uint32_t a = 0; a < kM; ++a) ea[c][a] = av[c][a] = 0.0f;
    for (uint32_t row = 0; row < rows; ++row) {
        const uint32_t py = band + row;
        if (px >= pv.width || py >= pv.height) break;
        const size_t p = static_cast<size_t>(py) * pv.width + px;
        const float4 r = reinterpret_cast<const float4*>(pv.raster)[p];
        const uint32_t id = decode_id(r.w);
        float gx = 0.0f, gy = 0.0f;
        if (id != 0 && id <= T) {
            if (id != cur) {
                uint32_t iv[3] = {idx[(id - 1) * 3 + 0], idx[(id - 1) * 3 + 1], idx[(id - 1) * 3 + 2]};
                in_range = iv[0] < V && iv[1] < V && iv[2] < V;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (iv[c] >= V) iv[c] = kNone;  // never matches (interpolate_grad_attributes.hlsl)
                // keep each accumulator whose vertex the new triangle has (each used once)
                int src[3] = {-1, -1, -1};
                uint32_t used = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int e = 0; e < 3; ++e)
                        if (src[c] < 0 && !((used >> e) & 1u) && ev[e] != kNone && ev[e] == iv[c]) {
                            src[c] = e;
                            used |= 1u << e;
                        }
                const bool evict[3] = {!(used & 1u) && ev[0] != kNone, !((used >> 1) & 1u) && ev[1] != kNone,
                                       !((used >> 2) & 1u) && ev[2] != kNone};
                flush_vertices(keys, acc, grad_attr, A, ev, evict, ea);
                float na[3][kM];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (uint32_t a = 0; a < kM; ++a)
                        na[c][a] = src[c] == 0 ? ea[0][a] : (src[c] == 1 ? ea[1][a] : (src[c] == 2 ? ea[2][a] : 0.0f));
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ev[c] = iv[c];
#pragma unroll
                    for (uint32_t a = 0; a < kM; ++a) ea[c][a] = na[c][a];
                }
                if (in_range) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const float* row_c = attr + static_cast<size_t>(iv[c]) * A;
                        if constexpr (kA == 8) {
                            const float4 x0 = reinterpret_cast<const float4*>(row_c)[0];
                            const float4 x1 = reinterpret_cast<const float4*>(row_c)[1];
                            av[c][0] = x0.x; av[c][1] = x0.y; av[c][2] = x0.z; av[c][3] = x0.w;
                            av[c][4] = x1.x; av[c][5] = x1.y; av[c][6] = x1.z; av[c][7] = x1.w;
                        } else {
#pragma unroll
                            for (uint32_t a = 0; a < kM; ++a)
                                if (a < A) av[c][a] = row_c[a];
                        }
                    }
                }
                cur = id;
            }
            float og[kM];
            const float* gin = pv.in0 + p * A;
            if constexpr (kA == 8) {
                const float4 g0 = reinterpret_cast<const float4*>(gin)[0];
                const float4 g1 = reinterpret_cast<const float4*>(gin)[1];
                og[0] = g0.x; og[1] = g0.y; og[2] = g0.z; og[3] = g0.w;
                og[4] = g1.x; og[5] = g1.y; og[6] = g1.z; og[7] = g1.w;
            } else {
#pragma unroll
                for (uint32_t a = 0; a < kM; ++a) og[a] = a < A ? gin[a] : 0.0f;
            }
            const float u = r.x, v = r.y;
            const float w = 1.0f - u - v;
#pragma unroll
            for (uint32_t a = 0; a < kM; ++a) {
                if (a >= A) break;
                if (in_range) {
                    // interpolate_grad_raster.hlsl:153-161
                    gx += og[a] * (av[0][a] - av[2][a]);
                    gy += og[a] * (av[1][a] - av[2][a]);
                }
                ea[0][a] += u * og[a];
                ea[1][a] += v * og[a];
                ea[2][a] += w * og[a];
            }
        }
        reinterpret_cast<float4*>(pv.out1)[p] = make_float4(gx, gy, 0.0f, 0.0f);
    }
    const bool held[3] = {ev[0] != kNone, ev[1] != kNone, ev[2] != kNone};
    flush_vertices(keys, acc, grad_attr, A, ev, held, ea);
    __syncthreads();
    // one (vertex, attribute) per lane, a vertex's attributes on neighbouring
    // lanes: each atomic instruction touches 8 vertex rows rather than 64
    for (uint32_t i = threadIdx.x; i < kVertexSlots * kMaxHashedAttrs; i += kBlock) {
        const uint32_t s = i / kMaxHashedAttrs, a = i % kMaxHashedAttrs;
        const uint32_t key = keys[s];
        if (!key || a >= A) continue;
        unsafeAtomicAdd(&grad_attr[static_cast<size_t>(key - 1u) * A + a], acc[a * kVertexSlots + s]);
    }
}

// A > kMaxHashedAttrs: per-pixel float atomics.
__global__ __launch_bounds__(kBlock) void k_interp_bwd_direct(PixelTable pt, const float* __restrict__ attr,
                                                              uint32_t V, uint32_t A, const uint32_t* __restrict__ idx,
                                                              uint32_t T, float* __restrict__ grad_attr) {
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= pt.total_pixels) return;
    bool split;
    const uint32_t kv = wave_view(pt, gp, split);
    auto body = [&](const PixelView& pv) {
        const uint32_t p = gp - pv.pixel_base;
        const float4 r = reinterpret_cast<const float4*>(pv.raster)[p];
        const uint32_t id = decode_id(r.w);
        const float* gin = pv.in0 + static_cast<size_t>(p) * A;
        float gx = 0.0f, gy = 0.0f;
        if (id != 0 && id <= T) {
            const uint32_t i0 = idx[(id - 1) * 3 + 0], i1 = idx[(id - 1) * 3 + 1], i2 = idx[(id - 1) * 3 + 2];
            const bool in_range = i0 < V && i1 < V && i2 < V;
            const float u = r.x, v = r.y;
            const float w = 1.0f - u - v;
            for (uint32_t a = 0; a < A; ++a) {
                const float og = gin[a];
                if (in_range) {
                    const float a0 = attr[static_cast<size_t>(i0) * A + a];
                    const float a1 = attr[static_cast<size_t>(i1) * A + a];
                    const float a2 = attr[static_cast<size_t>(i2) * A + a];
                    gx += og * (a0 - a2);
                    gy += og * (a1 - a2);
                }
                if (i0 < V) unsafeAtomicAdd(&grad_attr[static_cast<size_t>(i0) * A + a], u * og);
                if (i1 < V) unsafeAtomicAdd(&grad_attr[static_cast<size_t>(i1) * A + a], v * og);
                if (i2 < V) unsafeAtomicAdd(&grad_attr[static_cast<size_t>(i2) * A + a], w * og);
            }
        }
        reinterpret_cast<float4*>(pv.out1)[p] = make_float4(gx, gy, 0.0f, 0.0f);
    };
    if (!split)
        body(pt.v[kv]);
    else
        body(pt.v[find_pixel_view(pt, gp)]);
}

struct Bilinear {
    uint32_t o00, o10, o01, o11;  // texel indices
    float w0, w1, w2, w3;
    float fx, fy;
};

// texture_forward.hlsl:224-242 (float fractions)
__device__ __forceinline__ Bilinear bilinear_f32(float u, float v, uint32_t tw, uint32_t th, uint32_t mode) {
    const float tx = u * static_cast<float>(tw) - 0.5f;
    const float ty = v * static_cast<float>(th) - 0.5f;
    const float flx = floorf(tx), fly = floorf(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    Bilinear b;
    b.fx = tx - flx;
    b.fy = ty - fly;
    const int32_t x0 = address_coordinate(bx, tw, mode), x1 = address_coordinate(add1(bx), tw, mode);
    const int32_t y0 = address_coordinate(by, th, mode), y1 = address_coordinate(add1(by), th, mode);
    b.o00 = static_cast<uint32_t>(y0) * tw + static_cast<uint32_t>(x0);
    b.o10 = static_cast<uint32_t>(y0) * tw + static_cast<uint32_t>(x1);
    b.o01 = static_cast<uint32_t>(y1) * tw + static_cast<uint32_t>(x0);
    b.o11 = static_cast<uint32_t>(y1) * tw + static_cast<uint32_t>(x1);
    b.w0 = (1.0f - b.fx) * (1.0f - b.fy);
    b.w1 = b.fx * (1.0f - b.fy);
    b.w2 = (1.0f - b.fx) * b.fy;
    b.w3 = b.fx * b.fy;
    return b;
}

// build_texture_sample_adjacency (src/rasterizer.cpp:207-235): double fractions.
__device__ __forceinline__ Bilinear bilinear_f64(float u, float v, uint32_t tw, uint32_t th, uint32_t mode) {
    const double tx = static_cast<double>(u) * tw - 0.5;
    const double ty = static_cast<double>(v) * th - 0.5;
    const double flx = floor(tx), fly = floor(ty);
    const int32_t bx = d2i_sat(flx), by = d2i_sat(fly);
    Bilinear b;
    b.fx = static_cast<float>(tx - flx);
    b.fy = static_cast<float>(ty - fly);
    const int32_t x0 = address_coordinate(bx, tw, mode), x1 = address_coordinate(add1(bx), tw, mode);
    const int32_t y0 = address_coordinate(by, th, mode), y1 = address_coordinate(add1(by), th, mode);
    b.o00 = static_cast<uint32_t>(y0) * tw + static_cast<uint32_t>(x0);
    b.o10 = static_cast<uint32_t>(y0) * tw + static_cast<uint32_t>(x1);
    b.o01 = static_cast<uint32_t>(y1) * tw + static_cast<uint32_t>(x0);
    b.o11 = static_cast<uint32_t>(y1) * tw + static_cast<uint32_t>(x1);
    b.w0 = (1.0f - b.fx) * (1.0f - b.fy);
    b.w1 = b.fx * (1.0f - b.fy);
    b.w2 = (1.0f - b.fx) * b.fy;
    b.w3 = b.fx * b.fy;
    return b;
}

template <int kC>
__global__ __launch_bounds__(kBlock) void k_texture_fwd(PixelTable pt, const float* __restrict__ tex, uint32_t tw,
                                                        uint32_t th, uint32_t C_rt, uint32_t mode) {
    const uint32_t C = kC > 0 ? static_cast<uint32_t>(kC) : C_rt;
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= pt.total_pixels) return;
    bool split;
    const uint32_t kv = wave_view(pt, gp, split);
    auto body = [&](const PixelView& pv) {
        const uint32_t p = gp - pv.pixel_base;
        float* out = pv.out0 + static_cast<size_t>(p) * C;
        if (decode_id(pv.raster[static_cast<size_t>(p) * 4 + 3]) == 0) {
            if constexpr (kC == 4) {
                reinterpret_cast<float4*>(out)[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            } else {
                for (uint32_t c = 0; c < C; ++c) out[c] = 0.0f;
            }
            return;
        }
        const float2 uv = reinterpret_cast<const float2*>(pv.in0)[p];
        const Bilinear b = bilinear_f32(uv.x, uv.y, tw, th, mode);
        if constexpr (kC == 4) {
            const float4* T4 = reinterpret_cast<const float4*>(tex);
            const float4 c00 = T4[b.o00], c10 = T4[b.o10], c01 = T4[b.o01], c11 = T4[b.o11];
            reinterpret_cast<float4*>(out)[0] =
                make_float4(c00.x * b.w0 + c10.x * b.w1 + c01.x * b.w2 + c11.x * b.w3,
                            c00.y * b.w0 + c10.y * b.w1 + c01.y * b.w2 + c11.y * b.w3,
                            c00.z * b.w0 + c10.z * b.w1 + c01.z * b.w2 + c11.z * b.w3,
                            c00.w * b.w0 + c10.w * b.w1 + c01.w * b.w2 + c11.w * b.w3);
        } else {
            for (uint32_t c = 0; c < C; ++c)
                out[c] = tex[static_cast<size_t>(b.o00) * C + c] * b.w0 + tex[static_cast<size_t>(b.o10) * C + c] * b.w1 +
                         tex[static_cast<size_t>(b.o01) * C + c] * b.w2 + tex[static_cast<size_t>(b.o11) * C + c] * b.w3;
        }
    };
    if (!split)
        body(pt.v[kv]);
    else
        body(pt.v[find_pixel_view(pt, gp)]);
}

// texture_grad_texture (scatter form) + texture_grad_uv for one view.  For
// C = 4 the texel atomics are issued in (pixel, tap, channel) lane order, 4
// pixels per instruction, each pixel's bilinear taps and weights passed
// through a per-wave LDS stage: a texel's four channels sit on neighbouring
// lanes and an instruction covers the 2 x 2 footprints of 4 neighbouring
// pixels, so it touches few cache lines (MI355X performs these atomics at the
// memory side, at a cost per line touched).  (pixel, channel) order, 16 pixels
// per instruction, one tap each, took 34.0 us instead of 26.9 us on the 512^2
// chain.
template <int kC>
__global__ __launch_bounds__(kBlock) void k_texture_bwd(PixelTable pt, const float* __restrict__ tex, uint32_t tw,
                                                        uint32_t th, uint32_t C_rt, uint32_t mode, int compute_uv,
                                                        float* __restrict__ grad_tex) {
    const uint32_t C = kC > 0 ? static_cast<uint32_t>(kC) : C_rt;
    const uint32_t gp0 = blockIdx.x * kBlock + threadIdx.x;
    const bool live = gp0 < pt.total_pixels;  // every lane runs (the C = 4 scatter exchanges across lanes)
    const uint32_t gp = live ? gp0 : pt.total_pixels - 1u;
    bool split;
    const uint32_t kv = wave_view(pt, gp, split);
    auto body = [&](const PixelView& pv) {
        const uint32_t p = gp - pv.pixel_base;
        const float idf = pv.raster[static_cast<size_t>(p) * 4 + 3];
        const float2 uv = reinterpret_cast<const float2*>(pv.in0)[p];
        const float* g = pv.in1 + static_cast<size_t>(p) * C;
        const bool finite = __builtin_isfinite(uv.x) && __builtin_isfinite(uv.y);
        const bool scatter = live && idf != 0.0f && finite;
        if constexpr (kC == 4) {
            Bilinear b{};
            float4 g4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (scatter) {
                b = bilinear_f64(uv.x, uv.y, tw, th, mode);
                g4 = make_float4(g[0], g[1], g[2], g[3]);
            }
            __shared__ uint2 taps[kBlock / 64][64][4];
            __shared__ float4 grads[kBlock / 64][64];
            const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u, c = lane & 3u, k = (lane >> 2) & 3u;
            const uint32_t none = 0xFFFFFFFFu;
            taps[wv][lane][0] = make_uint2(scatter ? b.o00 : none, __float_as_uint(b.w0));
            taps[wv][lane][1] = make_uint2(scatter ? b.o10 : none, __float_as_uint(b.w1));
            taps[wv][lane][2] = make_uint2(scatter ? b.o01 : none, __float_as_uint(b.w2));
            taps[wv][lane][3] = make_uint2(scatter ? b.o11 : none, __float_as_uint(b.w3));
            grads[wv][lane] = g4;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 4
            for (uint32_t q = 0; q < 16; ++q) {
                const uint32_t src = q * 4u + (lane >> 4);
                const uint2 t = taps[wv][src][k];
                const float4 gg = grads[wv][src];
                const float gv = c == 0 ? gg.x : (c == 1 ? gg.y : (c == 2 ? gg.z : gg.w));
                if (t.x != none) unsafeAtomicAdd(&grad_tex[static_cast<size_t>(t.x) * 4 + c], __uint_as_float(t.y) * gv);
            }
        } else if (scatter) {
            const Bilinear b = bilinear_f64(uv.x, uv.y, tw, th, mode);
            for (uint32_t c = 0; c < C; ++c) {
                const float gv = g[c];
                unsafeAtomicAdd(&grad_tex[static_cast<size_t>(b.o00) * C + c], b.w0 * gv);
                unsafeAtomicAdd(&grad_tex[static_cast<size_t>(b.o10) * C + c], b.w1 * gv);
                unsafeAtomicAdd(&grad_tex[static_cast<size_t>(b.o01) * C + c], b.w2 * gv);
                unsafeAtomicAdd(&grad_tex[static_cast<size_t>(b.o11) * C + c], b.w3 * gv);
            }
        }
        if (!compute_uv || !live) return;
        float2 gu = make_float2(0.0f, 0.0f);
        if (decode_id(idf) != 0) {
            const Bilinear b = bilinear_f32(uv.x, uv.y, tw, th, mode);
            for (uint32_t c = 0; c < C; ++c) {
                const float c00 = tex[static_cast<size_t>(b.o00) * C + c];
                const float c10 = tex[static_cast<size_t>(b.o10) * C + c];
                const float c01 = tex[static_cast<size_t>(b.o01) * C + c];
                const float c11 = tex[static_cast<size_t>(b.o11) * C + c];
                const float gv = g[c];
                const float dx = (1.0f - b.fy) * (c10 - c00) + b.fy * (c11 - c01);
                const float dy = (1.0f - b.fx) * (c01 - c00) + b.fx * (c11 - c10);
                gu.x += gv * (dx * static_cast<float>(tw));
                gu.y += gv * (dy * static_cast<float>(th));
            }
        }
        reinterpret_cast<float2*>(pv.out1)[p] = gu;
    };
    if (!split)
        body(pt.v[kv]);
    else
        body(pt.v[find_pixel_view(pt, gp)]);
}

// src/rasterizer.cpp:198-211 checks, reported as the first offending pixel of
// each kind (flags[0]: non-finite uv, flags[1]: |texel coordinate| > 1e9).
__global__ __launch_bounds__(kBlock) void k_texture_check_uv(PixelTable pt, uint32_t tw, uint32_t th,
                                                             uint32_t* flags) {
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= pt.total_pixels) return;
    bool split;
    const uint32_t kv = wave_view(pt, gp, split);
    auto body = [&](const PixelView& pv) {
        const uint32_t p = gp - pv.pixel_base;
        if (pv.raster[static_cast<size_t>(p) * 4 + 3] == 0.0f) return;
        const float2 uv = reinterpret_cast<const float2*>(pv.in0)[p];
        if (!__builtin_isfinite(uv.x) || !__builtin_isfinite(uv.y)) {
            atomicMin(&flags[0], gp);
            return;
        }
        const double tx = static_cast<double>(uv.x) * tw - 0.5, ty = static_cast<double>(uv.y) * th - 0.5;
        if (fabs(tx) > 1.0e9 || fabs(ty) > 1.0e9) atomicMin(&flags[1], gp);
    };
    if (!split)
        body(pt.v[kv]);
    else
        body(pt.v[find_pixel_view(pt, gp)]);
}

uint32_t pixel_grid(uint32_t n) { return (n + kBlock - 1) / kBlock; }


// ------------------------------------------------------ deterministic gradients
// PDR_FLAG_DETERMINISTIC: the attribute and texel gradients as ordered gathers
// instead of atomics.  Each covered pixel emits one (vertex, pixel) entry per
// distinct vertex of its triangle, or one (texel, 4 * pixel + tap) entry per
// bilinear tap, in slots ordered by pixel; unused slots carry a key past the
// last vertex / texel.  A stable radix sort by key groups the entries with each
// group in ascending pixel order, a binary search gives each group's start,
// and one lane per vertex / texel adds its entries in that order: the sums of
// the reference's per-vertex pixel loop and of its texel adjacency lists
// (src/rasterizer.cpp:187-258, texture_grad_texture.hlsl:15-38), bit for bit.

__global__ void k_det_attr_entries(PixelTable pt, const uint32_t* __restrict__ idx, uint32_t T, uint32_t V,
                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= pt.total_pixels) return;
    const PixelView& pv = pt.v[find_pixel_view(pt, gp)];
    const uint32_t id = decode_id(pv.raster[static_cast<size_t>(gp - pv.pixel_base) * 4 + 3]);
    uint32_t tri[3] = {V, V, V};
    if (id != 0u && id <= T)
        for (int c = 0; c < 3; ++c) tri[c] = idx[static_cast<size_t>(id - 1u) * 3 + c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        bool seen = false;
        for (int k = 0; k < c; ++k) seen |= tri[k] == tri[c];
        keys[static_cast<size_t>(gp) * 3 + c] = !seen && tri[c] < V ? tri[c] : V;
        vals[static_cast<size_t>(gp) * 3 + c] = gp;
    }
}

__global__ void k_det_tex_entries(PixelTable pt, uint32_t tw, uint32_t th, uint32_t mode, uint32_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
    const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
    if (gp >= pt.total_pixels) return;
    const PixelView& pv = pt.v[find_pixel_view(pt, gp)];
    const uint32_t p = gp - pv.pixel_base;
    const float2 uv = reinterpret_cast<const float2*>(pv.in0)[p];
    const bool on = pv.raster[static_cast<size_t>(p) * 4 + 3] != 0.0f && __builtin_isfinite(uv.x) &&
                    __builtin_isfinite(uv.y);
    const uint32_t none = tw * th;
    uint32_t o[4] = {none, none, none, none};
    if (on) {
        const Bilinear b = bilinear_f64(uv.x, uv.y, tw, th, mode);
        o[0] = b.o00;
        o[1] = b.o10;
        o[2] = b.o01;
        o[3] = b.o11;
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        keys[static_cast<size_t>(gp) * 4 + k] = o[k];
        vals[static_cast<size_t>(gp) * 4 + k] = gp * 4u + k;
    }
}

// starts[q] = first entry with key >= q (sorted keys), q = 0 .. n
__global__ void k_det_starts(const uint32_t* __restrict__ keys, uint32_t E, uint32_t n, uint32_t* __restrict__ starts) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q > n) return;
    uint32_t lo = 0, hi = E;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (keys[mid] < q)
            lo = mid + 1u;
        else
            hi = mid;
    }
    starts[q] = lo;
}

// orc / reference order: weight = [tri0 == v] u + [tri1 == v] w + [tri2 == v] (1 - u - w), summed over the
// vertex's pixels in ascending order, per attribute.
__global__ void k_det_attr_gather(PixelTable pt, const uint32_t* __restrict__ idx, uint32_t V, uint32_t A,
                                  const uint32_t* __restrict__ starts, const uint32_t* __restrict__ vals,
                                  float* __restrict__ grad_attr) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const uint32_t e0 = starts[v], e1 = starts[v + 1];
    for (uint32_t a0 = 0; a0 < A; a0 += 8) {
        float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t gp = vals[e];
            const PixelView& pv = pt.v[find_pixel_view(pt, gp)];
            const size_t p = gp - pv.pixel_base;
            const float4 r = reinterpret_cast<const float4*>(pv.raster)[p];
            const uint32_t* tri = idx + static_cast<size_t>(decode_id(r.w) - 1u) * 3;
            float weight = 0.0f;
            if (tri[0] == v) weight += r.x;
            if (tri[1] == v) weight += r.y;
            if (tri[2] == v) weight += 1.0f - r.x - r.y;
            const float* g = pv.in0 + p * A;
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (a0 + j < A) acc[j] += weight * g[a0 + j];
        }
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j)
            if (a0 + j < A) grad_attr[static_cast<size_t>(v) * A + a0 + j] = acc[j];
    }
}

// texture_grad_texture.hlsl:15-38 over the texel's (pixel, tap) entries in ascending order.
__global__ void k_det_tex_gather(PixelTable pt, uint32_t tw, uint32_t th, uint32_t C, uint32_t mode,
                                 const uint32_t* __restrict__ starts, const uint32_t* __restrict__ vals,
                                 float* __restrict__ grad_tex) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= tw * th) return;
    const uint32_t e0 = starts[t], e1 = starts[t + 1];
    for (uint32_t c0 = 0; c0 < C; c0 += 4) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t gp = vals[e] >> 2, k = vals[e] & 3u;
            const PixelView& pv = pt.v[find_pixel_view(pt, gp)];
            const size_t p = gp - pv.pixel_base;
            const float2 uv = reinterpret_cast<const float2*>(pv.in0)[p];
            const Bilinear b = bilinear_f64(uv.x, uv.y, tw, th, mode);
            const float w = k == 0u ? b.w0 : (k == 1u ? b.w1 : (k == 2u ? b.w2 : b.w3));
            const float* g = pv.in1 + p * C;
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
                if (c0 + j < C) acc[j] += w * g[c0 + j];
        }
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
            if (c0 + j < C) grad_tex[static_cast<size_t>(t) * C + c0 + j] = acc[j];
    }
}

int key_bits_for(uint32_t n) {  // bits covering keys 0 .. n
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= n) ++bits;
    return bits;
}

}  // namespace

hipError_t interp_forward(const PixelTable& pt, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                          uint32_t T, hipStream_t st) {
    if (!pt.total_pixels) return hipSuccess;
    const uint32_t grid = pixel_grid(pt.total_pixels);
    bool aligned = (reinterpret_cast<uintptr_t>(attr) % 16) == 0;
    for (uint32_t k = 0; k < pt.n; ++k) aligned &= (reinterpret_cast<uintptr_t>(pt.v[k].out0) % 16) == 0;
    if (A == 8 && aligned) {
        const uint32_t part = (pt.total_pixels + kInterpPixels - 1u) / kInterpPixels;
        k_interp_fwd8_multi<kInterpPixels><<<pixel_grid(part), kBlock, 0, st>>>(pt, attr, V, idx, T, part);
    } else if (A == 2 && aligned)
        k_interp_fwd<2><<<grid, kBlock, 0, st>>>(pt, attr, V, A, idx, T);
    else
        k_interp_fwd<0><<<grid, kBlock, 0, st>>>(pt, attr, V, A, idx, T);
    return hipGetLastError();
}

hipError_t interp_backward(const PixelTable& table, const float* attr, uint32_t V, uint32_t A, const uint32_t* idx,
                           uint32_t T, float* grad_attr, hipStream_t st) {
    if (!table.total_tiles) return hipSuccess;
    // Tiles of kGradTile columns x 4 bands: 16-row bands, halved (down to 2
    // rows) while that leaves fewer than 2048 tiles, so small views still
    // spread over the chip.
    PixelTable pt = table;
    auto tiles_for = [&](uint32_t rows) {
        uint64_t n = 0;
        for (uint32_t k = 0; k < pt.n; ++k)
            n += static_cast<uint64_t>(pt.v[k].tiles_x) * ((pt.v[k].height + rows * 4 - 1) / (rows * 4));
        return n;
    };
    uint32_t rows = kGradTile / (kBlock / 64);
    while (rows > 2 && tiles_for(rows) < 2048) rows /= 2;
    pt.band_rows = rows;
    uint32_t base = 0;
    for (uint32_t k = 0; k < pt.n; ++k) {
        pt.v[k].tile_base = base;
        base += pt.v[k].tiles_x * ((pt.v[k].height + rows * 4 - 1) / (rows * 4));
    }
    pt.total_tiles = base;
    bool aligned = (reinterpret_cast<uintptr_t>(attr) % 16) == 0;
    for (uint32_t k = 0; k < pt.n; ++k) aligned &= (reinterpret_cast<uintptr_t>(pt.v[k].in0) % 16) == 0;
    if (A == 8 && aligned)
        k_interp_bwd<8><<<pt.total_tiles, kBlock, 0, st>>>(pt, attr, V, A, idx, T, grad_attr);
    else if (A <= kMaxHashedAttrs)
        k_interp_bwd<0><<<pt.total_tiles, kBlock, 0, st>>>(pt, attr, V, A, idx, T, grad_attr);
    else
        k_interp_bwd_direct<<<pixel_grid(pt.total_pixels), kBlock, 0, st>>>(pt, attr, V, A, idx, T, grad_attr);
    return hipGetLastError();
}

hipError_t texture_forward(const PixelTable& pt, const float* tex, uint32_t tw, uint32_t th, uint32_t C, uint32_t mode,
                           hipStream_t st) {
    if (!pt.total_pixels) return hipSuccess;
    const uint32_t grid = pixel_grid(pt.total_pixels);
    bool aligned = (reinterpret_cast<uintptr_t>(tex) % 16) == 0;
    for (uint32_t k = 0; k < pt.n; ++k) aligned &= (reinterpret_cast<uintptr_t>(pt.v[k].out0) % 16) == 0;
    if (C == 4 && aligned)
        k_texture_fwd<4><<<grid, kBlock, 0, st>>>(pt, tex, tw, th, C, mode);
    else
        k_texture_fwd<0><<<grid, kBlock, 0, st>>>(pt, tex, tw, th, C, mode);
    return hipGetLastError();
}

hipError_t texture_backward(const PixelTable& pt, const float* tex, uint32_t tw, uint32_t th, uint32_t C,
                            uint32_t mode, int compute_uv, float* grad_tex, hipStream_t st) {
    if (!pt.total_pixels) return hipSuccess;
    const uint32_t grid = pixel_grid(pt.total_pixels);
    if (C == 4)
        k_texture_bwd<4><<<grid, kBlock, 0, st>>>(pt, tex, tw, th, C, mode, compute_uv, grad_tex);
    else if (C == 3)
        k_texture_bwd<3><<<grid, kBlock, 0, st>>>(pt, tex, tw, th, C, mode, compute_uv, grad_tex);
    else
        k_texture_bwd<0><<<grid, kBlock, 0, st>>>(pt, tex, tw, th, C, mode, compute_uv, grad_tex);
    return hipGetLastError();
}

hipError_t texture_check_uv(const PixelTable& pt, uint32_t tw, uint32_t th, uint32_t* flags, hipStream_t st) {
    if (!pt.total_pixels) return hipSuccess;
    k_texture_check_uv<<<pixel_grid(pt.total_pixels), kBlock, 0, st>>>(pt, tw, th, flags);
    return hipGetLastError();
}

size_t deterministic_temp_bytes(uint32_t entries) {
    size_t a = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<int>(entries), 0, 32);
    return a;
}

static hipError_t det_sort_starts(const DeterministicScratch& s, uint32_t E, uint32_t nkeys, hipStream_t st) {
    size_t bytes = s.temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(s.temp, bytes, s.keys_in, s.keys_out, s.vals_in, s.vals_out,
                                                      static_cast<int>(E), 0, key_bits_for(nkeys), st);
    if (e != hipSuccess) return e;
    k_det_starts<<<(nkeys + 1 + kBlock - 1) / kBlock, kBlock, 0, st>>>(s.keys_out, E, nkeys, s.starts);
    return hipGetLastError();
}

hipError_t interp_backward_deterministic(const PixelTable& pt, uint32_t V, uint32_t A, const uint32_t* idx, uint32_t T,
                                         float* grad_attr, const DeterministicScratch& s, hipStream_t st) {
    const uint32_t E = pt.total_pixels * 3u;
    if (E) {
        k_det_attr_entries<<<pixel_grid(pt.total_pixels), kBlock, 0, st>>>(pt, idx, T, V, s.keys_in, s.vals_in);
        hipError_t e = det_sort_starts(s, E, V, st);
        if (e != hipSuccess) return e;
    } else {
        k_det_starts<<<(V + 1 + kBlock - 1) / kBlock, kBlock, 0, st>>>(s.keys_out, 0, V, s.starts);
    }
    k_det_attr_gather<<<(V + kBlock - 1) / kBlock, kBlock, 0, st>>>(pt, idx, V, A, s.starts, s.vals_out, grad_attr);
    return hipGetLastError();
}

hipError_t texture_backward_deterministic(const PixelTable& pt, uint32_t tw, uint32_t th, uint32_t C, uint32_t mode,
                                          float* grad_tex, const DeterministicScratch& s, hipStream_t st) {
    const uint32_t E = pt.total_pixels * 4u, texels = tw * th;
    if (E) {
        k_det_tex_entries<<<pixel_grid(pt.total_pixels), kBlock, 0, st>>>(pt, tw, th, mode, s.keys_in, s.vals_in);
        hipError_t e = det_sort_starts(s, E, texels, st);
        if (e != hipSuccess) return e;
    } else {
        k_det_starts<<<(texels + 1 + kBlock - 1) / kBlock, kBlock, 0, st>>>(s.keys_out, 0, texels, s.starts);
    }
    k_det_tex_gather<<<(texels + kBlock - 1) / kBlock, kBlock, 0, st>>>(pt, tw, th, C, mode, s.starts, s.vals_out,
                                                                      grad_tex);
    return hipGetLastError();
}

}  // namespace pdr
