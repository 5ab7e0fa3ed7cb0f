// Texture bake (TextureBaker::bake, src/texture_baker.cpp:156-361) for gfx950.
//
// The host side rasterizes the mesh into the atlas (uv * 2 - 1 as clip
// positions), interpolates world positions and normals per texel and
// rasterizes every view for its shadow map, all with the library's raster and
// interpolation kernels.  k_bake_project then runs
// shaders/project_texture.hlsl:117-247 for one texel per lane over the views
// in order, keeping the texel's blend state (accumulated colour, weight, best
// confidence, source view) in registers across the views of a launch and in
// global memory between launches; views are projected in the reference's
// dispatch order, so each texel sees the same sequence of updates.
#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr uint32_t kBlock = 256;

__global__ void k_atlas_clip(const float* __restrict__ uv, uint32_t V, float4* __restrict__ out) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock)
        out[v] = make_float4(uv[v * 2] * 2.0f - 1.0f, uv[v * 2 + 1] * 2.0f - 1.0f, 0.0f, 1.0f);
}

// lerp(a, b, t) = a + t * (b - a)
__device__ __forceinline__ float bake_lerp(float a, float b, float t) { return a + t * (b - a); }

// project_texture.hlsl:36-73: bilinear sample, coordinates clamped to the image,
// of every channel at once (a tap's channels are adjacent in memory and are
// loaded together); alpha = 1 for a 3-channel photo.
__device__ __forceinline__ void bake_photo_all(const BakeView& v, float px, float py, float out[4]) {
    const float tx = px - 0.5f, ty = py - 0.5f;
    const float flx = floorf(tx), fly = floorf(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    const float fx = tx - flx, fy = ty - fly;
    const int32_t wm = static_cast<int32_t>(v.width) - 1, hm = static_cast<int32_t>(v.height) - 1;
    auto tap = [&](int32_t x, int32_t y) {
        x = x < 0 ? 0 : (x > wm ? wm : x);
        y = y < 0 ? 0 : (y > hm ? hm : y);
        return v.photo + (static_cast<size_t>(y) * v.width + x) * v.channels;
    };
    const float* t00 = tap(bx, by);
    const float* t10 = tap(add1(bx), by);
    const float* t01 = tap(bx, add1(by));
    const float* t11 = tap(add1(bx), add1(by));
    const uint32_t nc = v.channels == 4 ? 4u : 3u;
    float a[4][4];
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) {
        if (c >= nc) break;
        a[0][c] = t00[c];
        a[1][c] = t10[c];
        a[2][c] = t01[c];
        a[3][c] = t11[c];
    }
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c)
        out[c] = c < nc ? bake_lerp(bake_lerp(a[0][c], a[1][c], fx), bake_lerp(a[2][c], a[3][c], fx), fy) : 1.0f;
}

__device__ __forceinline__ float bake_mask(const BakeView& v, float px, float py) {
    const float tx = px - 0.5f, ty = py - 0.5f;
    const float flx = floorf(tx), fly = floorf(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    const float fx = tx - flx, fy = ty - fly;
    const int32_t wm = static_cast<int32_t>(v.width) - 1, hm = static_cast<int32_t>(v.height) - 1;
    auto at = [&](int32_t x, int32_t y) {
        x = x < 0 ? 0 : (x > wm ? wm : x);
        y = y < 0 ? 0 : (y > hm ? hm : y);
        const float m = v.mask[static_cast<size_t>(y) * v.width + x];
        return v.mask_floor > 0.0f ? (m < v.mask_floor ? v.mask_floor : m) : m;  // texture_baker.cpp:282-291
    };
    return bake_lerp(bake_lerp(at(bx, by), at(add1(bx), by), fx), bake_lerp(at(bx, add1(by)), at(add1(bx), add1(by)), fx),
                     fy);
}

// A shadow-map texel from the view's raster key: (depth, 1) where a triangle
// covers the pixel, (0, 0) where none does.  The key's high word is the
// winner's depth clamp(iz / iw, -1, 1) under an order-preserving map (-0 as
// +0), inverted here exactly - the depth the raster's resolve pass would
// write, which the bake then needs no pass for.
__device__ __forceinline__ float2 shadow_texel(unsigned long long key) {
    if (key == ~0ull) return make_float2(0.0f, 0.0f);
    const uint32_t hi = static_cast<uint32_t>(key >> 32);
    const uint32_t u = (hi & 0x80000000u) ? (hi & 0x7FFFFFFFu) : ~hi;
    return make_float2(__uint_as_float(u), 1.0f);
}

// project_texture.hlsl:75-99: share of the PCF neighbourhood whose shadow-map
// depth plus the view's depth bias is not in front of the sample.  kR >= 0 is
// the radius fixed at compile time: every tap is loaded (outside taps from a
// clamped address) before any is tested.  kR < 0 takes the radius r at run time.
template <int kR>
__device__ __forceinline__ float bake_shadow(const BakeView& v, float px, float py, float depth, int32_t r) {
    const int32_t cx = f2i_sat(floorf(px)), cy = f2i_sat(floorf(py));
    const int32_t W = static_cast<int32_t>(v.width), H = static_cast<int32_t>(v.height);
    const unsigned long long* sm = v.shadow;
    float visible = 0.0f, count = 0.0f;
    if constexpr (kR >= 0) {
        constexpr int kD = 2 * kR + 1;
        float2 s[kD * kD];
        bool in[kD * kD];
#pragma unroll
        for (int y = 0; y < kD; ++y)
#pragma unroll
            for (int x = 0; x < kD; ++x) {
                const int32_t qx = cx + x - kR, qy = cy + y - kR;
                in[y * kD + x] = qx >= 0 && qy >= 0 && qx < W && qy < H;
                const int32_t lx = qx < 0 ? 0 : (qx >= W ? W - 1 : qx), ly = qy < 0 ? 0 : (qy >= H ? H - 1 : qy);
                s[y * kD + x] = shadow_texel(sm[key_index(lx, ly, v.key_tx)]);
            }
#pragma unroll
        for (int k = 0; k < kD * kD; ++k)
            if (in[k] && s[k].y != 0.0f) {
                visible += depth <= s[k].x + v.depth_bias ? 1.0f : 0.0f;
                count += 1.0f;
            }
    } else {
        for (int32_t y = -r; y <= r; ++y)
            for (int32_t x = -r; x <= r; ++x) {
                const int32_t qx = cx + x, qy = cy + y;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                const float2 s = shadow_texel(sm[key_index(qx, qy, v.key_tx)]);
                if (s.y != 0.0f) {
                    visible += depth <= s.x + v.depth_bias ? 1.0f : 0.0f;
                    count += 1.0f;
                }
            }
    }
    return count > 0.0f ? visible / count : 0.0f;
}

// One view's sample at a texel (project_texture.hlsl:126-175): projection into
// the view, visibility (shadow map, mask), view cosine and the photo's colour;
// false when the view contributes nothing.
struct BakeSample {
    float vis, cosv, dl, cr, cg, cb, alpha;
};

template <int kR>
__device__ __forceinline__ bool bake_sample(const BakeView& v, float x, float y, float z, float nx, float ny, float nz,
                                            int32_t pcf, BakeSample& s) {
    const float* m = v.m;
    float c[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) c[r] = m[r * 4] * x + m[r * 4 + 1] * y + m[r * 4 + 2] * z + m[r * 4 + 3] * 1.0f;
    if (c[3] <= 1e-8f) return false;
    const float ndx = c[0] / c[3], ndy = c[1] / c[3], ndz = c[2] / c[3];
    if (ndx < -1.0f || ndy < -1.0f || ndx > 1.0f || ndy > 1.0f || ndz < -1.0f || ndz > 1.0f) return false;
    const float px = v.vp[0] + (ndx * 0.5f + 0.5f) * v.vp[2];
    const float py = v.vp[1] + (ndy * 0.5f + 0.5f) * v.vp[3];
    if (px < v.vp[0] || py < v.vp[1] || px >= v.vp[0] + v.vp[2] || py >= v.vp[1] + v.vp[3]) return false;
    // the view-angle test first: a texel facing away needs no shadow or mask reads
    const float dx = v.cam[0] - x, dy = v.cam[1] - y, dz = v.cam[2] - z;
    const float dl = sqrtf(dx * dx + dy * dy + dz * dz);
    const float vx = dx / dl, vy = dy / dl, vz = dz / dl;
    const float cosv = nx * vx + ny * vy + nz * vz;
    if (cosv < v.min_cos) return false;
    // the photo taps (always inside the image) are read with the shadow and
    // mask taps, before the visibility that may discard them is known
    float col[4];
    bake_photo_all(v, px, py, col);
    float vis = bake_shadow<kR>(v, px, py, ndz, pcf);
    if (v.mask) vis *= clampf(bake_mask(v, px, py), 0.0f, 1.0f);
    if (vis <= 0.0f) return false;
    s.vis = vis;
    s.cosv = cosv;
    s.dl = dl;
    s.cr = col[0];
    s.cg = col[1];
    s.cb = col[2];
    s.alpha = col[3];
    return true;
}

// Texel position and unit normal (project_texture.hlsl:126-131).
__device__ __forceinline__ void texel_frame(const float* __restrict__ apos, const float* __restrict__ anrm, uint32_t p,
                                            float pos[3], float nrm[3]) {
    pos[0] = apos[p * 3];
    pos[1] = apos[p * 3 + 1];
    pos[2] = apos[p * 3 + 2];
    const float n0 = anrm[p * 3], n1 = anrm[p * 3 + 1], n2 = anrm[p * 3 + 2];
    const float nl = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
    nrm[0] = n0 / nl;
    nrm[1] = n1 / nl;
    nrm[2] = n2 / nl;
}

template <int kR>
__global__ __launch_bounds__(kBlock) void k_bake_project(const float* __restrict__ araster,
                                                         const float* __restrict__ apos,
                                                         const float* __restrict__ anrm, uint32_t P,
                                                         const BakeView* __restrict__ views, uint32_t nviews,
                                                         uint32_t blend, int32_t pcf, float* __restrict__ acc3,
                                                         float* __restrict__ accw, float* __restrict__ best,
                                                         uint32_t* __restrict__ src) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P || decode_id(araster[static_cast<size_t>(p) * 4 + 3]) == 0u) return;  // project_texture.hlsl:121-125
    float pos[3], nrm[3];
    texel_frame(apos, anrm, p, pos, nrm);
    float a0 = acc3[p * 3], a1 = acc3[p * 3 + 1], a2 = acc3[p * 3 + 2];
    float aw = accw[p], bst = best[p];
    uint32_t sv = src[p];
    for (uint32_t k = 0; k < nviews; ++k) {
        const BakeView& v = views[k];
        BakeSample s;
        if (!bake_sample<kR>(v, pos[0], pos[1], pos[2], nrm[0], nrm[1], nrm[2], pcf, s)) continue;
        const float vis = s.vis, cosv = s.cosv, dl = s.dl, cr = s.cr, cg = s.cg, cb = s.cb, alpha = s.alpha;
        if (blend == 2) {  // project_texture.hlsl:176-206, online softmax
            const float dist = maxf_std(dl, 1e-6f);
            const float footprint = cosv / (dist * dist);
            const float e = clampf(vis * alpha * footprint * v.weight, 0.0f, 64.0f);
            const float prev = bst;
            const float cur = maxf_std(prev, e);
            if (cur > prev) {
                const float rs = expf(prev - cur);
                a0 *= rs;
                a1 *= rs;
                a2 *= rs;
                aw *= rs;
                bst = cur;
            }
            if (e >= prev) sv = v.index;
            const float w = expf(e - cur);
            if (w <= 0.0f) continue;
            a0 += cr * w;
            a1 += cg * w;
            a2 += cb * w;
            aw += w;
            continue;
        }
        const float conf = vis * alpha * cosv * v.weight;  // project_texture.hlsl:208-246
        if (conf <= 0.0f) continue;
        if (conf > bst) {
            bst = conf;
            sv = v.index;
        }
        if (blend == 0) {
            if (conf >= aw) {
                a0 = cr;
                a1 = cg;
                a2 = cb;
                aw = conf;
            }
        } else {
            a0 += cr * conf;
            a1 += cg * conf;
            a2 += cb * conf;
            aw += conf;
        }
    }
    acc3[p * 3] = a0;
    acc3[p * 3 + 1] = a1;
    acc3[p * 3 + 2] = a2;
    accw[p] = aw;
    best[p] = bst;
    src[p] = sv;
}

// Warm-start features: for each covered texel the kFeatureViews samples of
// highest weighted-average confidence vis * alpha * cos * weight, in
// descending order (ties keep view order), each slot [R, G, B, confidence,
// cos]; unused slots are zero.  feats [P, 40] carries the list between launches.
constexpr int kFeatureViews = 8;

template <int kR>
__global__ __launch_bounds__(kBlock) void k_bake_features(const float* __restrict__ araster,
                                                          const float* __restrict__ apos,
                                                          const float* __restrict__ anrm, uint32_t P,
                                                          const BakeView* __restrict__ views, uint32_t nviews,
                                                          int32_t pcf, float* __restrict__ feats) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P || decode_id(araster[static_cast<size_t>(p) * 4 + 3]) == 0u) return;
    float pos[3], nrm[3];
    texel_frame(apos, anrm, p, pos, nrm);
    float f[kFeatureViews * 5];
    float4* row = reinterpret_cast<float4*>(feats + static_cast<size_t>(p) * kFeatureViews * 5);
#pragma unroll
    for (int i = 0; i < kFeatureViews * 5 / 4; ++i) {
        const float4 q = row[i];
        f[4 * i] = q.x;
        f[4 * i + 1] = q.y;
        f[4 * i + 2] = q.z;
        f[4 * i + 3] = q.w;
    }
    for (uint32_t k = 0; k < nviews; ++k) {
        const BakeView& v = views[k];
        BakeSample s;
        if (!bake_sample<kR>(v, pos[0], pos[1], pos[2], nrm[0], nrm[1], nrm[2], pcf, s)) continue;
        const float conf = s.vis * s.alpha * s.cosv * v.weight;  // project_texture.hlsl:208
        if (conf <= 0.0f) continue;
        // insertion into the descending list: the carried entry swaps with every smaller slot
        float c[5] = {s.cr, s.cg, s.cb, conf, s.cosv};
#pragma unroll
        for (int i = 0; i < kFeatureViews; ++i) {
            if (c[3] > f[5 * i + 3]) {
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    const float t = f[5 * i + e];
                    f[5 * i + e] = c[e];
                    c[e] = t;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kFeatureViews * 5 / 4; ++i)
        row[i] = make_float4(f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
}

// src/texture_baker.cpp:193-204 and :339-357: outputs from the blend state.
__global__ void k_bake_finish(uint32_t P, uint32_t blend, const float* __restrict__ araster,
                              const float* __restrict__ acc3, const float* __restrict__ accw, float4* __restrict__ color,
                              float* __restrict__ confidence, float* __restrict__ valid, float* __restrict__ coverage,
                              uint32_t* __restrict__ face_ids) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const float id = rintf(araster[static_cast<size_t>(p) * 4 + 3]);
    const float cov = id != 0.0f ? 1.0f : 0.0f;
    coverage[p] = cov;
    if (face_ids) face_ids[p] = cov > 0.5f ? static_cast<uint32_t>(id) - 1u : 0xFFFFFFFFu;
    const float w = accw[p];
    confidence[p] = w;
    const float ok = w > 0.0f ? 1.0f : 0.0f;
    valid[p] = ok;
    const float d = blend != 0 && w > 0.0f ? w : 1.0f;
    color[p] = make_float4(acc3[p * 3] / d, acc3[p * 3 + 1] / d, acc3[p * 3 + 2] / d, ok);
}

__global__ void k_bake_init(uint32_t P, float* __restrict__ acc3, float* __restrict__ accw, float* __restrict__ best,
                            uint32_t* __restrict__ src) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    acc3[p * 3] = acc3[p * 3 + 1] = acc3[p * 3 + 2] = 0.0f;
    accw[p] = 0.0f;
    best[p] = 0.0f;
    src[p] = 0xFFFFFFFFu;
}

uint32_t grid_of(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t atlas_clip_positions(const float* uv, uint32_t V, float* out4, hipStream_t st) {
    const uint32_t g = grid_of(V);
    k_atlas_clip<<<g < 65536 ? g : 65536, kBlock, 0, st>>>(uv, V, reinterpret_cast<float4*>(out4));
    return hipGetLastError();
}

hipError_t bake_init(uint32_t P, float* acc3, float* accw, float* best, uint32_t* src, hipStream_t st) {
    k_bake_init<<<grid_of(P), kBlock, 0, st>>>(P, acc3, accw, best, src);
    return hipGetLastError();
}

hipError_t bake_project(const float* araster, const float* apos, const float* anrm, uint32_t P, const BakeView* views,
                        uint32_t nviews, uint32_t blend, uint32_t pcf, float* acc3, float* accw, float* best,
                        uint32_t* src, hipStream_t st) {
    auto go = [&](auto kernel) {
        kernel<<<grid_of(P), kBlock, 0, st>>>(araster, apos, anrm, P, views, nviews, blend, static_cast<int32_t>(pcf),
                                              acc3, accw, best, src);
    };
    if (pcf == 0)
        go(k_bake_project<0>);
    else if (pcf == 1)
        go(k_bake_project<1>);
    else if (pcf == 2)
        go(k_bake_project<2>);
    else
        go(k_bake_project<-1>);
    return hipGetLastError();
}

hipError_t bake_features(const float* araster, const float* apos, const float* anrm, uint32_t P,
                         const BakeView* views, uint32_t nviews, uint32_t pcf, float* feats, hipStream_t st) {
    auto go = [&](auto kernel) {
        kernel<<<grid_of(P), kBlock, 0, st>>>(araster, apos, anrm, P, views, nviews, static_cast<int32_t>(pcf), feats);
    };
    if (pcf == 0)
        go(k_bake_features<0>);
    else if (pcf == 1)
        go(k_bake_features<1>);
    else if (pcf == 2)
        go(k_bake_features<2>);
    else
        go(k_bake_features<-1>);
    return hipGetLastError();
}

hipError_t bake_finish(uint32_t P, uint32_t blend, const float* araster, const float* acc3, const float* accw,
                       float* color4, float* confidence, float* valid, float* coverage, uint32_t* face_ids,
                       hipStream_t st) {
    k_bake_finish<<<grid_of(P), kBlock, 0, st>>>(P, blend, araster, acc3, accw, reinterpret_cast<float4*>(color4),
                                                 confidence, valid, coverage, face_ids);
    return hipGetLastError();
}

}  // namespace pdr
