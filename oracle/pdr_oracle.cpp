// pdr_oracle — CPU restatement of the reference raster/interp/texture/refine path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  The product
// (photara_drender_amd) never links or calls this file.
//
// Build with -ffp-contract=off (no FMA contraction) so that every float
// expression rounds exactly as written, in the reference's operation order.
// Out-of-range float->int conversions saturate (NaN -> 0); the reference leaves
// them undefined.  Triangle ids decode as round-half-even of the float id, with
// negative/NaN ids treated as uncovered (the reference's (uint)round() of a
// negative value is undefined).
#include "pdr_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

int fail(int code, const char* msg) {
    g_error = msg;
    return code;
}

constexpr uint32_t TILE = 16;  // src/rasterizer.cpp:19

inline int32_t f2i_sat(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
    if (f <= -2147483648.0f) return std::numeric_limits<int32_t>::min();
    return static_cast<int32_t>(f);
}

inline int32_t d2i_sat(double f) {
    if (f != f) return 0;
    if (f >= 2147483648.0) return std::numeric_limits<int32_t>::max();
    if (f <= -2147483648.0) return std::numeric_limits<int32_t>::min();
    return static_cast<int32_t>(f);
}

// (uint)round(x) of a stored triangle id (HLSL round = half-to-even).
inline uint32_t decode_id(float f) {
    const float r = std::nearbyint(f);
    if (!(r > 0.0f)) return 0;
    if (r >= 4294967296.0f) return 0xFFFFFFFFu;
    return static_cast<uint32_t>(r);
}

inline float cross2(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

inline float clampf(float x, float lo, float hi) { return std::min(std::max(x, lo), hi); }

struct Viewport {
    float x, y, w, h;
};

Viewport resolve_viewport(uint32_t W, uint32_t H, const float* vp) {
    if (vp) return {vp[0], vp[1], vp[2], vp[3]};
    return {0.0f, 0.0f, static_cast<float>(W), static_cast<float>(H)};
}

// src/rasterizer.cpp:101-116 pixel bounds of a binned triangle (inclusive).
struct PixelRect {
    int32_t x0, x1, y0, y1;
};

bool bin_rect(const float* pos, const uint32_t* tri, uint32_t W, uint32_t H, const Viewport& vp,
              PixelRect* out) {
    float min_x = std::numeric_limits<float>::infinity();
    float min_y = std::numeric_limits<float>::infinity();
    float max_x = -std::numeric_limits<float>::infinity();
    float max_y = -std::numeric_limits<float>::infinity();
    for (int c = 0; c < 3; ++c) {
        const float* p = pos + static_cast<size_t>(tri[c]) * 4;
        const float w = p[3];
        if (!(w > 1e-8f) || !std::isfinite(w)) return false;
        const float x = p[0] / w;
        const float y = p[1] / w;
        if (!std::isfinite(x) || !std::isfinite(y)) return false;
        min_x = std::min(min_x, x);
        min_y = std::min(min_y, y);
        max_x = std::max(max_x, x);
        max_y = std::max(max_y, y);
    }
    if (max_x < -1.0f || min_x > 1.0f || max_y < -1.0f || min_y > 1.0f) return false;
    const int64_t wm = static_cast<int64_t>(W) - 1, hm = static_cast<int64_t>(H) - 1;
    auto clampi = [](int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
    out->x0 = static_cast<int32_t>(clampi(f2i_sat(std::floor((min_x * 0.5f + 0.5f) * vp.w + vp.x)), 0, wm));
    out->x1 = static_cast<int32_t>(
        clampi(static_cast<int64_t>(f2i_sat(std::ceil((max_x * 0.5f + 0.5f) * vp.w + vp.x))) - 1, 0, wm));
    out->y0 = static_cast<int32_t>(clampi(f2i_sat(std::floor((min_y * 0.5f + 0.5f) * vp.h + vp.y)), 0, hm));
    out->y1 = static_cast<int32_t>(
        clampi(static_cast<int64_t>(f2i_sat(std::ceil((max_y * 0.5f + 0.5f) * vp.h + vp.y))) - 1, 0, hm));
    return true;
}

struct Bins {
    uint32_t tiles_x = 0;
    std::vector<uint32_t> offsets;
    std::vector<uint32_t> list;
};

int build_bins(const float* pos, const uint32_t* idx, uint32_t T, uint32_t W, uint32_t H, const Viewport& vp,
               Bins* bins) {
    bins->tiles_x = (W + TILE - 1) / TILE;
    const uint32_t tiles_y = (H + TILE - 1) / TILE;
    const size_t ntiles = static_cast<size_t>(bins->tiles_x) * tiles_y;
    std::vector<uint32_t> count(ntiles, 0);
    std::vector<PixelRect> rects(T);
    std::vector<uint8_t> valid(T, 0);
    for (uint32_t t = 0; t < T; ++t) {
        valid[t] = bin_rect(pos, idx + static_cast<size_t>(t) * 3, W, H, vp, &rects[t]) ? 1 : 0;
        if (!valid[t]) continue;
        const PixelRect& r = rects[t];
        for (uint32_t ty = r.y0 / TILE; ty <= static_cast<uint32_t>(r.y1) / TILE; ++ty)
            for (uint32_t tx = r.x0 / TILE; tx <= static_cast<uint32_t>(r.x1) / TILE; ++tx)
                ++count[static_cast<size_t>(ty) * bins->tiles_x + tx];
    }
    bins->offsets.assign(ntiles + 1, 0);
    uint64_t total = 0;
    for (size_t i = 0; i < ntiles; ++i) {
        total += count[i];
        if (total > 0xFFFFFFFFull) return fail(3, "Tile triangle list exceeds backend limits");
        bins->offsets[i + 1] = static_cast<uint32_t>(total);
    }
    bins->list.resize(total);
    std::vector<uint32_t> cursor(bins->offsets.begin(), bins->offsets.end() - 1);
    for (uint32_t t = 0; t < T; ++t) {  // triangles in ascending order, as the reference
        if (!valid[t]) continue;
        const PixelRect& r = rects[t];
        for (uint32_t ty = r.y0 / TILE; ty <= static_cast<uint32_t>(r.y1) / TILE; ++ty)
            for (uint32_t tx = r.x0 / TILE; tx <= static_cast<uint32_t>(r.x1) / TILE; ++tx)
                bins->list[cursor[static_cast<size_t>(ty) * bins->tiles_x + tx]++] = t;
    }
    return 0;
}

int address_coordinate(int32_t c, uint32_t size, uint32_t mode) {
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

inline int32_t add1(int32_t c) { return static_cast<int32_t>(static_cast<uint32_t>(c) + 1u); }

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_build_tile_bins(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T, uint32_t W, uint32_t H,
                        const float* viewport4, uint32_t* offsets, uint32_t* list) {
    (void)V;
    Bins b;
    const int rc = build_bins(pos, idx, T, W, H, resolve_viewport(W, H, viewport4), &b);
    if (rc) return rc;
    std::memcpy(offsets, b.offsets.data(), b.offsets.size() * sizeof(uint32_t));
    if (list) std::memcpy(list, b.list.data(), b.list.size() * sizeof(uint32_t));
    return 0;
}

// shaders/rasterize_forward.hlsl:45-175, one loop iteration per pixel invocation.
int orc_rasterize_forward(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T, uint32_t W, uint32_t H,
                          uint32_t cull_mode, int want_deriv, const float* viewport4, float* raster, float* deriv,
                          float* viewport_out4) {
    const Viewport vp = resolve_viewport(W, H, viewport4);
    if (viewport_out4) {
        viewport_out4[0] = vp.x;
        viewport_out4[1] = vp.y;
        viewport_out4[2] = vp.w;
        viewport_out4[3] = vp.h;
    }
    Bins bins;
    const int rc = build_bins(pos, idx, T, W, H, vp, &bins);
    if (rc) return rc;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t py = 0; py < static_cast<int64_t>(H); ++py) {
        for (uint32_t px = 0; px < W; ++px) {
            const size_t pixel = static_cast<size_t>(py) * W + px;
            const float sx = (static_cast<float>(px) - vp.x + 0.5f) * 2.0f / vp.w - 1.0f;
            const float sy = (static_cast<float>(py) - vp.y + 0.5f) * 2.0f / vp.h - 1.0f;
            float best_depth = 3.402823466e+38f;
            uint32_t best_tri = 0;
            float best_b0 = 0.0f, best_b1 = 0.0f;
            float bd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const size_t tile = static_cast<size_t>(py / TILE) * bins.tiles_x + px / TILE;
            for (uint32_t e = bins.offsets[tile]; e < bins.offsets[tile + 1]; ++e) {
                const uint32_t t = bins.list[e];
                const uint32_t* tri = idx + static_cast<size_t>(t) * 3;
                if (tri[0] >= V || tri[1] >= V || tri[2] >= V) continue;
                const float* p0 = pos + static_cast<size_t>(tri[0]) * 4;
                const float* p1 = pos + static_cast<size_t>(tri[1]) * 4;
                const float* p2 = pos + static_cast<size_t>(tri[2]) * 4;
                if (p0[3] <= 1e-8f || p1[3] <= 1e-8f || p2[3] <= 1e-8f) continue;
                const float s0x = p0[0] / p0[3], s0y = p0[1] / p0[3];
                const float s1x = p1[0] / p1[3], s1y = p1[1] / p1[3];
                const float s2x = p2[0] / p2[3], s2y = p2[1] / p2[3];
                const float screen_area = cross2(s1x - s0x, s1y - s0y, s2x - s0x, s2y - s0y);
                if (std::fabs(screen_area) <= 1e-12f) continue;
                if ((cull_mode == 1 && screen_area <= 0.0f) || (cull_mode == 2 && screen_area >= 0.0f)) continue;
                const float e0 = cross2(s1x - sx, s1y - sy, s2x - sx, s2y - sy);
                const float e1 = cross2(s2x - sx, s2y - sy, s0x - sx, s0y - sy);
                const float e2 = cross2(s0x - sx, s0y - sy, s1x - sx, s1y - sy);
                const float sgn = screen_area > 0.0f ? 1.0f : -1.0f;
                if (e0 * sgn < -1e-7f || e1 * sgn < -1e-7f || e2 * sgn < -1e-7f) continue;
                const float q0x = p0[0] - sx * p0[3], q0y = p0[1] - sy * p0[3];
                const float q1x = p1[0] - sx * p1[3], q1y = p1[1] - sy * p1[3];
                const float q2x = p2[0] - sx * p2[3], q2y = p2[1] - sy * p2[3];
                const float a0 = cross2(q1x, q1y, q2x, q2y);
                const float a1 = cross2(q2x, q2y, q0x, q0y);
                const float a2 = cross2(q0x, q0y, q1x, q1y);
                const float area = a0 + a1 + a2;
                if (std::fabs(area) <= 1e-12f) continue;
                const float inv = 1.0f / area;
                const float b0 = clampf(a0 * inv, 0.0f, 1.0f);
                const float b1 = clampf(a1 * inv, 0.0f, 1.0f);
                const float iz = p0[2] * a0 + p1[2] * a1 + p2[2] * a2;
                const float iw = p0[3] * a0 + p1[3] * a1 + p2[3] * a2;
                if (std::fabs(iw) <= 1e-12f) continue;
                const float depth = clampf(iz / iw, -1.0f, 1.0f);
                if (depth > best_depth || (depth == best_depth && best_tri != 0 && t + 1 >= best_tri)) continue;
                best_depth = depth;
                best_tri = t + 1;
                best_b0 = b0;
                best_b1 = b1;
                if (want_deriv) {
                    const float da0_dx = -p1[3] * p2[1] + p2[3] * p1[1];
                    const float da0_dy = -p1[0] * p2[3] + p1[3] * p2[0];
                    const float da1_dx = -p2[3] * p0[1] + p2[1] * p0[3];
                    const float da1_dy = -p2[0] * p0[3] + p2[3] * p0[0];
                    const float da2_dx = -p0[3] * p1[1] + p0[1] * p1[3];
                    const float da2_dy = -p0[0] * p1[3] + p0[3] * p1[0];
                    const float dA_dx = da0_dx + da1_dx + da2_dx;
                    const float dA_dy = da0_dy + da1_dy + da2_dy;
                    bd[0] = (da0_dx - b0 * dA_dx) * inv * 2.0f / vp.w;
                    bd[2] = (da1_dx - b1 * dA_dx) * inv * 2.0f / vp.w;
                    bd[1] = (da0_dy - b0 * dA_dy) * inv * 2.0f / vp.h;
                    bd[3] = (da1_dy - b1 * dA_dy) * inv * 2.0f / vp.h;
                }
            }
            float* r = raster + pixel * 4;
            if (best_tri == 0) {
                r[0] = r[1] = r[2] = r[3] = 0.0f;
                if (deriv) deriv[pixel * 4 + 0] = deriv[pixel * 4 + 1] = deriv[pixel * 4 + 2] = deriv[pixel * 4 + 3] = 0.0f;
            } else {
                r[0] = best_b0;
                r[1] = best_b1;
                r[2] = best_depth;
                r[3] = static_cast<float>(best_tri);
                if (deriv) {
                    // with derivatives disabled the shader still stores its zero-initialised value
                    for (int k = 0; k < 4; ++k) deriv[pixel * 4 + k] = want_deriv ? bd[k] : 0.0f;
                }
            }
        }
    }
    return 0;
}

// shaders/rasterize_backward.hlsl:36-158 then reduce_vertex_gradients.hlsl:12-29
// over build_vertex_adjacency (src/rasterizer.cpp:145-162).
int orc_rasterize_backward(const float* pos, uint32_t V, const uint32_t* idx, uint32_t T, const float* raster,
                           uint32_t W, uint32_t H, const float* grad_raster, const float* viewport4,
                           float* grad_pos) {
    const Viewport vp = resolve_viewport(W, H, viewport4);
    std::vector<float> corner(static_cast<size_t>(T) * 3 * 4, 0.0f);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ti = 0; ti < static_cast<int64_t>(T); ++ti) {
        const uint32_t t = static_cast<uint32_t>(ti);
        const uint32_t* tri = idx + static_cast<size_t>(t) * 3;
        float g[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        float* out = corner.data() + static_cast<size_t>(t) * 12;
        if (tri[0] >= V || tri[1] >= V || tri[2] >= V) continue;
        const float* p[3] = {pos + static_cast<size_t>(tri[0]) * 4, pos + static_cast<size_t>(tri[1]) * 4,
                             pos + static_cast<size_t>(tri[2]) * 4};
        if (p[0][3] <= 1e-8f || p[1][3] <= 1e-8f || p[2][3] <= 1e-8f) continue;
        const float s0x = p[0][0] / p[0][3], s0y = p[0][1] / p[0][3];
        const float s1x = p[1][0] / p[1][3], s1y = p[1][1] / p[1][3];
        const float s2x = p[2][0] / p[2][3], s2y = p[2][1] / p[2][3];
        const float mnx = std::min(s0x, std::min(s1x, s2x)), mny = std::min(s0y, std::min(s1y, s2y));
        const float mxx = std::max(s0x, std::max(s1x, s2x)), mxy = std::max(s0y, std::max(s1y, s2y));
        auto clampi = [](int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
        const int64_t x0 = clampi(f2i_sat(std::floor((mnx * 0.5f + 0.5f) * vp.w + vp.x)), 0, W);
        const int64_t x1 = clampi(f2i_sat(std::ceil((mxx * 0.5f + 0.5f) * vp.w + vp.x)), 0, W);
        const int64_t y0 = clampi(f2i_sat(std::floor((mny * 0.5f + 0.5f) * vp.h + vp.y)), 0, H);
        const int64_t y1 = clampi(f2i_sat(std::ceil((mxy * 0.5f + 0.5f) * vp.h + vp.y)), 0, H);
        for (int64_t y = y0; y < y1; ++y) {
            for (int64_t x = x0; x < x1; ++x) {
                const size_t pix = static_cast<size_t>(y) * W + static_cast<size_t>(x);
                if (decode_id(raster[pix * 4 + 3]) != t + 1) continue;
                const float lgx = grad_raster[pix * 4 + 0], lgy = grad_raster[pix * 4 + 1];
                if (lgx == 0.0f && lgy == 0.0f) continue;
                const float sx = (static_cast<float>(x) - vp.x + 0.5f) * 2.0f / vp.w - 1.0f;
                const float sy = (static_cast<float>(y) - vp.y + 0.5f) * 2.0f / vp.h - 1.0f;
                const float q0x = p[0][0] - sx * p[0][3], q0y = p[0][1] - sy * p[0][3];
                const float q1x = p[1][0] - sx * p[1][3], q1y = p[1][1] - sy * p[1][3];
                const float q2x = p[2][0] - sx * p[2][3], q2y = p[2][1] - sy * p[2][3];
                const float area = cross2(q1x, q1y, q2x, q2y) + cross2(q2x, q2y, q0x, q0y) + cross2(q0x, q0y, q1x, q1y);
                if (std::fabs(area) <= 1e-12f) continue;
                const float eps = area >= 0.0f ? 1e-6f : -1e-6f;
                const float inv = 1.0f / (area + eps);
                const float a0 = cross2(q1x, q1y, q2x, q2y);
                const float a1 = cross2(q2x, q2y, q0x, q0y);
                const float b0 = a0 * inv, b1 = a1 * inv;
                const float gb0 = lgx * inv, gb1 = lgy * inv;
                const float gc = gb0 * b0 + gb1 * b1;
                const float gp0x = gc * (q2y - q1y) - gb1 * q2y;
                const float gp1x = gc * (q0y - q2y) + gb0 * q2y;
                const float gp2x = gc * (q1y - q0y) - gb0 * q1y + gb1 * q0y;
                const float gp0y = gc * (q1x - q2x) + gb1 * q2x;
                const float gp1y = gc * (q2x - q0x) - gb0 * q2x;
                const float gp2y = gc * (q0x - q1x) + gb0 * q1x - gb1 * q0x;
                g[0][0] += gp0x; g[0][1] += gp0y; g[0][2] += 0.0f; g[0][3] += -sx * gp0x - sy * gp0y;
                g[1][0] += gp1x; g[1][1] += gp1y; g[1][2] += 0.0f; g[1][3] += -sx * gp1x - sy * gp1y;
                g[2][0] += gp2x; g[2][1] += gp2y; g[2][2] += 0.0f; g[2][3] += -sx * gp2x - sy * gp2y;
            }
        }
        std::memcpy(out, g, sizeof(g));
    }
    // build_vertex_adjacency: counting sort, corners ascending per vertex.
    std::vector<uint32_t> offsets(static_cast<size_t>(V) + 1, 0);
    const size_t ncorner = static_cast<size_t>(T) * 3;
    for (size_t c = 0; c < ncorner; ++c) ++offsets[idx[c] + 1];
    for (uint32_t v = 1; v <= V; ++v) offsets[v] += offsets[v - 1];
    std::vector<uint32_t> corners(ncorner);
    std::vector<uint32_t> cursor(offsets);
    for (size_t c = 0; c < ncorner; ++c) corners[cursor[idx[c]]++] = static_cast<uint32_t>(c);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < static_cast<int64_t>(V); ++v) {
        float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t a = offsets[v]; a < offsets[v + 1]; ++a) {
            const float* cg = corner.data() + static_cast<size_t>(corners[a]) * 4;
            s[0] += cg[0]; s[1] += cg[1]; s[2] += cg[2]; s[3] += cg[3];
        }
        std::memcpy(grad_pos + static_cast<size_t>(v) * 4, s, sizeof(s));
    }
    return 0;
}

// shaders/interpolate_forward.hlsl:15-54
int orc_interpolate_forward(const float* attr, uint32_t V, uint32_t A, const uint32_t* idx, uint32_t T,
                            const float* raster, uint32_t P, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < static_cast<int64_t>(P); ++p) {
        const uint32_t id = decode_id(raster[p * 4 + 3]);
        float* o = out + static_cast<size_t>(p) * A;
        const uint32_t* tri = idx + static_cast<size_t>(id ? id - 1 : 0) * 3;
        if (id == 0 || id > T || tri[0] >= V || tri[1] >= V || tri[2] >= V) {
            for (uint32_t a = 0; a < A; ++a) o[a] = 0.0f;
            continue;
        }
        const float u = raster[p * 4 + 0], v = raster[p * 4 + 1];
        const float w = 1.0f - u - v;
        for (uint32_t a = 0; a < A; ++a) {
            const float a0 = attr[static_cast<size_t>(tri[0]) * A + a];
            const float a1 = attr[static_cast<size_t>(tri[1]) * A + a];
            const float a2 = attr[static_cast<size_t>(tri[2]) * A + a];
            o[a] = u * a0 + v * a1 + w * a2;
        }
    }
    return 0;
}

// shaders/interpolate_grad_attributes.hlsl:15-58 + interpolate_grad_raster.hlsl:16-54
// shaders/interpolate_grad_attributes.hlsl:15-58: one invocation per output
// o = vertex * A + attribute, looping over every pixel; outputs [o_begin, o_end)
// only (the CPU baseline times a slice and extrapolates).
void orc_interpolate_grad_attributes_range(uint32_t V, uint32_t A, const uint32_t* idx, uint32_t T,
                                           const float* raster, uint32_t P, const float* grad, uint64_t o_begin,
                                           uint64_t o_end, float* out) {
    const uint64_t end = std::min<uint64_t>(o_end, static_cast<uint64_t>(V) * A);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t o = static_cast<int64_t>(o_begin); o < static_cast<int64_t>(end); ++o) {
        const uint32_t v = static_cast<uint32_t>(o / A), a = static_cast<uint32_t>(o % A);
        float g = 0.0f;
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t id = decode_id(raster[static_cast<size_t>(p) * 4 + 3]);
            if (id == 0 || id > T) continue;
            const uint32_t* tri = idx + static_cast<size_t>(id - 1) * 3;
            const float u = raster[static_cast<size_t>(p) * 4], w = raster[static_cast<size_t>(p) * 4 + 1];
            float weight = 0.0f;
            if (tri[0] == v) weight += u;
            if (tri[1] == v) weight += w;
            if (tri[2] == v) weight += 1.0f - u - w;
            g += weight * grad[static_cast<size_t>(p) * A + a];
        }
        out[o - static_cast<int64_t>(o_begin)] = g;
    }
}

int orc_interpolate_backward(const float* attr, uint32_t V, uint32_t A, const uint32_t* idx, uint32_t T,
                             const float* raster, uint32_t P, const float* grad, float* grad_attr,
                             float* grad_raster, int reference_mode) {
    if (reference_mode) {
        orc_interpolate_grad_attributes_range(V, A, idx, T, raster, P, grad, 0, static_cast<uint64_t>(V) * A,
                                              grad_attr);
    } else {
        // pixel lists per vertex, ascending pixel order (each pixel once per distinct vertex)
        std::vector<uint32_t> count(static_cast<size_t>(V) + 1, 0);
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t id = decode_id(raster[static_cast<size_t>(p) * 4 + 3]);
            if (id == 0 || id > T) continue;
            const uint32_t* tri = idx + static_cast<size_t>(id - 1) * 3;
            for (int c = 0; c < 3; ++c) {
                bool seen = false;
                for (int k = 0; k < c; ++k) seen |= tri[k] == tri[c];
                if (!seen && tri[c] < V) ++count[tri[c] + 1];
            }
        }
        for (uint32_t v = 1; v <= V; ++v) count[v] += count[v - 1];
        std::vector<uint32_t> plist(count[V]);
        std::vector<uint32_t> cursor(count);
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t id = decode_id(raster[static_cast<size_t>(p) * 4 + 3]);
            if (id == 0 || id > T) continue;
            const uint32_t* tri = idx + static_cast<size_t>(id - 1) * 3;
            for (int c = 0; c < 3; ++c) {
                bool seen = false;
                for (int k = 0; k < c; ++k) seen |= tri[k] == tri[c];
                if (!seen && tri[c] < V) plist[cursor[tri[c]]++] = p;
            }
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t v = 0; v < static_cast<int64_t>(V); ++v) {
            for (uint32_t a = 0; a < A; ++a) {
                float g = 0.0f;
                for (uint32_t e = count[v]; e < count[v + 1]; ++e) {
                    const uint32_t p = plist[e];
                    const uint32_t id = decode_id(raster[static_cast<size_t>(p) * 4 + 3]);
                    const uint32_t* tri = idx + static_cast<size_t>(id - 1) * 3;
                    const float u = raster[static_cast<size_t>(p) * 4], w = raster[static_cast<size_t>(p) * 4 + 1];
                    float weight = 0.0f;
                    if (tri[0] == static_cast<uint32_t>(v)) weight += u;
                    if (tri[1] == static_cast<uint32_t>(v)) weight += w;
                    if (tri[2] == static_cast<uint32_t>(v)) weight += 1.0f - u - w;
                    g += weight * grad[static_cast<size_t>(p) * A + a];
                }
                grad_attr[static_cast<size_t>(v) * A + a] = g;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < static_cast<int64_t>(P); ++p) {
        const uint32_t id = decode_id(raster[p * 4 + 3]);
        float gx = 0.0f, gy = 0.0f;
        if (id != 0 && id <= T) {
            const uint32_t* tri = idx + static_cast<size_t>(id - 1) * 3;
            if (tri[0] < V && tri[1] < V && tri[2] < V) {
                for (uint32_t a = 0; a < A; ++a) {
                    const float a0 = attr[static_cast<size_t>(tri[0]) * A + a];
                    const float a1 = attr[static_cast<size_t>(tri[1]) * A + a];
                    const float a2 = attr[static_cast<size_t>(tri[2]) * A + a];
                    const float og = grad[static_cast<size_t>(p) * A + a];
                    gx += og * (a0 - a2);
                    gy += og * (a1 - a2);
                }
            }
        }
        grad_raster[p * 4 + 0] = gx;
        grad_raster[p * 4 + 1] = gy;
        grad_raster[p * 4 + 2] = 0.0f;
        grad_raster[p * 4 + 3] = 0.0f;
    }
    return 0;
}

// shaders/texture_forward.hlsl:37-79
int orc_texture_forward(const float* tex, uint32_t tw, uint32_t th, uint32_t C, uint32_t mode, const float* uv,
                        const float* raster, uint32_t P, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < static_cast<int64_t>(P); ++p) {
        float* o = out + static_cast<size_t>(p) * C;
        if (decode_id(raster[p * 4 + 3]) == 0) {
            for (uint32_t c = 0; c < C; ++c) o[c] = 0.0f;
            continue;
        }
        const float tx = uv[p * 2 + 0] * static_cast<float>(tw) - 0.5f;
        const float ty = uv[p * 2 + 1] * static_cast<float>(th) - 0.5f;
        const float flx = std::floor(tx), fly = std::floor(ty);
        const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
        const float fx = tx - flx, fy = ty - fly;
        const int32_t x0 = address_coordinate(bx, tw, mode), x1 = address_coordinate(add1(bx), tw, mode);
        const int32_t y0 = address_coordinate(by, th, mode), y1 = address_coordinate(add1(by), th, mode);
        const size_t o00 = (static_cast<size_t>(y0) * tw + x0) * C, o10 = (static_cast<size_t>(y0) * tw + x1) * C;
        const size_t o01 = (static_cast<size_t>(y1) * tw + x0) * C, o11 = (static_cast<size_t>(y1) * tw + x1) * C;
        const float w0 = (1.0f - fx) * (1.0f - fy), w1 = fx * (1.0f - fy);
        const float w2 = (1.0f - fx) * fy, w3 = fx * fy;
        for (uint32_t c = 0; c < C; ++c)
            o[c] = tex[o00 + c] * w0 + tex[o10 + c] * w1 + tex[o01 + c] * w2 + tex[o11 + c] * w3;
    }
    return 0;
}

int orc_texture_backward(const float* tex, uint32_t tw, uint32_t th, uint32_t C, uint32_t mode, const float* uv,
                         const float* raster, uint32_t P, const float* grad, int compute_uv, float* grad_tex,
                         float* grad_uv) {
    // build_texture_sample_adjacency, src/rasterizer.cpp:187-258 (double-precision fractions)
    const size_t texels = static_cast<size_t>(tw) * th;
    std::vector<uint32_t> toff(texels + 1, 0);
    std::vector<float> sw(static_cast<size_t>(P) * 4, 0.0f);
    std::vector<uint32_t> stx(static_cast<size_t>(P) * 4, 0);
    size_t nvalid = 0;
    for (size_t p = 0; p < P; ++p) {
        if (raster[p * 4 + 3] == 0.0f) continue;
        const float u = uv[p * 2], v = uv[p * 2 + 1];
        if (!std::isfinite(u) || !std::isfinite(v)) return fail(1, "uv contains a non-finite value in a covered pixel");
        const double tx = static_cast<double>(u) * tw - 0.5;
        const double ty = static_cast<double>(v) * th - 0.5;
        if (std::fabs(tx) > 1.0e9 || std::fabs(ty) > 1.0e9) return fail(2, "uv is too large for portable texture addressing");
        const int32_t bx = d2i_sat(std::floor(tx)), by = d2i_sat(std::floor(ty));
        const float fx = static_cast<float>(tx - std::floor(tx));
        const float fy = static_cast<float>(ty - std::floor(ty));
        const int32_t xs[2] = {address_coordinate(bx, tw, mode), address_coordinate(bx + 1, tw, mode)};
        const int32_t ys[2] = {address_coordinate(by, th, mode), address_coordinate(by + 1, th, mode)};
        const float w[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
        const uint32_t tx4[4] = {ys[0] * tw + xs[0], ys[0] * tw + xs[1], ys[1] * tw + xs[0], ys[1] * tw + xs[1]};
        for (int k = 0; k < 4; ++k) {
            sw[p * 4 + k] = w[k];
            stx[p * 4 + k] = tx4[k];
            ++toff[tx4[k] + 1];
            ++nvalid;
        }
    }
    for (size_t t = 1; t <= texels; ++t) toff[t] += toff[t - 1];
    std::vector<uint32_t> entries(nvalid);
    std::vector<uint32_t> cur(toff);
    for (size_t e = 0; e < static_cast<size_t>(P) * 4; ++e) {
        if (raster[(e / 4) * 4 + 3] == 0.0f) continue;
        entries[cur[stx[e]]++] = static_cast<uint32_t>(e);
    }
    // shaders/texture_grad_texture.hlsl:15-38
#pragma omp parallel for schedule(static)
    for (int64_t o = 0; o < static_cast<int64_t>(texels * C); ++o) {
        const size_t t = static_cast<size_t>(o) / C, c = static_cast<size_t>(o) % C;
        float g = 0.0f;
        for (uint32_t a = toff[t]; a < toff[t + 1]; ++a) {
            const uint32_t e = entries[a];
            g += sw[e] * grad[static_cast<size_t>(e / 4) * C + c];
        }
        grad_tex[o] = g;
    }
    if (!compute_uv || !grad_uv) return 0;
    // shaders/texture_grad_uv.hlsl:38-80
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < static_cast<int64_t>(P); ++p) {
        if (decode_id(raster[p * 4 + 3]) == 0) {
            grad_uv[p * 2] = grad_uv[p * 2 + 1] = 0.0f;
            continue;
        }
        const float tx = uv[p * 2 + 0] * static_cast<float>(tw) - 0.5f;
        const float ty = uv[p * 2 + 1] * static_cast<float>(th) - 0.5f;
        const float flx = std::floor(tx), fly = std::floor(ty);
        const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
        const float fx = tx - flx, fy = ty - fly;
        const int32_t x0 = address_coordinate(bx, tw, mode), x1 = address_coordinate(add1(bx), tw, mode);
        const int32_t y0 = address_coordinate(by, th, mode), y1 = address_coordinate(add1(by), th, mode);
        float gx = 0.0f, gy = 0.0f;
        for (uint32_t c = 0; c < C; ++c) {
            const float c00 = tex[(static_cast<size_t>(y0) * tw + x0) * C + c];
            const float c10 = tex[(static_cast<size_t>(y0) * tw + x1) * C + c];
            const float c01 = tex[(static_cast<size_t>(y1) * tw + x0) * C + c];
            const float c11 = tex[(static_cast<size_t>(y1) * tw + x1) * C + c];
            const float gv = grad[static_cast<size_t>(p) * C + c];
            const float dx = (1.0f - fy) * (c10 - c00) + fy * (c11 - c01);
            const float dy = (1.0f - fx) * (c01 - c00) + fx * (c11 - c10);
            gx += gv * (dx * static_cast<float>(tw));
            gy += gv * (dy * static_cast<float>(th));
        }
        grad_uv[p * 2] = gx;
        grad_uv[p * 2 + 1] = gy;
    }
    return 0;
}

void orc_transform_positions(const float* pos3, uint32_t V, const float* m, float* out4) {
    for (size_t v = 0; v < V; ++v) {
        const float x = pos3[v * 3], y = pos3[v * 3 + 1], z = pos3[v * 3 + 2];
        for (int r = 0; r < 4; ++r) out4[v * 4 + r] = m[r * 4] * x + m[r * 4 + 1] * y + m[r * 4 + 2] * z + m[r * 4 + 3];
    }
}

namespace {
// clamp-mode bilinear layout of shaders/atlas_photometric_gradient.hlsl:98-117
inline void atlas_layout(float u, float v, uint32_t tw, uint32_t th, uint32_t texels[4], float w[4]) {
    const float tx = u * static_cast<float>(tw) - 0.5f, ty = v * static_cast<float>(th) - 0.5f;
    const float flx = std::floor(tx), fly = std::floor(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    const float fx = tx - flx, fy = ty - fly;
    const int32_t wm = static_cast<int32_t>(tw) - 1, hm = static_cast<int32_t>(th) - 1;
    auto cl = [](int32_t c, int32_t hi) { return static_cast<uint32_t>(c < 0 ? 0 : (c > hi ? hi : c)); };
    const uint32_t x0 = cl(bx, wm), x1 = cl(add1(bx), wm), y0 = cl(by, hm), y1 = cl(add1(by), hm);
    texels[0] = y0 * tw + x0;
    texels[1] = y0 * tw + x1;
    texels[2] = y1 * tw + x0;
    texels[3] = y1 * tw + x1;
    w[0] = (1.0f - fx) * (1.0f - fy);
    w[1] = fx * (1.0f - fy);
    w[2] = (1.0f - fx) * fy;
    w[3] = fx * fy;
}

// groupshared tree reduction of 64 values (stride 32..1), as the atlas shaders do.
float group_tree_sum(float* g) {
    for (int s = 32; s > 0; s >>= 1)
        for (int i = 0; i < s; ++i) g[i] += g[i + s];
    return g[0];
}
}  // namespace

void orc_photometric_gradient(const float* tex, uint32_t tw, uint32_t th, const float* uv, const float* target,
                              const float* mask, uint32_t P, float eps, float norm, float* grad, float* loss) {
    const uint32_t groups = (P + 63) / 64;
    for (uint32_t gi = 0; gi < groups; ++gi) {
        float gl[64];
        for (uint32_t li = 0; li < 64; ++li) {
            const uint32_t p = gi * 64 + li;
            float local = 0.0f;
            if (p < P && mask[p] > 0.0f) {
                const float m = mask[p];
                uint32_t tx[4];
                float w[4];
                atlas_layout(uv[static_cast<size_t>(p) * 2], uv[static_cast<size_t>(p) * 2 + 1], tw, th, tx, w);
                for (uint32_t ch = 0; ch < 3; ++ch) {
                    float rendered = 0.0f;
                    for (int k = 0; k < 4; ++k) rendered += tex[static_cast<size_t>(tx[k]) * 3 + ch] * w[k];
                    const float d = rendered - target[static_cast<size_t>(p) * 3 + ch];
                    const float r = std::sqrt(d * d + eps * eps);
                    const float og = d / r * m * norm;
                    local += r * m * norm;
                    for (int k = 0; k < 4; ++k) grad[static_cast<size_t>(tx[k]) * 3 + ch] += og * w[k];
                }
            }
            gl[li] = local;
        }
        const float s = group_tree_sum(gl);
        if (s != 0.0f) *loss += s;
    }
}

void orc_seam_gradient(const float* tex, uint32_t tw, uint32_t th, const float* pairs, uint32_t N, float eps,
                       float norm, float* grad, float* loss) {
    const uint32_t groups = (N + 63) / 64;
    for (uint32_t gi = 0; gi < groups; ++gi) {
        float gl[64];
        for (uint32_t li = 0; li < 64; ++li) {
            const uint32_t q = gi * 64 + li;
            float local = 0.0f;
            if (q < N) {
                const float* pr = pairs + static_cast<size_t>(q) * 4;
                uint32_t t0[4], t1[4];
                float w0[4], w1[4];
                atlas_layout(pr[0], pr[1], tw, th, t0, w0);
                atlas_layout(pr[2], pr[3], tw, th, t1, w1);
                for (uint32_t ch = 0; ch < 3; ++ch) {
                    float c0 = 0.0f, c1 = 0.0f;
                    for (int k = 0; k < 4; ++k) {
                        c0 += tex[static_cast<size_t>(t0[k]) * 3 + ch] * w0[k];
                        c1 += tex[static_cast<size_t>(t1[k]) * 3 + ch] * w1[k];
                    }
                    const float d = c0 - c1;
                    const float r = std::sqrt(d * d + eps * eps);
                    const float g = d / r * norm;
                    local += r * norm;
                    for (int k = 0; k < 4; ++k) {
                        grad[static_cast<size_t>(t0[k]) * 3 + ch] += g * w0[k];
                        grad[static_cast<size_t>(t1[k]) * 3 + ch] += -g * w1[k];
                    }
                }
            }
            gl[li] = local;
        }
        const float s = group_tree_sum(gl);
        if (s != 0.0f) *loss += s;
    }
}

void orc_adam_update(uint32_t n, float lr, float b1, float b2, float eps, float bc1, float bc2, float cmin,
                     float cmax, float* tex, float* grad, float* m, float* v) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
        const float g = grad[i];
        const float m1 = b1 * m[i] + (1.0f - b1) * g;
        const float m2 = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = m1;
        v[i] = m2;
        const float c1 = m1 / bc1, c2 = m2 / bc2;
        tex[i] = clampf(tex[i] - lr * c1 / (std::sqrt(c2) + eps), cmin, cmax);
        grad[i] = 0.0f;
    }
}

void orc_refine_schedule(uint32_t steps, uint32_t batch, uint32_t nviews, const double* vw, float lr, float min_lr,
                         float b1, float b2, uint32_t seam_steps, float* lr_out, float* bc1_out, float* bc2_out,
                         float* norm_out, uint32_t* view_idx_out, float* seam_bc1_out, float* seam_bc2_out) {
    const float pi = std::acos(-1.0F);
    for (uint32_t s = 0; s < steps; ++s) {
        double bw = 0.0;
        for (uint32_t b = 0; b < batch; ++b) {
            const size_t vi = (static_cast<size_t>(s) * batch + b) % nviews;
            bw += vw[vi];
            view_idx_out[static_cast<size_t>(s) * batch + b] = static_cast<uint32_t>(vi);
        }
        norm_out[s] = bw > 0.0 ? static_cast<float>(1.0 / (bw * 3.0)) : 0.0F;
        const float phase = static_cast<float>(s) / static_cast<float>(steps);
        lr_out[s] = min_lr + 0.5F * (lr - min_lr) * (1.0F + std::cos(pi * phase));
        bc1_out[s] = 1.0F - std::pow(b1, static_cast<float>(s + 1));
        bc2_out[s] = 1.0F - std::pow(b2, static_cast<float>(s + 1));
    }
    for (uint32_t k = 0; k < seam_steps; ++k) {
        seam_bc1_out[k] = 1.0F - std::pow(b1, static_cast<float>(k + 1));
        seam_bc2_out[k] = 1.0F - std::pow(b2, static_cast<float>(k + 1));
    }
}

void orc_refine_view_inputs(const float* image, uint32_t channels, const float* vis, const float* raster, uint32_t P,
                            float* target3, float* mask, double* valid_weight) {
    double wsum = 0.0;
    for (size_t p = 0; p < P; ++p) {
        target3[p * 3] = image[p * channels];
        target3[p * 3 + 1] = image[p * channels + 1];
        target3[p * 3 + 2] = image[p * channels + 2];
        mask[p] = 0.0f;
        if (raster[p * 4 + 3] != 0.0f) {
            const float ext = vis ? clampf(vis[p], 0.0f, 1.0f) : 1.0f;
            mask[p] = ext;
            wsum += ext;
        }
    }
    *valid_weight = wsum;
}

void orc_raster_fixed_f64(const double* pos, uint32_t V, const uint32_t* idx, uint32_t T, const float* raster_ids,
                          uint32_t W, uint32_t H, const double* vp4, double* out) {
    const double vx = vp4 ? vp4[0] : 0.0, vy = vp4 ? vp4[1] : 0.0;
    const double vw = vp4 ? vp4[2] : static_cast<double>(W), vh = vp4 ? vp4[3] : static_cast<double>(H);
    for (size_t py = 0; py < H; ++py) {
        for (size_t px = 0; px < W; ++px) {
            const size_t pix = py * W + px;
            const uint32_t id = decode_id(raster_ids[pix * 4 + 3]);
            double* o = out + pix * 4;
            o[0] = o[1] = o[2] = o[3] = 0.0;
            if (id == 0 || id > T) continue;
            const uint32_t* tri = idx + static_cast<size_t>(id - 1) * 3;
            if (tri[0] >= V || tri[1] >= V || tri[2] >= V) continue;
            const double* p0 = pos + static_cast<size_t>(tri[0]) * 4;
            const double* p1 = pos + static_cast<size_t>(tri[1]) * 4;
            const double* p2 = pos + static_cast<size_t>(tri[2]) * 4;
            const double sx = (static_cast<double>(px) - vx + 0.5) * 2.0 / vw - 1.0;
            const double sy = (static_cast<double>(py) - vy + 0.5) * 2.0 / vh - 1.0;
            const double q0x = p0[0] - sx * p0[3], q0y = p0[1] - sy * p0[3];
            const double q1x = p1[0] - sx * p1[3], q1y = p1[1] - sy * p1[3];
            const double q2x = p2[0] - sx * p2[3], q2y = p2[1] - sy * p2[3];
            const double a0 = q1x * q2y - q1y * q2x, a1 = q2x * q0y - q2y * q0x, a2 = q0x * q1y - q0y * q1x;
            const double inv = 1.0 / (a0 + a1 + a2);
            o[0] = std::min(std::max(a0 * inv, 0.0), 1.0);
            o[1] = std::min(std::max(a1 * inv, 0.0), 1.0);
            const double iz = p0[2] * a0 + p1[2] * a1 + p2[2] * a2, iw = p0[3] * a0 + p1[3] * a1 + p2[3] * a2;
            o[2] = std::min(std::max(iz / iw, -1.0), 1.0);
            o[3] = static_cast<double>(id);
        }
    }
}

void orc_interpolate_forward_f64(const double* attr, uint32_t V, uint32_t A, const uint32_t* idx, uint32_t T,
                                 const double* raster, uint32_t P, double* out) {
    for (size_t p = 0; p < P; ++p) {
        const double idf = raster[p * 4 + 3];
        const uint32_t id = idf > 0.5 ? static_cast<uint32_t>(idf + 0.5) : 0;
        const uint32_t* tri = idx + static_cast<size_t>(id ? id - 1 : 0) * 3;
        for (uint32_t a = 0; a < A; ++a) {
            if (id == 0 || id > T || tri[0] >= V || tri[1] >= V || tri[2] >= V) {
                out[p * A + a] = 0.0;
                continue;
            }
            const double u = raster[p * 4], v = raster[p * 4 + 1];
            out[p * A + a] = u * attr[tri[0] * A + a] + v * attr[tri[1] * A + a] + (1.0 - u - v) * attr[tri[2] * A + a];
        }
    }
}

void orc_texture_forward_f64(const double* tex, uint32_t tw, uint32_t th, uint32_t C, uint32_t mode, const double* uv,
                             const double* raster, uint32_t P, double* out) {
    for (size_t p = 0; p < P; ++p) {
        if (!(raster[p * 4 + 3] > 0.5)) {
            for (uint32_t c = 0; c < C; ++c) out[p * C + c] = 0.0;
            continue;
        }
        const double tx = uv[p * 2] * tw - 0.5, ty = uv[p * 2 + 1] * th - 0.5;
        const double flx = std::floor(tx), fly = std::floor(ty);
        const int32_t bx = d2i_sat(flx), by = d2i_sat(fly);
        const double fx = tx - flx, fy = ty - fly;
        const int32_t x0 = address_coordinate(bx, tw, mode), x1 = address_coordinate(bx + 1, tw, mode);
        const int32_t y0 = address_coordinate(by, th, mode), y1 = address_coordinate(by + 1, th, mode);
        for (uint32_t c = 0; c < C; ++c) {
            out[p * C + c] = tex[(static_cast<size_t>(y0) * tw + x0) * C + c] * (1 - fx) * (1 - fy) +
                             tex[(static_cast<size_t>(y0) * tw + x1) * C + c] * fx * (1 - fy) +
                             tex[(static_cast<size_t>(y1) * tw + x0) * C + c] * (1 - fx) * fy +
                             tex[(static_cast<size_t>(y1) * tw + x1) * C + c] * fx * fy;
        }
    }
}

namespace {

// shaders/project_texture.hlsl:36-54: bilinear sample with coordinates clamped
// to the image; lerp(a, b, t) = a + t * (b - a).
inline float bake_lerp(float a, float b, float t) { return a + t * (b - a); }

float bake_photo(const orc_bake_view& v, float px, float py, uint32_t ch) {
    const float tx = px - 0.5f, ty = py - 0.5f;
    const float flx = std::floor(tx), fly = std::floor(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    const float fx = tx - flx, fy = ty - fly;
    const int32_t wm = static_cast<int32_t>(v.width) - 1, hm = static_cast<int32_t>(v.height) - 1;
    auto at = [&](int32_t x, int32_t y) {
        x = std::min(std::max(x, 0), wm);
        y = std::min(std::max(y, 0), hm);
        return v.image[(static_cast<size_t>(y) * v.width + x) * v.channels + ch];
    };
    return bake_lerp(bake_lerp(at(bx, by), at(add1(bx), by), fx), bake_lerp(at(bx, add1(by)), at(add1(bx), add1(by)), fx),
                     fy);
}

// shaders/project_texture.hlsl:56-73
float bake_mask(const orc_bake_view& v, float px, float py) {
    const float tx = px - 0.5f, ty = py - 0.5f;
    const float flx = std::floor(tx), fly = std::floor(ty);
    const int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    const float fx = tx - flx, fy = ty - fly;
    const int32_t wm = static_cast<int32_t>(v.width) - 1, hm = static_cast<int32_t>(v.height) - 1;
    auto at = [&](int32_t x, int32_t y) {
        x = std::min(std::max(x, 0), wm);
        y = std::min(std::max(y, 0), hm);
        return v.mask[static_cast<size_t>(y) * v.width + x];
    };
    return bake_lerp(bake_lerp(at(bx, by), at(add1(bx), by), fx), bake_lerp(at(bx, add1(by)), at(add1(bx), add1(by)), fx),
                     fy);
}

// shaders/project_texture.hlsl:75-99: share of the PCF neighbourhood whose
// shadow-raster depth (plus the depth bias) is not in front of the sample.
float bake_shadow(const orc_bake_view& v, const float* shadow, float px, float py, float depth, uint32_t r) {
    const int32_t cx = f2i_sat(std::floor(px)), cy = f2i_sat(std::floor(py));
    float visible = 0.0f, count = 0.0f;
    for (int32_t y = -static_cast<int32_t>(r); y <= static_cast<int32_t>(r); ++y)
        for (int32_t x = -static_cast<int32_t>(r); x <= static_cast<int32_t>(r); ++x) {
            const int32_t qx = cx + x, qy = cy + y;
            if (qx < 0 || qy < 0 || qx >= static_cast<int32_t>(v.width) || qy >= static_cast<int32_t>(v.height)) continue;
            const size_t o = (static_cast<size_t>(qy) * v.width + qx) * 4;
            if (decode_id(shadow[o + 3]) != 0u) {
                visible += depth <= shadow[o + 2] + v.depth_bias ? 1.0f : 0.0f;
                count += 1.0f;
            }
        }
    return count > 0.0f ? visible / count : 0.0f;
}

// One view's sample at atlas texel p (shaders/project_texture.hlsl:126-175):
// false when the view contributes nothing.
struct BakeSample {
    float vis, cosv, dl, col[3], alpha;
};

bool project_sample(const orc_bake_view& v, const float* shadow, const float* apos, const float* anrm, size_t p,
                    uint32_t pcf, BakeSample& s) {
    const float x = apos[p * 3], y = apos[p * 3 + 1], z = apos[p * 3 + 2];
    const float n0 = anrm[p * 3], n1 = anrm[p * 3 + 1], n2 = anrm[p * 3 + 2];
    const float nl = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    const float nx = n0 / nl, ny = n1 / nl, nz = n2 / nl;
    const float* m = v.world_to_clip;
    float c[4];
    for (int r = 0; r < 4; ++r) c[r] = m[r * 4] * x + m[r * 4 + 1] * y + m[r * 4 + 2] * z + m[r * 4 + 3] * 1.0f;
    if (c[3] <= 1e-8f) return false;
    const float ndx = c[0] / c[3], ndy = c[1] / c[3], ndz = c[2] / c[3];
    if (ndx < -1.0f || ndy < -1.0f || ndx > 1.0f || ndy > 1.0f || ndz < -1.0f || ndz > 1.0f) return false;
    const float px = v.viewport[0] + (ndx * 0.5f + 0.5f) * v.viewport[2];
    const float py = v.viewport[1] + (ndy * 0.5f + 0.5f) * v.viewport[3];
    if (px < v.viewport[0] || py < v.viewport[1] || px >= v.viewport[0] + v.viewport[2] ||
        py >= v.viewport[1] + v.viewport[3])
        return false;
    float vis = bake_shadow(v, shadow, px, py, ndz, pcf);
    if (v.mask) vis *= clampf(bake_mask(v, px, py), 0.0f, 1.0f);
    const float dx = v.camera[0] - x, dy = v.camera[1] - y, dz = v.camera[2] - z;
    const float dl = std::sqrt(dx * dx + dy * dy + dz * dz);
    const float vx = dx / dl, vy = dy / dl, vz = dz / dl;
    const float cosv = nx * vx + ny * vy + nz * vz;
    if (cosv < v.min_view_cosine || vis <= 0.0f) return false;
    s.vis = vis;
    s.cosv = cosv;
    s.dl = dl;
    for (uint32_t ch = 0; ch < 3; ++ch) s.col[ch] = bake_photo(v, px, py, ch);
    s.alpha = v.channels == 4 ? bake_photo(v, px, py, 3) : 1.0f;
    return true;
}

// Atlas raster of uv * 2 - 1 and the texels' interpolated positions and
// normals (src/texture_baker.cpp:85-94, 183-192).
int atlas_frames(const float* pos3, const float* normal3, const float* uv2, uint32_t V, const uint32_t* idx, uint32_t T,
                 uint32_t W, uint32_t H, std::vector<float>& araster, std::vector<float>& apos,
                 std::vector<float>& anrm) {
    const size_t P = static_cast<size_t>(W) * H;
    std::vector<float> aclip(static_cast<size_t>(V) * 4);
    for (size_t v = 0; v < V; ++v) {
        aclip[v * 4] = uv2[v * 2] * 2.0f - 1.0f;
        aclip[v * 4 + 1] = uv2[v * 2 + 1] * 2.0f - 1.0f;
        aclip[v * 4 + 2] = 0.0f;
        aclip[v * 4 + 3] = 1.0f;
    }
    araster.assign(P * 4, 0.0f);
    apos.assign(P * 3, 0.0f);
    anrm.assign(P * 3, 0.0f);
    float vp_out[4];
    const int rc = orc_rasterize_forward(aclip.data(), V, idx, T, W, H, 0, 0, nullptr, araster.data(), nullptr, vp_out);
    if (rc) return rc;
    orc_interpolate_forward(pos3, V, 3, idx, T, araster.data(), static_cast<uint32_t>(P), apos.data());
    orc_interpolate_forward(normal3, V, 3, idx, T, araster.data(), static_cast<uint32_t>(P), anrm.data());
    return 0;
}

// Shadow raster of one view (src/texture_baker.cpp:266-276).
int shadow_map(const float* pos3, uint32_t V, const uint32_t* idx, uint32_t T, const orc_bake_view& v,
               std::vector<float>& shadow) {
    std::vector<float> clip(static_cast<size_t>(V) * 4);
    shadow.assign(static_cast<size_t>(v.width) * v.height * 4, 0.0f);
    orc_transform_positions(pos3, V, v.world_to_clip, clip.data());
    float vp_out[4];
    return orc_rasterize_forward(clip.data(), V, idx, T, v.width, v.height, 0, 0, v.viewport, shadow.data(), nullptr,
                                 vp_out);
}

}  // namespace

float orc_suggest_softmax_scale(const float* pos3, uint32_t V, const float* cams, uint32_t nc) {
    const size_t stride = std::max<size_t>(1, V / 10000);
    double best = std::numeric_limits<double>::infinity();
    for (size_t v = 0; v < V; v += stride) {
        const double x = pos3[v * 3], y = pos3[v * 3 + 1], z = pos3[v * 3 + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) continue;
        for (size_t c = 0; c < nc; ++c) {
            const double dx = cams[c * 3] - x, dy = cams[c * 3 + 1] - y, dz = cams[c * 3 + 2] - z;
            best = std::min(best, dx * dx + dy * dy + dz * dz);
        }
    }
    if (!std::isfinite(best) || best <= 0.0) return 0.0f;
    return static_cast<float>(20.0 * best);
}

int orc_bake(const float* pos3, const float* normal3, const float* uv2, uint32_t V, const uint32_t* idx, uint32_t T,
             const orc_bake_view* views, uint32_t nviews, uint32_t W, uint32_t H, uint32_t blend, uint32_t pcf,
             float* color4, float* confidence, uint32_t* source_view, float* valid, float* coverage,
             uint32_t* face_ids) {
    const size_t P = static_cast<size_t>(W) * H;
    std::vector<float> araster, apos, anrm;
    int rc = atlas_frames(pos3, normal3, uv2, V, idx, T, W, H, araster, apos, anrm);
    if (rc) return rc;
    // src/texture_baker.cpp:193-204 coverage and face ids
    for (size_t p = 0; p < P; ++p) {
        const float id = std::round(araster[p * 4 + 3]);
        coverage[p] = id != 0.0f ? 1.0f : 0.0f;
        if (face_ids) face_ids[p] = coverage[p] > 0.5f ? static_cast<uint32_t>(id) - 1u : 0xFFFFFFFFu;
    }
    std::vector<float> acc(P * 3, 0.0f), accw(P, 0.0f), best(P, 0.0f), shadow;
    for (size_t p = 0; p < P; ++p) source_view[p] = 0xFFFFFFFFu;
    for (uint32_t vi = 0; vi < nviews; ++vi) {
        const orc_bake_view& v = views[vi];
        rc = shadow_map(pos3, V, idx, T, v, shadow);
        if (rc) return rc;
#pragma omp parallel for schedule(static)
        for (int64_t pi = 0; pi < static_cast<int64_t>(P); ++pi) {
            const size_t p = static_cast<size_t>(pi);
            if (decode_id(araster[p * 4 + 3]) == 0u) continue;  // project_texture.hlsl:121-125
            BakeSample s;
            if (!project_sample(v, shadow.data(), apos.data(), anrm.data(), p, pcf, s)) continue;
            const float vis = s.vis, cosv = s.cosv, dl = s.dl, alpha = s.alpha;
            const float* col = s.col;
            if (blend == 2) {  // project_texture.hlsl:176-206
                const float dist = std::max(dl, 1e-6f);
                const float footprint = cosv / (dist * dist);
                const float e = clampf(vis * alpha * footprint * v.weight, 0.0f, 64.0f);
                const float prev = best[p];
                const float cur = std::max(prev, e);
                if (cur > prev) {
                    const float rs = std::exp(prev - cur);
                    acc[p * 3] *= rs;
                    acc[p * 3 + 1] *= rs;
                    acc[p * 3 + 2] *= rs;
                    accw[p] *= rs;
                    best[p] = cur;
                }
                if (e >= prev) source_view[p] = vi;
                const float w = std::exp(e - cur);
                if (w <= 0.0f) continue;
                acc[p * 3] += col[0] * w;
                acc[p * 3 + 1] += col[1] * w;
                acc[p * 3 + 2] += col[2] * w;
                accw[p] += w;
                continue;
            }
            const float conf = vis * alpha * cosv * v.weight;  // project_texture.hlsl:208-246
            if (conf <= 0.0f) continue;
            if (conf > best[p]) {
                best[p] = conf;
                source_view[p] = vi;
            }
            if (blend == 0) {
                if (conf >= accw[p]) {
                    acc[p * 3] = col[0];
                    acc[p * 3 + 1] = col[1];
                    acc[p * 3 + 2] = col[2];
                    accw[p] = conf;
                }
            } else {
                acc[p * 3] += col[0] * conf;
                acc[p * 3 + 1] += col[1] * conf;
                acc[p * 3 + 2] += col[2] * conf;
                accw[p] += conf;
            }
        }
    }
    // src/texture_baker.cpp:339-357
    for (size_t p = 0; p < P; ++p) {
        const float w = accw[p];
        confidence[p] = w;
        valid[p] = w > 0.0f ? 1.0f : 0.0f;
        const float d = blend != 0 && w > 0.0f ? w : 1.0f;
        color4[p * 4] = acc[p * 3] / d;
        color4[p * 4 + 1] = acc[p * 3 + 1] / d;
        color4[p * 4 + 2] = acc[p * 3 + 2] / d;
        color4[p * 4 + 3] = valid[p];
    }
    return 0;
}

int orc_bake_features(const float* pos3, const float* normal3, const float* uv2, uint32_t V, const uint32_t* idx,
                      uint32_t T, const orc_bake_view* views, uint32_t nviews, uint32_t W, uint32_t H, uint32_t pcf,
                      float* features40) {
    const size_t P = static_cast<size_t>(W) * H;
    const int K = 8;
    std::vector<float> araster, apos, anrm, shadow;
    int rc = atlas_frames(pos3, normal3, uv2, V, idx, T, W, H, araster, apos, anrm);
    if (rc) return rc;
    std::fill(features40, features40 + P * 5 * K, 0.0f);
    for (uint32_t vi = 0; vi < nviews; ++vi) {
        const orc_bake_view& v = views[vi];
        rc = shadow_map(pos3, V, idx, T, v, shadow);
        if (rc) return rc;
#pragma omp parallel for schedule(static)
        for (int64_t pi = 0; pi < static_cast<int64_t>(P); ++pi) {
            const size_t p = static_cast<size_t>(pi);
            if (decode_id(araster[p * 4 + 3]) == 0u) continue;
            BakeSample s;
            if (!project_sample(v, shadow.data(), apos.data(), anrm.data(), p, pcf, s)) continue;
            const float conf = s.vis * s.alpha * s.cosv * v.weight;  // project_texture.hlsl:208
            if (conf <= 0.0f) continue;
            // the list is sorted by confidence, descending; a new sample goes after every slot of equal or
            // higher confidence, and the last slot falls off
            float* f = features40 + p * 5 * K;
            int at = 0;
            while (at < K && f[at * 5 + 3] >= conf) ++at;
            if (at == K) continue;
            for (int i = K - 1; i > at; --i)
                for (int e = 0; e < 5; ++e) f[i * 5 + e] = f[(i - 1) * 5 + e];
            const float entry[5] = {s.col[0], s.col[1], s.col[2], conf, s.cosv};
            for (int e = 0; e < 5; ++e) f[at * 5 + e] = entry[e];
        }
    }
    return 0;
}


int orc_pad_texture_atlas(float* color, uint32_t W, uint32_t H, uint32_t C, const float* valid, const float* coverage,
                          float* filled, uint32_t margin, int fill_unobserved, uint32_t max_fill, float threshold,
                          uint64_t* st) {
    const int32_t kSide = 16384, kFar = 12000;  // src/texture_padding.cpp:18-21
    if (W == 0 || H == 0) return fail(1, "cannot pad an empty atlas");
    if (W > static_cast<uint32_t>(kSide) || H > static_cast<uint32_t>(kSide))
        return fail(1, "atlas dimensions exceed the supported maximum");
    if (C != 3 && C != 4) return fail(1, "atlas color must have 3 or 4 channels");
    if (!(threshold > 0.0f && threshold <= 1.0f)) return fail(1, "valid_threshold must be within (0, 1]");
    const size_t P = static_cast<size_t>(W) * H;
    std::vector<float> no_fill;
    if (!filled) {
        no_fill.assign(P, 0.0f);
        filled = no_fill.data();
    }
    for (int i = 0; i < 5; ++i) st[i] = 0;
    // sources: valid texels and earlier covered fills (:259-276)
    std::vector<uint8_t> src(P, 0);
    for (size_t p = 0; p < P; ++p) {
        if (valid[p] >= threshold) {
            src[p] = 1;
            ++st[0];
        } else if (filled[p] > 0.5f && coverage && coverage[p] > 0.5f) {
            src[p] = 1;
            ++st[2];
            ++st[4];
        } else {
            filled[p] = 0.0f;
        }
    }
    if (st[0] == 0 && st[4] == 0) return 0;
    // nearest-source offsets (:37-97)
    std::vector<int32_t> ox(P), oy(P);
    for (size_t p = 0; p < P; ++p) ox[p] = oy[p] = src[p] ? 0 : kFar;
    auto try_from = [&](int32_t x, int32_t y, int32_t nx, int32_t ny) {
        if (nx < 0 || ny < 0 || nx >= static_cast<int32_t>(W) || ny >= static_cast<int32_t>(H)) return;
        const size_t c = static_cast<size_t>(y) * W + x, n = static_cast<size_t>(ny) * W + nx;
        const int64_t cx = ox[n] + (nx - x), cy = oy[n] + (ny - y);
        if (cx * cx + cy * cy >= static_cast<int64_t>(ox[c]) * ox[c] + static_cast<int64_t>(oy[c]) * oy[c]) return;
        ox[c] = static_cast<int32_t>(std::min<int64_t>(std::max<int64_t>(cx, -kSide), kSide));
        oy[c] = static_cast<int32_t>(std::min<int64_t>(std::max<int64_t>(cy, -kSide), kSide));
    };
    for (int32_t y = 0; y < static_cast<int32_t>(H); ++y)
        for (int32_t x = 0; x < static_cast<int32_t>(W); ++x) {
            try_from(x, y, x - 1, y);
            try_from(x, y, x, y - 1);
            try_from(x, y, x - 1, y - 1);
            try_from(x, y, x + 1, y - 1);
        }
    for (int32_t y = static_cast<int32_t>(H) - 1; y >= 0; --y)
        for (int32_t x = static_cast<int32_t>(W) - 1; x >= 0; --x) {
            try_from(x, y, x + 1, y);
            try_from(x, y, x, y + 1);
            try_from(x, y, x + 1, y + 1);
            try_from(x, y, x - 1, y + 1);
        }
    // fills (:289-348); sources are never written, so reading in place is order-independent
    const int64_t margin2 = static_cast<int64_t>(margin) * margin;
    const int64_t limit2 = max_fill == 0 ? INT64_MAX : static_cast<int64_t>(max_fill) * max_fill;
    const int64_t none2 = 2LL * kFar * kFar;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const size_t p = static_cast<size_t>(y) * W + x;
            if (src[p]) continue;
            const int64_t d2 = static_cast<int64_t>(ox[p]) * ox[p] + static_cast<int64_t>(oy[p]) * oy[p];
            if (d2 >= none2) continue;
            const int64_t sx = static_cast<int64_t>(x) + ox[p], sy = static_cast<int64_t>(y) + oy[p];
            if (sx < 0 || sy < 0 || sx >= W || sy >= H) continue;
            const size_t s = static_cast<size_t>(sy) * W + static_cast<size_t>(sx);
            const bool covered = coverage && coverage[p] > 0.5f;
            const bool fill = covered ? (fill_unobserved && d2 <= limit2) : (margin > 0 && d2 <= margin2);
            if (!fill) {
                if (covered) ++st[3];
                continue;
            }
            for (uint32_t c = 0; c < C; ++c) color[p * C + c] = color[s * C + c];
            if (C == 4) color[p * C + 3] = 1.0f;
            filled[p] = 1.0f;
            ++st[4];
            ++st[covered ? 2 : 1];
        }
    return 0;
}

}  // extern "C"
