// Screen-space triangle setup shared by the raster kernels (and callable on the
// host, where the library's self-test hooks use it).
//
// setup_screen restates the reference's per-triangle validity and bounds:
//   binning validity    src/rasterizer.cpp:73-100 (w > 1e-8, finite screen
//                       coordinates, screen bbox overlapping [-1, 1]^2);
//   shader validity     shaders/rasterize_forward.hlsl:75-101 (|area| > 1e-12,
//                       cull mode);
//   bin rectangle       src/rasterizer.cpp:101-116, expanded to whole 16x16
//                       tiles (a pixel is a candidate of a triangle iff the
//                       triangle is in the pixel's tile list).
// It then bounds the pixels whose float edge test (rasterize_forward.hlsl:
// 103-113, e * sign >= -1e-7 for all three edges) can pass:
//   * the exact edge lines are shifted outward by 1e-7 plus a bound on the
//     float evaluation error of the edge test (32 u M^2, M the largest
//     vertex-to-candidate distance) and on the double rounding of the line
//     coefficients, giving three half-planes a x + b y >= c in NDC;
//   * the rectangle of their intersection (halo rectangle) cuts the bin
//     rectangle;
//   * per line of pixels (row, or column for triangles taller than wide), the
//     half-planes give conservative bounds along the line (LineSpans): with
//     the line's float sample coordinate, each edge not nearly parallel to
//     the line bounds the span from one side, evaluated in float and padded
//     for the rounding of the coefficients, of the evaluation and of the
//     float pixel samples.
// Every pixel that the float test can accept lies in its line's span; the
// spans are a superset, never a subset, of the covered pixels.
#pragma once

#include "pdr_internal.hpp"

namespace pdr {

struct ScreenTri {
    float4 p0, p1, p2;
    float s0x, s0y, s1x, s1y, s2x, s2y, sgn;
    int32_t x0, x1, y0, y1;  // inclusive candidate rectangle (bins and halo)
    float work;              // estimated candidate pixels + lines
};

// Conservative spans of the candidate region along lines of pixels: rows
// (cols == 0; span over x for the row's sample y) or columns (cols == 1; span
// over y for the column's sample x).  Edge i bounds the span from below
// (bit i of lower) or above (upper) at u[i] + v[i] * s, in pixels, s the
// line's float sample coordinate.
struct LineSpans {
    float u[3], v[3];
    uint32_t lower, upper, cols;
};

// Span [*lo, *hi] of the line with sample s, inside [a0, a1]; empty when
// *lo > *hi.  alpha, beta: pixel = alpha * ndc + beta along the span's axis.
// The float evaluation is padded by 1e-6 of the magnitudes involved (it
// rounds the coefficients and two operations, < 2^-22 relative) and by the
// 1e-5 NDC allowance for the rounding of the float pixel samples.
__host__ __device__ __forceinline__ void line_span(const LineSpans& s, float sv, float alpha, float beta, int32_t a0,
                                                   int32_t a1, int32_t* lo, int32_t* hi) {
    const float inf = __builtin_inff();
    float l = -inf, h = inf, el = 0.0f, eh = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float vs = s.v[i] * sv;
        const float b = s.u[i] + vs;
        const float e = fabsf(s.u[i]) + fabsf(vs);
        if (((s.lower >> i) & 1u) && b > l) {
            l = b;
            el = e;
        }
        if (((s.upper >> i) & 1u) && b < h) {
            h = b;
            eh = e;
        }
    }
    int32_t a = a0, z = a1;
    if (l > -inf) {
        const float c = ceilf(l - (1e-5f * (alpha + fabsf(l - beta)) + 1e-6f * el + 1e-6f));
        if (c > static_cast<float>(a)) a = c > static_cast<float>(a1) ? add1(a1) : static_cast<int32_t>(c);
    }
    if (h < inf) {
        const float c = floorf(h + (1e-5f * (alpha + fabsf(h - beta)) + 1e-6f * eh + 1e-6f));
        if (c < static_cast<float>(z)) z = c < static_cast<float>(a0) ? a0 - 1 : static_cast<int32_t>(c);
    }
    *lo = a;
    *hi = z;
}

// Shifted half-planes a x + b y >= c (NDC) of r and their rectangle, for the
// candidate rectangle [cx0, cx1] x [cy0, cy1].  Returns false when the
// shifted lines are ill-conditioned; the caller then keeps the rectangle and
// uses no spans.
__host__ __device__ inline bool halo_bounds(const ScreenTri& r, double vx, double vy, double vw, double vh,
                                            int32_t cx0, int32_t cx1, int32_t cy0, int32_t cy1, int32_t* hx0,
                                            int32_t* hx1, int32_t* hy0, int32_t* hy1, double a[3], double b[3],
                                            double c[3]) {
    const double S[3][2] = {{r.s0x, r.s0y}, {r.s1x, r.s1y}, {r.s2x, r.s2y}};
    // reciprocals instead of divisions: one more rounding (2^-53 relative) is far inside the pads below
    const double sxv = 2.0 / vw, syv = 2.0 / vh;
    auto ndc_x = [&](double px) { return ((px - vx) + 0.5) * sxv - 1.0; };
    auto ndc_y = [&](double py) { return ((py - vy) + 0.5) * syv - 1.0; };
    const double nx0 = ndc_x(cx0), nx1 = ndc_x(cx1), ny0 = ndc_y(cy0), ny1 = ndc_y(cy1);
    double M = 0.0, cmax = 0.0;
    for (int j = 0; j < 3; ++j) {
        M = fmax(M, fmax(fmax(fabs(S[j][0] - nx0), fabs(S[j][0] - nx1)), fmax(fabs(S[j][1] - ny0), fabs(S[j][1] - ny1))));
        cmax = fmax(cmax, fmax(fabs(S[j][0]), fabs(S[j][1])));
    }
    // float edge test error <= ~4u (|ax||by| + |ay||bx|) <= 8u M^2; take 32u M^2
    const double eps = 1.0000001e-7 + 32.0 * 5.9604644775390625e-8 * M * M + 1e-15 * cmax * cmax;
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
        const double inv = 1.0 / det;
        const double x = (c[i] * b[j] - c[j] * b[i]) * inv;
        const double y = (a[i] * c[j] - a[j] * c[i]) * inv;
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

// Span coefficients along rows (cols = 0: x >= (c - b y) / a for a > 0, <= for
// a < 0) or columns (cols = 1: the same in y with a and b exchanged), in
// pixels of the span's axis.  Edges nearly parallel to the lines (|coefficient
// of the span's axis| < 1e-3 of the other) bound the line range, not the span,
// and are skipped.
__host__ __device__ inline void make_spans(const double a[3], const double b[3], const double c[3], uint32_t cols,
                                           double v0, double vsize, LineSpans& sp) {
    const double alpha = vsize * 0.5, beta = vsize * 0.5 + v0 - 0.5;
    sp.lower = sp.upper = 0u;
    sp.cols = cols;
    for (int i = 0; i < 3; ++i) {
        const double along = cols ? b[i] : a[i], across = cols ? a[i] : b[i];
        sp.u[i] = 0.0f;
        sp.v[i] = 0.0f;
        if (!(fabs(along) >= 1e-3 * fabs(across)) || along == 0.0) continue;
        const double r = 1.0 / along;  // the float conversion below rounds far more than the reciprocal
        sp.u[i] = static_cast<float>(alpha * (c[i] * r) + beta);
        sp.v[i] = static_cast<float>(-alpha * (across * r));
        if (along > 0.0)
            sp.lower |= 1u << i;
        else
            sp.upper |= 1u << i;
    }
}

// Binning and shader validity of (p0, p1, p2) (src/rasterizer.cpp:73-100,
// rasterize_forward.hlsl:75-101): fills r's corners, screen coordinates and
// winding sign, and the screen bounding box bb = (min x, min y, max x, max y).
__host__ __device__ inline bool screen_coords(const float4& p0, const float4& p1, const float4& p2, uint32_t cull,
                                              ScreenTri& r, float bb[4]) {
    r.p0 = p0;
    r.p1 = p1;
    r.p2 = p2;
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
    r.work = fabsf(sa);
    bb[0] = min_x;
    bb[1] = min_y;
    bb[2] = max_x;
    bb[3] = max_y;
    return true;
}

// The reference's float edge test (rasterize_forward.hlsl:103-113).
template <class Tri>
__host__ __device__ __forceinline__ bool edge_pass(const Tri& r, float sx, float sy) {
    const float e0 = cross2(r.s1x - sx, r.s1y - sy, r.s2x - sx, r.s2y - sy);
    const float e1 = cross2(r.s2x - sx, r.s2y - sy, r.s0x - sx, r.s0y - sy);
    const float e2 = cross2(r.s0x - sx, r.s0y - sy, r.s1x - sx, r.s1y - sy);
    return !(e0 * r.sgn < -1e-7f || e1 * r.sgn < -1e-7f || e2 * r.sgn < -1e-7f);
}

// Forward setup of triangle (p0, p1, p2) in view vd.  Returns false when the
// triangle is rejected or has no candidate pixel.  With allow_cols, a
// triangle whose candidate rectangle is taller than wide gets column spans.
__host__ __device__ inline bool setup_screen(const ViewDesc& vd, const float4& p0, const float4& p1, const float4& p2,
                                             uint32_t cull, bool allow_cols, ScreenTri& r, LineSpans& sp) {
    float bb[4];
    if (!screen_coords(p0, p1, p2, cull, r, bb)) return false;
    const float min_x = bb[0], min_y = bb[1], max_x = bb[2], max_y = bb[3];
    const float sa = r.work;
    // src/rasterizer.cpp:101-116 bounds, expanded to whole 16x16 tiles
    const int64_t wm = static_cast<int64_t>(vd.width) - 1, hm = static_cast<int64_t>(vd.height) - 1;
    auto cl = [](int64_t v, int64_t hi) { return static_cast<int32_t>(v < 0 ? 0 : (v > hi ? hi : v)); };
    const int32_t bx0 = cl(f2i_sat(floorf((min_x * 0.5f + 0.5f) * vd.vw + vd.vx)), wm);
    const int32_t bx1 = cl(static_cast<int64_t>(f2i_sat(ceilf((max_x * 0.5f + 0.5f) * vd.vw + vd.vx))) - 1, wm);
    const int32_t by0 = cl(f2i_sat(floorf((min_y * 0.5f + 0.5f) * vd.vh + vd.vy)), hm);
    const int32_t by1 = cl(static_cast<int64_t>(f2i_sat(ceilf((max_y * 0.5f + 0.5f) * vd.vh + vd.vy))) - 1, hm);
    const int32_t tile = static_cast<int32_t>(kTile);
    int32_t cx0 = (bx0 / tile) * tile;
    int32_t cx1 = static_cast<int32_t>((bx1 / tile) * tile + tile - 1 < wm ? (bx1 / tile) * tile + tile - 1 : wm);
    int32_t cy0 = (by0 / tile) * tile;
    int32_t cy1 = static_cast<int32_t>((by1 / tile) * tile + tile - 1 < hm ? (by1 / tile) * tile + tile - 1 : hm);
    int32_t hx0, hx1, hy0, hy1;
    double a[3], b[3], c[3];
    const bool spans = halo_bounds(r, vd.vx, vd.vy, vd.vw, vd.vh, cx0, cx1, cy0, cy1, &hx0, &hx1, &hy0, &hy1, a, b, c);
    if (spans) {
        cx0 = cx0 > hx0 ? cx0 : hx0;
        cx1 = cx1 < hx1 ? cx1 : hx1;
        cy0 = cy0 > hy0 ? cy0 : hy0;
        cy1 = cy1 < hy1 ? cy1 : hy1;
    }
    r.x0 = cx0;
    r.x1 = cx1;
    r.y0 = cy0;
    r.y1 = cy1;
    if (cx0 > cx1 || cy0 > cy1) return false;
    const float rows = static_cast<float>(cy1 - cy0 + 1), cols = static_cast<float>(cx1 - cx0 + 1);
    const uint32_t by_cols = allow_cols && cols < rows ? 1u : 0u;
    if (spans) {
        if (by_cols)
            make_spans(a, b, c, 1u, vd.vy, vd.vh, sp);
        else
            make_spans(a, b, c, 0u, vd.vx, vd.vw, sp);
    } else {
        sp.lower = sp.upper = 0u;
        sp.cols = by_cols;
        for (int i = 0; i < 3; ++i) sp.u[i] = sp.v[i] = 0.0f;
    }
    // screen area in pixels (|sa| / 2 in NDC^2) plus a pixel of halo per line
    const float area = sa * vd.vw * vd.vh * 0.125f;
    const float lines = by_cols ? cols : rows;
    r.work = spans ? minf_std(area + 3.0f * lines, rows * cols) : rows * cols;
    return true;
}

// Reference screen bounding box of the backward pass (rasterize_backward.hlsl:
// 66-87): [x0, x1) x [y0, y1), clamped to the image; false when some w <= 1e-8.
__host__ __device__ inline bool backward_box(const ViewDesc& vd, const float4& p0, const float4& p1, const float4& p2,
                                             int32_t* x0, int32_t* x1, int32_t* y0, int32_t* y1) {
    if (p0.w <= 1e-8f || p1.w <= 1e-8f || p2.w <= 1e-8f) return false;
    const float s0x = p0.x / p0.w, s0y = p0.y / p0.w;
    const float s1x = p1.x / p1.w, s1y = p1.y / p1.w;
    const float s2x = p2.x / p2.w, s2y = p2.y / p2.w;
    const float mnx = minf_std(s0x, minf_std(s1x, s2x)), mny = minf_std(s0y, minf_std(s1y, s2y));
    const float mxx = maxf_std(s0x, maxf_std(s1x, s2x)), mxy = maxf_std(s0y, maxf_std(s1y, s2y));
    auto cl = [](int32_t v, int32_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    const int32_t W = static_cast<int32_t>(vd.width), H = static_cast<int32_t>(vd.height);
    *x0 = cl(f2i_sat(floorf((mnx * 0.5f + 0.5f) * vd.vw + vd.vx)), W);
    *x1 = cl(f2i_sat(ceilf((mxx * 0.5f + 0.5f) * vd.vw + vd.vx)), W);
    *y0 = cl(f2i_sat(floorf((mny * 0.5f + 0.5f) * vd.vh + vd.vy)), H);
    *y1 = cl(f2i_sat(ceilf((mxy * 0.5f + 0.5f) * vd.vh + vd.vy)), H);
    return true;
}

// pixel = alpha * ndc + beta along x (cols = 0) or y (cols = 1)
__host__ __device__ __forceinline__ float span_alpha(const ViewDesc& vd, uint32_t cols) {
    return (cols ? vd.vh : vd.vw) * 0.5f;
}
__host__ __device__ __forceinline__ float span_beta(const ViewDesc& vd, uint32_t cols) {
    return (cols ? vd.vh : vd.vw) * 0.5f + (cols ? vd.vy : vd.vx) - 0.5f;
}

}  // namespace pdr
