// Rasterize forward and backward for gfx950, triangle-parallel.
//
// Forward (shaders/rasterize_forward.hlsl:45-175; candidates from the tile bins
// of src/rasterizer.cpp:61-138):
//   k_raster_small    one lane per (view, triangle) sets the triangle up
//                     (raster_geom.hpp): its candidate pixels are those of its
//                     16x16 bins (the reference's pixel bounds, tile-expanded)
//                     that lie in conservative per-row spans of every pixel
//                     able to pass the reference's epsilon-inclusive edge
//                     test.  Each candidate runs the reference's coverage /
//                     depth expressions and, if covered, does a 64-bit
//                     atomicMin of (depth, id) into the pixel's key.  Small
//                     triangles are walked by their own lane, medium ones by
//                     the wave (rows' spans concatenated, 64 pixels per step),
//                     large ones are queued;
//   k_raster_large    the queued triangles' 16-row slabs, dealt round-robin to
//                     every wave of a persistent grid;
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
//   k_bwd_prepass     per pixel: the id the backward compares (0 when the
//                     pixel's grad_raster.xy is zero) and a per-view flag when
//                     a pixel names a triangle whose candidate region does not
//                     hold it (only possible for a raster not produced by the
//                     forward; such views are walked over whole boxes), and
//                     a per-(view, triangle) mark of the triangles some pixel
//                     names (most of an overdrawn view's triangles name none:
//                     their corner sums are zero and are written without a walk);
//   k_corner_small    per (view, triangle): the corner gradients summed over
//                     the reference's screen bounding box (restricted to the
//                     candidate region, outside which no pixel names the
//                     triangle) in the reference's row-major order, by the
//                     triangle's lane or, for larger regions, by the wave with
//                     an ordered ballot/readlane fold (bit-exact per-triangle
//                     sums); boxes over kWaveBox pixels are queued;
//   k_corner_large    queued boxes, one workgroup each, fixed-order reduction;
//   k_vertex_reduce   per vertex, corner gradients summed in ascending corner
//                     order over a vertex -> corner adjacency built on the GPU
//                     as the reference's CSR (counts, exclusive scan, atomic
//                     placement, then each vertex's short list sorted by one
//                     lane).  The corner kernels write each corner's gradient
//                     at its position in that list, so the reduce streams.
// No float atomics: the backward is deterministic.
#include <hipcub/hipcub.hpp>

#include "pdr_internal.hpp"
#include "raster_geom.hpp"

namespace pdr {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kSlabRows = 16;       // forward: rows per queued work item
constexpr uint32_t kWaveBox = 65536;     // backward: box pixels summed by one wave in pixel order (else queued)
// The forward's wave-walk limit (TuneParams::wave_work, else queued) and the
// backward's concatenated-walk limit (TuneParams::flat_work, else the ordered
// wave fold) are launch parameters chosen per workload class.
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

// 16-byte stores over a 16-byte aligned p: n / 4 float4, then the n % 4 tail.
__global__ void k_fill_f32x4(float* p, float v, size_t n) {
    const size_t n4 = n / 4;
    float4* q = reinterpret_cast<float4*>(p);
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n4;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        q[i] = make_float4(v, v, v, v);
    if (blockIdx.x == 0 && threadIdx.x < n % 4) p[n4 * 4 + threadIdx.x] = v;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

uint32_t grid_for(size_t n, uint32_t block = kBlock) {
    const size_t g = (n + block - 1) / block;
    return static_cast<uint32_t>(g < 65536 ? (g ? g : 1) : 65536);
}

// Pixel-centre NDC of column x / row y (rasterize_forward.hlsl:55-57); every
// kernel evaluates it inline (a correctly rounded division), so no per-call
// table launch is needed.
__device__ __forceinline__ float sample_x(const ViewDesc& vd, int32_t x) {
    return (static_cast<float>(x) - vd.vx + 0.5f) * 2.0f / vd.vw - 1.0f;
}
__device__ __forceinline__ float sample_y(const ViewDesc& vd, int32_t y) {
    return (static_cast<float>(y) - vd.vy + 0.5f) * 2.0f / vd.vh - 1.0f;
}

__device__ __forceinline__ uint32_t order_key(float depth) {
    const uint32_t u = __float_as_uint(depth + 0.0f);  // -0 -> +0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// rasterize_forward.hlsl:103-135 for one candidate pixel; returns the key or ~0.
template <class Tri>
__device__ __forceinline__ unsigned long long test_pixel(const Tri& r, uint32_t t, float sx, float sy) {
    if (!edge_pass(r, sx, sy)) return ~0ull;
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

// Keys only decrease, so a stale read can only make the filter let an atomic through.
__device__ __forceinline__ void key_min(unsigned long long* p, unsigned long long key) {
    if (key < __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMin(p, key);
}

// The large-work queue's counters (counters[0] queue length, [1] next entry)
// are zero between calls: the producers count up from zero, and the kernel
// that follows the queue's consumer in the same call (the resolve, the tile
// pass or the vertex reduce) clears them from one lane, so no fill launch is
// needed before the next call.
__device__ __forceinline__ void clear_queue_counters(uint32_t* counters) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        counters[0] = 0u;
        counters[1] = 0u;
    }
}

// Intra-wave hand-off of LDS data written by other lanes of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- wave-uniform broadcast of lane j's values (j uniform)
__device__ __forceinline__ float bcf(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ int32_t bci(int32_t v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ float4 bcf4(const float4& v, int j) {
    return make_float4(bcf(v.x, j), bcf(v.y, j), bcf(v.z, j), bcf(v.w, j));
}
__device__ __forceinline__ void bc_spans(const LineSpans& s, int j, LineSpans& o) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        o.u[i] = bcf(s.u[i], j);
        o.v[i] = bcf(s.v[i], j);
    }
    o.lower = static_cast<uint32_t>(bci(static_cast<int32_t>(s.lower), j));
    o.upper = static_cast<uint32_t>(bci(static_cast<int32_t>(s.upper), j));
    o.cols = static_cast<uint32_t>(bci(static_cast<int32_t>(s.cols), j));
}

template <int Ctrl, int RowMask>
__device__ __forceinline__ int32_t dpp_move(int32_t old, int32_t v) {
    return __builtin_amdgcn_update_dpp(old, v, Ctrl, RowMask, 0xf, false);
}

// Inclusive prefix sum over the 64 lanes (all lanes active): Hillis-Steele
// within each 16-lane row (DPP row_shr 1, 2, 4, 8), then the row totals
// (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3).
__device__ __forceinline__ int32_t wave_incl_scan(int32_t v) {
    v += dpp_move<0x111, 0xf>(0, v);
    v += dpp_move<0x112, 0xf>(0, v);
    v += dpp_move<0x114, 0xf>(0, v);
    v += dpp_move<0x118, 0xf>(0, v);
    v += dpp_move<0x142, 0xa>(0, v);
    v += dpp_move<0x143, 0xc>(0, v);
    return v;
}

// Inclusive prefix maximum over the 64 lanes (values >= -1; all lanes active).
__device__ __forceinline__ int32_t wave_incl_max(int32_t v) {
    v = max(v, dpp_move<0x111, 0xf>(-1, v));
    v = max(v, dpp_move<0x112, 0xf>(-1, v));
    v = max(v, dpp_move<0x114, 0xf>(-1, v));
    v = max(v, dpp_move<0x118, 0xf>(-1, v));
    v = max(v, dpp_move<0x142, 0xa>(-1, v));
    v = max(v, dpp_move<0x143, 0xc>(-1, v));
    return v;
}

// The lanes' segments [incl - len, incl) concatenated in lane order: the lane
// owning position pc + lane (meaningful when it is below the total).  The
// owner of pc is the number of lanes ending at or before pc; every lane whose
// non-empty segment starts inside [pc, pc + 64) writes its id at its start in
// the wave's 64 LDS slots, and a prefix maximum carries it forward.
__device__ __forceinline__ int32_t chunk_owner(int32_t incl, int32_t len, int32_t pc, uint32_t lane,
                                               int32_t* __restrict__ slot) {
    const int32_t first = static_cast<int32_t>(__popcll(__ballot(incl <= pc)));
    slot[lane] = -1;
    wave_lds_sync();
    const int32_t s = incl - len - pc;
    if (len > 0 && s >= 0 && s < 64) slot[s] = static_cast<int32_t>(lane);
    wave_lds_sync();
    const int32_t v = wave_incl_max(slot[lane]);
    wave_lds_sync();
    return v > first ? v : first;
}

// ------------------------------------------------------------------ forward

__device__ __forceinline__ bool load_corners(const ViewDesc& vd, const uint32_t* __restrict__ idx, uint32_t t,
                                             uint32_t V, float4& p0, float4& p1, float4& p2) {
    const uint32_t i0 = idx[t * 3 + 0], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
    if (i0 >= V || i1 >= V || i2 >= V) return false;
    const float4* P = reinterpret_cast<const float4*>(vd.pos);
    p0 = P[i0];
    p1 = P[i1];
    p2 = P[i2];
    return true;
}

__device__ __forceinline__ bool setup_forward_tri(const ViewDesc& vd, const uint32_t* __restrict__ idx, uint32_t t,
                                                  uint32_t V, uint32_t cull, bool allow_cols, ScreenTri& r,
                                                  LineSpans& sp) {
    float4 p0, p1, p2;
    if (!load_corners(vd, idx, t, V, p0, p1, p2)) return false;
    return setup_screen(vd, p0, p1, p2, cull, allow_cols, r, sp);
}

// The triangles of a block's four 64-triangle batches in LDS, one array per
// field group (lanes reading different triangles hit different banks).
struct FwdRecs {
    float4 p0[kBlock], p1[kBlock], p2[kBlock];
    float4 sa[kBlock];  // s0x, s0y, s1x, s1y
    float4 sb[kBlock];  // s2x, s2y, sgn, triangle index bits
    float4 su[kBlock];  // span coefficients u0..u2, lower-bound mask bits
    float4 sv[kBlock];  // span coefficients v0..v2, upper-bound mask bits
    int4 ln[kBlock];    // span range along the lines a0, a1; line base; lines are columns
    int32_t slot[kBlock];
};

struct RecTri {  // what test_pixel reads
    float4 p0, p1, p2;
    float s0x, s0y, s1x, s1y, s2x, s2y, sgn;
};

// Stores the record of a set-up triangle (lane i of the block) whose lines
// [line_lo, ...] carry spans inside [span_lo, span_hi].
__device__ __forceinline__ void put_record(FwdRecs& R, uint32_t i, const ScreenTri& r, const LineSpans& sp, uint32_t t,
                                           int32_t span_lo, int32_t span_hi, int32_t line_lo) {
    R.p0[i] = r.p0;
    R.p1[i] = r.p1;
    R.p2[i] = r.p2;
    R.sa[i] = make_float4(r.s0x, r.s0y, r.s1x, r.s1y);
    R.sb[i] = make_float4(r.s2x, r.s2y, r.sgn, __uint_as_float(t));
    R.su[i] = make_float4(sp.u[0], sp.u[1], sp.u[2], __uint_as_float(sp.lower));
    R.sv[i] = make_float4(sp.v[0], sp.v[1], sp.v[2], __uint_as_float(sp.upper));
    R.ln[i] = make_int4(span_lo, span_hi, line_lo, static_cast<int32_t>(sp.cols));
}

// The walk of one wave's batch of up to 64 triangles (records w0 .. w0 + 63 of
// R; lane `lane` owns `lines` lines of its triangle): the lines of the batch
// are concatenated and dealt 64 at a time to the lanes, each lane computing its
// line's span; the spans of those 64 lines are concatenated in turn and dealt
// 64 pixels at a time, so every lane tests one candidate pixel per step
// whatever the triangles' sizes.  sink(x, y, key) receives every covered
// candidate's (depth, id) key.
template <class Sink>
__device__ __forceinline__ void walk_batch(const ViewDesc& vd, FwdRecs& R, uint32_t w0, uint32_t lane, int32_t lines,
                                           int32_t* slot, const float ax[2], const float bx[2], Sink&& sink) {
    const int32_t incl_l = wave_incl_scan(lines);
    const int32_t nlines = __builtin_amdgcn_readlane(incl_l, 63);
    if (lines > 0) R.ln[w0 + lane].z -= incl_l - lines;  // line coordinate = base + list position
    wave_lds_sync();
    for (int32_t lc = 0; lc < nlines; lc += 64) {
        const int32_t f = lc + static_cast<int32_t>(lane);
        const int32_t o = chunk_owner(incl_l, lines, lc, lane, slot);
        int32_t lo = 0, hi = -1, line = 0, cols = 0;
        float sl = 0.0f;
        if (f < nlines) {
            const uint32_t i = w0 + static_cast<uint32_t>(o);
            const int4 ln = R.ln[i];
            const float4 su = R.su[i], sv = R.sv[i];
            LineSpans sp;
            sp.u[0] = su.x;
            sp.u[1] = su.y;
            sp.u[2] = su.z;
            sp.lower = __float_as_uint(su.w);
            sp.v[0] = sv.x;
            sp.v[1] = sv.y;
            sp.v[2] = sv.z;
            sp.upper = __float_as_uint(sv.w);
            sp.cols = static_cast<uint32_t>(ln.w);
            cols = ln.w;
            line = ln.z + f;
            sl = cols ? sample_x(vd, line) : sample_y(vd, line);
            line_span(sp, sl, ax[cols], bx[cols], ln.x, ln.y, &lo, &hi);
        }
        const int32_t len = hi >= lo ? hi - lo + 1 : 0;
        const int32_t incl_p = wave_incl_scan(len);
        const int32_t npix = __builtin_amdgcn_readlane(incl_p, 63);
        const int32_t aoff = lo - (incl_p - len);
        for (int32_t pc = 0; pc < npix; pc += 64) {
            const int32_t m = pc + static_cast<int32_t>(lane);
            const int32_t k = chunk_owner(incl_p, len, pc, lane, slot);
            const int32_t a = __shfl(aoff, k, 64) + m;
            const int32_t ln = __shfl(line, k, 64);
            const int32_t oo = __shfl(o, k, 64);
            const int32_t cc = __shfl(cols, k, 64);
            const float sln = __shfl(sl, k, 64);
            if (m < npix) {
                const uint32_t i = w0 + static_cast<uint32_t>(oo);
                RecTri q;
                q.p0 = R.p0[i];
                q.p1 = R.p1[i];
                q.p2 = R.p2[i];
                const float4 sa = R.sa[i], sb = R.sb[i];
                q.s0x = sa.x;
                q.s0y = sa.y;
                q.s1x = sa.z;
                q.s1y = sa.w;
                q.s2x = sb.x;
                q.s2y = sb.y;
                q.sgn = sb.z;
                const int32_t x = cc ? ln : a, y = cc ? a : ln;
                const float sx = cc ? sln : sample_x(vd, x), sy = cc ? sample_y(vd, y) : sln;
                const unsigned long long key = test_pixel(q, __float_as_uint(sb.w), sx, sy);
                if (key != ~0ull) sink(x, y, key);
            }
        }
    }
    wave_lds_sync();
}

// One batch of up to 64 triangles per wave (walk_batch), with the keys kept
// by a global atomicMin per pixel.  Triangles above wave_work candidates are
// queued in 16-row slabs for k_raster_large.  With a list, the view's
// triangles are list[view * T + i], i < list_count[view] (else every triangle).
// Traversal order (without a list): order 0 takes the triangles in index
// order; order 1 starts from the end of the index range nearer the camera
// (first vs last triangle's first vertex clip w) in each view, so that with a
// mesh stored in spatial order (a grid) the near geometry lands first and the
// key filter turns away more of the far geometry's atomics.  The winner does
// not depend on the order.
__global__ __launch_bounds__(kBlock) void k_raster_small(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, uint32_t cull, unsigned long long* __restrict__ keys,
                                                         uint4* __restrict__ queue, uint32_t* __restrict__ counters,
                                                         uint32_t bsz, float wave_work,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ list_count, uint32_t order) {
    __shared__ FwdRecs R;
    const uint32_t view = blockIdx.y;
    const uint32_t count = list ? list_count[view] : T;
    const uint32_t* vlist = list ? list + static_cast<size_t>(view) * T : nullptr;
    const ViewDesc& vd = vt.v[view];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t w0 = wave * 64;
    int32_t* slot = R.slot + w0;
    unsigned long long* kbase = keys + vd.key_base;
    const float ax[2] = {span_alpha(vd, 0), span_alpha(vd, 1)};
    const float bx[2] = {span_beta(vd, 0), span_beta(vd, 1)};
    const uint32_t stride = gridDim.x * (kBlock / 64) * bsz;
    // each lane's corners are loaded one batch ahead, so the index and position
    // loads of the next batch are in flight while this one is rasterised
    float4 c0, c1, c2;
    bool cvalid = false;
    uint32_t ctri = 0;
    bool rev = false;
    if (order == 1u && !vlist && count > 1u) {
        const uint32_t ia = idx[0], ib = idx[static_cast<size_t>(count - 1u) * 3];
        const float4* P = reinterpret_cast<const float4*>(vd.pos);
        rev = ia < V && ib < V && P[ib].w < P[ia].w;
    }
    auto fetch = [&](uint32_t b) {
        const uint32_t i = b + lane;
        ctri = vlist && lane < bsz && b < count && i < count ? vlist[i] : (rev && i < count ? count - 1u - i : i);
        cvalid = lane < bsz && b < count && i < count && load_corners(vd, idx, ctri, V, c0, c1, c2);
    };
    uint32_t batch = (blockIdx.x * (kBlock / 64) + wave) * bsz;
    fetch(batch);
    for (; batch < count; batch += stride) {
        const uint32_t t = ctri;
        const float4 q0 = c0, q1 = c1, q2 = c2;
        const bool have = cvalid;
        fetch(batch + stride);
        int32_t lines = 0;
        {
            ScreenTri r;
            LineSpans sp;
            const bool ok = have && setup_screen(vd, q0, q1, q2, cull, true, r, sp);
            if (ok && r.work > wave_work) {
                const uint32_t slabs = static_cast<uint32_t>(r.y1 - r.y0) / kSlabRows + 1u;
                queue[atomicAdd(&counters[0], 1u)] = make_uint4(view, t, slabs, 0u);
            } else if (ok) {
                lines = sp.cols ? r.x1 - r.x0 + 1 : r.y1 - r.y0 + 1;
                put_record(R, w0 + lane, r, sp, t, sp.cols ? r.y0 : r.x0, sp.cols ? r.y1 : r.x1,
                           sp.cols ? r.x0 : r.y0);
            }
        }
        walk_batch(vd, R, w0, lane, lines, slot, ax, bx, [&](int32_t x, int32_t y, unsigned long long key) {
            key_min(&kbase[key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y), vd.key_tx)], key);
        });
    }
}

// Rows [ya, yb] of a wave-uniform triangle with row spans, 64 rows at a time.
__device__ __forceinline__ void wave_forward(const ScreenTri& r, const LineSpans& sp, uint32_t t, int32_t ya,
                                             int32_t yb, const ViewDesc& vd, uint32_t lane, int32_t* slot,
                                             unsigned long long* __restrict__ kbase) {
    const float alpha = span_alpha(vd, 0), beta = span_beta(vd, 0);
    for (int32_t y0 = ya; y0 <= yb; y0 += 64) {
        const int32_t y = y0 + static_cast<int32_t>(lane);
        int32_t lo = 0, hi = -1;
        float sy = 0.0f;
        if (y <= yb) {
            sy = sample_y(vd, y);
            line_span(sp, sy, alpha, beta, r.x0, r.x1, &lo, &hi);
        }
        const int32_t len = hi >= lo ? hi - lo + 1 : 0;
        const int32_t incl = wave_incl_scan(len);
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t xoff = lo - (incl - len);
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t m = base + static_cast<int32_t>(lane);
            const int32_t rl = chunk_owner(incl, len, base, lane, slot);
            const int32_t x = __shfl(xoff, rl, 64) + m;
            const float syr = __shfl(sy, rl, 64);
            if (m < total) {
                const unsigned long long key = test_pixel(r, t, sample_x(vd, x), syr);
                if (key != ~0ull)
                    key_min(&kbase[key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y0 + rl), vd.key_tx)], key);
            }
        }
    }
}

// Queued triangles, split into 16-row slabs dealt round-robin to every wave of
// the grid (slab k of the concatenated slab list goes to wave k mod waves).
__global__ __launch_bounds__(kBlock) void k_raster_large(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, uint32_t cull, unsigned long long* __restrict__ keys,
                                                         const uint4* __restrict__ queue,
                                                         const uint32_t* __restrict__ counters) {
    __shared__ uint32_t nslab[kBlock];
    __shared__ unsigned long long before[kBlock];
    __shared__ unsigned long long batch_total;
    __shared__ int32_t slots[kBlock];
    const uint32_t count = counters[0];
    const uint32_t waves = gridDim.x * (kBlock / 64);
    const uint32_t wid = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    unsigned long long cum = 0;
    for (uint32_t b0 = 0; b0 < count; b0 += kBlock) {
        nslab[threadIdx.x] = b0 + threadIdx.x < count ? queue[b0 + threadIdx.x].z : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long s = 0;
            for (uint32_t i = 0; i < kBlock; ++i) {
                before[i] = s;
                s += nslab[i];
            }
            batch_total = s;
        }
        __syncthreads();
        const uint32_t n_in = count - b0 < kBlock ? count - b0 : kBlock;
        for (uint32_t j = 0; j < n_in; ++j) {
            const uint32_t n = nslab[j];
            const uint32_t first = static_cast<uint32_t>((cum + before[j]) % waves);
            uint32_t s = wid >= first ? wid - first : wid + waves - first;
            if (s >= n) continue;
            const uint4 q = queue[b0 + j];
            const ViewDesc& vd = vt.v[q.x];
            ScreenTri r;
            LineSpans sp;
            if (!setup_forward_tri(vd, idx, q.y, V, cull, false, r, sp)) continue;
            for (; s < n; s += waves) {
                const int32_t ya = r.y0 + static_cast<int32_t>(s * kSlabRows);
                const int32_t yb = ya + static_cast<int32_t>(kSlabRows) - 1 < r.y1 ? ya + static_cast<int32_t>(kSlabRows) - 1
                                                                                  : r.y1;
                wave_forward(r, sp, q.y, ya, yb, vd, lane, slots + (threadIdx.x & ~63u), keys + vd.key_base);
            }
        }
        cum += batch_total;
        __syncthreads();
    }
}

// Winner's outputs of pixel p with sample (sx, sy) from its key
// (rasterize_forward.hlsl:115-162, 165-174): raster and, with want_deriv and a
// derivative output, the barycentric derivatives (zeros where key is ~0).
__device__ __forceinline__ void resolve_key(const ViewDesc& vd, const uint32_t* __restrict__ idx, int want_deriv,
                                            unsigned long long key, uint32_t p, float sx, float sy) {
        float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (key != ~0ull) {
            const uint32_t t = static_cast<uint32_t>(key & 0xFFFFFFFFull) - 1u;
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
        // nontemporal stores: the outputs stream past the caches instead of
        // displacing lines the next kernel reads (resolve + interpolate forward
        // 0.455 -> 0.432 ms on 15.2 M px, with the interpolate's own output
        // also streamed)
        typedef float nt4 __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(nt4{r.x, r.y, r.z, r.w}, reinterpret_cast<nt4*>(vd.raster) + p);
        if (vd.deriv) __builtin_nontemporal_store(nt4{d.x, d.y, d.z, d.w}, reinterpret_cast<nt4*>(vd.deriv) + p);
}

// Winner's outputs of every pixel from the global keys; resets the keys.  One
// lane per key, in key order (a wave covers four 4 x 4 tiles of one view: its
// key reads are one contiguous 512-byte run and each output row it writes is
// 16 pixels long); keys past the view's last row or column are skipped.
__global__ __launch_bounds__(kBlock) void k_raster_resolve(ViewTable vt, const uint32_t* __restrict__ idx,
                                                           int want_deriv, unsigned long long* __restrict__ keys,
                                                           uint32_t* __restrict__ counters) {
    clear_queue_counters(counters);  // k_raster_large, before this pass, was the queue's consumer
    const uint32_t gk = blockIdx.x * kBlock + threadIdx.x;
    if (gk >= vt.total_keys) return;
    // views start at multiples of 64 keys: the wave's first key names its view
    const uint32_t first = __builtin_amdgcn_readfirstlane(gk);
    uint32_t view = 0;
    while (view + 1 < vt.n && first >= vt.v[view + 1].key_base) ++view;
    const ViewDesc& vd = vt.v[view];
    const uint32_t k = gk - vd.key_base, tile = k >> 4;
    const uint32_t x = (tile % vd.key_tx) * kKeyTile + (k & 3u);
    const uint32_t y = (tile / vd.key_tx) * kKeyTile + ((k >> 2) & 3u);
    if (x >= vd.width || y >= vd.height) return;
    const unsigned long long key = keys[gk];
    keys[gk] = ~0ull;
    resolve_key(vd, idx, want_deriv, key, y * vd.width + x, sample_x(vd, static_cast<int32_t>(x)),
                sample_y(vd, static_cast<int32_t>(y)));
}

// ------------------------------------------------------------ tiled forward
//
// k_bin_count    per (view, triangle): the forward setup; a triangle whose
//                candidate rectangle touches at most bin_tiles raster tiles is
//                counted in each of them, a larger one is queued in 16-row
//                slabs for k_raster_large (global keys) and its tiles marked
//                as seeded;
// (scan)         bin offsets = exclusive sum of the counts;
// k_bin_fill     the same setup, each binned triangle's id into its tiles'
//                lists (counting the counts back down to zero);
// k_raster_large the slab triangles' keys (global atomicMin);
// k_tile_raster  one workgroup per raster tile: 1,024 LDS keys (seeded from
//                the global keys where slab triangles reach the tile), the
//                tile's triangles walked as the triangle-parallel path walks
//                them but with LDS atomicMin and every line clamped to the
//                tile, then every pixel of the tile resolved from LDS.
// The candidate set of every triangle is the one of the triangle-parallel
// path, split between tiles, and the (depth, id) minimum does not depend on
// order: raster and derivatives are the same bit for bit.
constexpr int32_t kRT = static_cast<int32_t>(kRasterTile);
constexpr uint32_t kRTPix = kRasterTile * kRasterTile;

__device__ __forceinline__ uint32_t rtile_count(const ScreenTri& r) {
    return static_cast<uint32_t>(r.x1 / kRT - r.x0 / kRT + 1) * static_cast<uint32_t>(r.y1 / kRT - r.y0 / kRT + 1);
}

// Block-level aggregation of per-tile increments: the block's entries (tile
// keys) are counted in an LDS hash table (open addressing, kBinSlots slots);
// each slot then takes one global atomic.  Neighbouring triangles share tiles,
// and same-address global atomics serialise.  An entry that finds no free slot
// (more distinct tiles in a block than the table holds, which needs very large
// triangles) goes to global memory on its own.
constexpr uint32_t kBinSlots = 1024;
struct BinHash {
    uint32_t key[kBinSlots];
    uint32_t count[kBinSlots];
    uint32_t base[kBinSlots];
};

__device__ __forceinline__ void bin_hash_clear(BinHash& h) {
    for (uint32_t i = threadIdx.x; i < kBinSlots; i += kBlock) {
        h.key[i] = ~0u;
        h.count[i] = 0u;
    }
}

// Inserts tile `key`: returns its slot (or kBinSlots when the table is full)
// and the entry's rank among the block's entries of that tile.
__device__ __forceinline__ uint32_t bin_hash_add(BinHash& h, uint32_t key, uint32_t* rank) {
    uint32_t slot = (key * 2654435761u) >> 22;  // top 10 bits: kBinSlots = 1024
    for (uint32_t probe = 0; probe < kBinSlots; ++probe, slot = (slot + 1u) & (kBinSlots - 1u)) {
        const uint32_t prev = atomicCAS(&h.key[slot], ~0u, key);
        if (prev == ~0u || prev == key) {
            *rank = atomicAdd(&h.count[slot], 1u);
            return slot;
        }
    }
    return kBinSlots;
}


constexpr uint32_t kMaxBinTiles = 8;  // largest bin depth (tiles per binned triangle)

// Raster tile k (row-major over the triangle's candidate rectangle) of r.
__device__ __forceinline__ uint32_t rtile_of(const ViewDesc& vd, const ScreenTri& r, uint32_t k) {
    const uint32_t nx = static_cast<uint32_t>(r.x1 / kRT - r.x0 / kRT + 1);
    return vd.rtile_base + (static_cast<uint32_t>(r.y0 / kRT) + k / nx) * vd.rtiles_x +
           static_cast<uint32_t>(r.x0 / kRT) + k % nx;
}

// A binned triangle's forward setup as k_tile_raster needs it (written once by
// k_bin_count): corners, screen coordinates and winding, spans (lower / upper
// bits, columns flag in bit 8 of lower) and the candidate rectangle.
struct TriRecord {
    float4 p0, p1, p2;
    float4 sa;  // s0x, s0y, s1x, s1y
    float4 sb;  // s2x, s2y, sgn, 0
    float4 su;  // u0..u2, lower | cols << 8 (bits)
    float4 sv;  // v0..v2, upper (bits)
    int4 rect;  // x0, x1, y0, y1
};

// Tiles of a binned triangle for the fill pass: first tile x, y (16 bits each
// in the low word), tiles across and count (high word); 0 when not binned.
__device__ __forceinline__ unsigned long long tile_span_word(const ScreenTri& r, uint32_t n) {
    const uint32_t nx = static_cast<uint32_t>(r.x1 / kRT - r.x0 / kRT + 1);
    return (static_cast<unsigned long long>((nx << 16) | n) << 32) |
           static_cast<unsigned long long>((static_cast<uint32_t>(r.y0 / kRT) << 16) |
                                           static_cast<uint32_t>(r.x0 / kRT));
}

// Binning pass 1 (grid.y = view; every lane of a block works in one view): the
// forward setup of each triangle, stored as its record.  A triangle touching
// at most bin_tiles raster tiles is counted in each (block LDS hash, then one
// atomic per distinct tile); a larger one goes to its view's list for the
// global-key walk (k_raster_small) and marks its tiles as seeded.
__global__ __launch_bounds__(kBlock) void k_bin_count(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                      uint32_t V, uint32_t cull, uint32_t bin_tiles,
                                                      uint32_t* __restrict__ counts, uint8_t* __restrict__ seeded,
                                                      uint32_t* __restrict__ mlist, uint32_t* __restrict__ mcount,
                                                      TriRecord* __restrict__ recs,
                                                      unsigned long long* __restrict__ spans) {
    __shared__ BinHash h;
    __shared__ uint32_t medium_n, medium_base;
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    for (uint32_t t0 = blockIdx.x * kBlock; t0 < T; t0 += gridDim.x * kBlock) {  // block-uniform
        bin_hash_clear(h);
        if (threadIdx.x == 0) medium_n = 0u;
        __syncthreads();
        const uint32_t t = t0 + threadIdx.x;
        ScreenTri r;
        LineSpans sp;
        const bool ok = t < T && setup_forward_tri(vd, idx, t, V, cull, true, r, sp);
        const uint32_t n = ok ? rtile_count(r) : 0u;
        const bool binned = ok && n <= bin_tiles, medium = ok && n > bin_tiles;
        if (t < T) {
            const size_t slot = static_cast<size_t>(view) * T + t;
            spans[slot] = binned ? tile_span_word(r, n) : 0ull;
            if (binned) {
                TriRecord rec;
                rec.p0 = r.p0;
                rec.p1 = r.p1;
                rec.p2 = r.p2;
                rec.sa = make_float4(r.s0x, r.s0y, r.s1x, r.s1y);
                rec.sb = make_float4(r.s2x, r.s2y, r.sgn, 0.0f);
                rec.su = make_float4(sp.u[0], sp.u[1], sp.u[2], __uint_as_float(sp.lower | (sp.cols << 8)));
                rec.sv = make_float4(sp.v[0], sp.v[1], sp.v[2], __uint_as_float(sp.upper));
                rec.rect = make_int4(r.x0, r.x1, r.y0, r.y1);
                recs[slot] = rec;
            }
        }
        uint32_t mrank = 0;
        if (medium) mrank = atomicAdd(&medium_n, 1u);
        for (uint32_t k = 0; binned && k < n; ++k) {
            const uint32_t tile = rtile_of(vd, r, k);
            uint32_t rank;
            if (bin_hash_add(h, tile, &rank) == kBinSlots) atomicAdd(&counts[tile], 1u);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kBinSlots; i += kBlock)
            if (h.key[i] != ~0u) atomicAdd(&counts[h.key[i]], h.count[i]);
        if (threadIdx.x == 0 && medium_n) medium_base = atomicAdd(&mcount[view], medium_n);
        __syncthreads();
        if (medium) {
            mlist[static_cast<size_t>(view) * T + medium_base + mrank] = t;
            for (uint32_t k = 0; k < n; ++k) seeded[rtile_of(vd, r, k)] = 1u;
        }
        __syncthreads();
    }
}

// Binning pass 2: each binned triangle's id into its tiles' lists at
// offsets[tile] + (a block base taken per distinct tile, counting the counts
// back down to zero) + the entry's rank within the block.
__global__ __launch_bounds__(kBlock) void k_bin_fill(ViewTable vt, uint32_t T, uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offsets, uint32_t* __restrict__ list,
                                                     const unsigned long long* __restrict__ spans) {
    __shared__ BinHash h;
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    for (uint32_t t0 = blockIdx.x * kBlock; t0 < T; t0 += gridDim.x * kBlock) {  // block-uniform
        bin_hash_clear(h);
        __syncthreads();
        const uint32_t t = t0 + threadIdx.x;
        const unsigned long long w = t < T ? spans[static_cast<size_t>(view) * T + t] : 0ull;
        const uint32_t hi = static_cast<uint32_t>(w >> 32), lo = static_cast<uint32_t>(w);
        const uint32_t n = hi & 0xFFFFu, nx = hi > 0xFFFFu ? hi >> 16 : 1u;
        uint32_t slot_rank[kMaxBinTiles];  // (slot << 16) | rank of tile k, ~0 when placed on its own
#pragma unroll
        for (uint32_t k = 0; k < kMaxBinTiles; ++k) {
            slot_rank[k] = ~0u;
            if (k < n) {
                const uint32_t tile = vd.rtile_base + ((lo >> 16) + k / nx) * vd.rtiles_x + (lo & 0xFFFFu) + k % nx;
                uint32_t rank = 0;
                const uint32_t slot = bin_hash_add(h, tile, &rank);
                if (slot == kBinSlots)
                    list[offsets[tile] + atomicSub(&counts[tile], 1u) - 1u] = t;  // table full: on its own
                else
                    slot_rank[k] = (slot << 16) | rank;
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kBinSlots; i += kBlock)
            if (h.key[i] != ~0u) h.base[i] = atomicSub(&counts[h.key[i]], h.count[i]) - h.count[i];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kMaxBinTiles; ++k) {
            const uint32_t sr = slot_rank[k];
            if (sr == ~0u) continue;
            const uint32_t slot = sr >> 16;
            list[offsets[h.key[slot]] + h.base[slot] + (sr & 0xFFFFu)] = t;
        }
        __syncthreads();
    }
}

// One workgroup per raster tile: LDS keys (seeded from the global keys where
// unbinned triangles reach the tile), the tile's triangles from their records
// walked with LDS atomicMin and every line clamped to the tile, then every
// pixel of the tile resolved from LDS.
__global__ __launch_bounds__(kBlock) void k_tile_raster(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                        int want_deriv, const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ list,
                                                        const uint8_t* __restrict__ seeded,
                                                        const TriRecord* __restrict__ recs,
                                                        unsigned long long* __restrict__ keys,
                                                        uint32_t* __restrict__ counters) {
    __shared__ unsigned long long lk[kRTPix];
    __shared__ FwdRecs R;
    clear_queue_counters(counters);  // k_raster_large, before this pass, was the queue's consumer
    const uint32_t tile = blockIdx.x;
    uint32_t view = 0;
    while (view + 1 < vt.n && tile >= vt.v[view + 1].rtile_base) ++view;
    const ViewDesc& vd = vt.v[view];
    const uint32_t lt = tile - vd.rtile_base;
    const int32_t tx0 = static_cast<int32_t>(lt % vd.rtiles_x) * kRT, ty0 = static_cast<int32_t>(lt / vd.rtiles_x) * kRT;
    const int32_t W = static_cast<int32_t>(vd.width), H = static_cast<int32_t>(vd.height);
    const int32_t tx1 = tx0 + kRT - 1 < W - 1 ? tx0 + kRT - 1 : W - 1;
    const int32_t ty1 = ty0 + kRT - 1 < H - 1 ? ty0 + kRT - 1 : H - 1;
    const bool seed = seeded[tile] != 0u;
    unsigned long long* gk = keys + vd.key_base;
    for (uint32_t i = threadIdx.x; i < kRTPix; i += kBlock) {
        const int32_t x = tx0 + static_cast<int32_t>(i % kRasterTile), y = ty0 + static_cast<int32_t>(i / kRasterTile);
        lk[i] = seed && x <= tx1 && y <= ty1
                    ? gk[key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y), vd.key_tx)]
                    : ~0ull;
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t w0 = wave * 64;
    int32_t* slot = R.slot + w0;
    const float ax[2] = {span_alpha(vd, 0), span_alpha(vd, 1)};
    const float bx[2] = {span_beta(vd, 0), span_beta(vd, 1)};
    const uint32_t e0 = offsets[tile], e1 = offsets[tile + 1];
    for (uint32_t b = e0 + w0; b < e1; b += kBlock) {  // wave-uniform: the waves take batches of 64 entries
        int32_t lines = 0;
        if (b + lane < e1) {
            const uint32_t t = list[b + lane];
            const TriRecord& rec = recs[static_cast<size_t>(view) * T + t];
            const int4 rc = rec.rect;
            const int32_t xa = rc.x > tx0 ? rc.x : tx0, xb = rc.y < tx1 ? rc.y : tx1;
            const int32_t ya = rc.z > ty0 ? rc.z : ty0, yb = rc.w < ty1 ? rc.w : ty1;
            if (xa <= xb && ya <= yb) {
                const float4 su = rec.su;
                const uint32_t bits = __float_as_uint(su.w), cols = bits >> 8;
                lines = cols ? xb - xa + 1 : yb - ya + 1;
                const uint32_t i = w0 + lane;
                R.p0[i] = rec.p0;
                R.p1[i] = rec.p1;
                R.p2[i] = rec.p2;
                R.sa[i] = rec.sa;
                const float4 sb = rec.sb;
                R.sb[i] = make_float4(sb.x, sb.y, sb.z, __uint_as_float(t));
                R.su[i] = make_float4(su.x, su.y, su.z, __uint_as_float(bits & 0xFFu));
                R.sv[i] = rec.sv;
                R.ln[i] = make_int4(cols ? ya : xa, cols ? yb : xb, cols ? xa : ya, static_cast<int32_t>(cols));
            }
        }
        walk_batch(vd, R, w0, lane, lines, slot, ax, bx, [&](int32_t x, int32_t y, unsigned long long key) {
            unsigned long long* k = &lk[(y - ty0) * kRT + (x - tx0)];
            if (key < *k) atomicMin(k, key);
        });
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kRTPix; i += kBlock) {
        const int32_t x = tx0 + static_cast<int32_t>(i % kRasterTile), y = ty0 + static_cast<int32_t>(i / kRasterTile);
        if (x > tx1 || y > ty1) continue;
        const uint32_t p = static_cast<uint32_t>(y) * vd.width + static_cast<uint32_t>(x);
        if (seed) gk[key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y), vd.key_tx)] = ~0ull;
        resolve_key(vd, idx, want_deriv, lk[i], p, sample_x(vd, x), sample_y(vd, y));
    }
}

// ----------------------------------------------------------------- backward

// Per pixel: the id a triangle's backward loop compares against
// (rasterize_backward.hlsl:96-100: decoded id, 0 when grad_raster.xy is zero).
// The backward visits only the forward candidate region of each triangle
// (rows and spans of raster_geom.hpp) inside its bounding box; a pixel that
// names triangle t, lies in t's box and is outside that region (which the
// forward never produces, but a caller-supplied raster can) marks its view,
// and the view's triangles are then walked over their whole boxes.  A pixel
// that passes t's float edge test is inside t's region, which skips the
// region computation for every pixel of a forward-produced raster.
//
// The compared ids and the pixels' grad_raster.xy go to cid / cgrad in the
// keys' 4 x 4-tile layout (key_base + key_index; cgrad only where the id is
// not 0), so the corner walks, which visit a few pixels on each of a few rows
// per triangle, load from few cache lines.  One lane per key, in key order.
__global__ __launch_bounds__(kBlock) void k_bwd_prepass(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                        uint32_t V, uint32_t* __restrict__ cid,
                                                        float2* __restrict__ cgrad, uint32_t* __restrict__ view_flags,
                                                        uint8_t* __restrict__ won) {
    const uint32_t gk = blockIdx.x * kBlock + threadIdx.x;
    if (gk >= vt.total_keys) return;  // whole waves: views start at multiples of 64 keys
    const uint32_t first = __builtin_amdgcn_readfirstlane(gk);
    uint32_t view = 0;
    while (view + 1 < vt.n && first >= vt.v[view + 1].key_base) ++view;
    const ViewDesc& vd = vt.v[view];
    const uint32_t k = gk - vd.key_base, tile = k >> 4;
    const uint32_t xu = (tile % vd.key_tx) * kKeyTile + (k & 3u);
    const uint32_t yu = (tile / vd.key_tx) * kKeyTile + ((k >> 2) & 3u);
    uint32_t c = 0;
    if (xu < vd.width && yu < vd.height) {  // keys past the last row or column name no pixel
        const size_t p = static_cast<size_t>(yu) * vd.width + xu;
        const uint32_t id = decode_id(vd.raster[p * 4 + 3]);
        const float2 lg = reinterpret_cast<const float2*>(vd.grad_raster)[p * 2];
        if (id != 0u && id <= T && !(lg.x == 0.0f && lg.y == 0.0f)) {
            c = id;
            cgrad[gk] = lg;
            float4 p0, p1, p2;
            int32_t bx0, bx1, by0, by1;
            const int32_t x = static_cast<int32_t>(xu), y = static_cast<int32_t>(yu);
            if (load_corners(vd, idx, id - 1u, V, p0, p1, p2) && backward_box(vd, p0, p1, p2, &bx0, &bx1, &by0, &by1) &&
                x >= bx0 && x < bx1 && y >= by0 && y < by1) {
                const float sx = sample_x(vd, x), sy = sample_y(vd, y);
                ScreenTri r;
                float bb[4];
                bool inside = screen_coords(p0, p1, p2, 0u, r, bb) && edge_pass(r, sx, sy);
                if (!inside) {
                    LineSpans sp;
                    if (setup_screen(vd, p0, p1, p2, 0u, false, r, sp) && y >= r.y0 && y <= r.y1) {
                        int32_t lo, hi;
                        line_span(sp, sy, span_alpha(vd, 0), span_beta(vd, 0), r.x0, r.x1, &lo, &hi);
                        inside = x >= lo && x <= hi;
                    }
                }
                if (!inside) atomicOr(&view_flags[view], 1u);
            }
        }
        cid[gk] = c;
    }
    // mark the named triangle once per run of equal neighbours in the wave
    // (every lane takes part, with no mark where no pixel is named)
    const uint32_t mark = c ? view * T + (c - 1u) : 0xFFFFFFFFu;
    const uint32_t prev = __shfl_up(mark, 1, 64);
    if (c && ((threadIdx.x & 63u) == 0 || prev != mark)) won[mark] = 1;
}

// Terms of one pixel (rasterize_backward.hlsl:101-151): false when the pixel's
// area is degenerate; t9 = (x, y, w) for corners 0..2.
__device__ __forceinline__ bool pixel_terms(const float4* p, float sx, float sy, float2 lg, float t9[9]) {
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

// Rows [ya, yb] and columns [xa, xb] of triangle t visited by the backward
// (inside the row spans unless the view is walked over whole boxes).
struct BwdRegion {
    float4 p[3];
    LineSpans sp;
    int32_t ya, yb, xa, xb;
    uint32_t box;  // bounding-box pixel count
    float work;    // estimated visited pixels + rows
};

__device__ __forceinline__ bool setup_backward(const ViewDesc& vd, const uint32_t* __restrict__ idx, uint32_t t,
                                               uint32_t V, bool full, BwdRegion& g) {
    int32_t bx0, bx1, by0, by1;
    if (!load_corners(vd, idx, t, V, g.p[0], g.p[1], g.p[2])) return false;
    if (!backward_box(vd, g.p[0], g.p[1], g.p[2], &bx0, &bx1, &by0, &by1) || bx1 <= bx0 || by1 <= by0) return false;
    g.box = static_cast<uint32_t>(bx1 - bx0) * static_cast<uint32_t>(by1 - by0);
    if (full) {
        g.ya = by0;
        g.yb = by1 - 1;
        g.xa = bx0;
        g.xb = bx1 - 1;
        g.sp.lower = g.sp.upper = g.sp.cols = 0u;
#pragma unroll
        for (int i = 0; i < 3; ++i) g.sp.u[i] = g.sp.v[i] = 0.0f;
        g.work = static_cast<float>(g.box);
        return true;
    }
    ScreenTri r;
    if (!setup_screen(vd, g.p[0], g.p[1], g.p[2], 0u, false, r, g.sp)) return false;
    g.ya = r.y0 > by0 ? r.y0 : by0;
    g.yb = r.y1 < by1 - 1 ? r.y1 : by1 - 1;
    g.xa = r.x0 > bx0 ? r.x0 : bx0;
    g.xb = r.x1 < bx1 - 1 ? r.x1 : bx1 - 1;
    g.work = r.work;
    return g.ya <= g.yb && g.xa <= g.xb;
}

// The flat-class triangles of a block's four batches in LDS (one array per
// field group), the per-wave owner slots and the per-wave lists of matched
// pixels' terms.
constexpr uint32_t kList = 64;  // matched pixels buffered per wave before the per-triangle fold

struct BwdRecs {
    float4 p0[kBlock], p1[kBlock], p2[kBlock];
    float4 su[kBlock];  // span coefficients u0..u2, lower-bound mask bits
    float4 sv[kBlock];  // span coefficients v0..v2, upper-bound mask bits
    int4 rg[kBlock];    // xa, xb, row base, unused
    int32_t slot[kBlock];
    int32_t seg_first[kBlock];  // per lane: first list entry of its triangle (-1: none)
    int32_t seg_end[kBlock];    // per lane: one past its triangle's last list entry
    float term[kBlock / 64][kList * 9];
};

// Lane `lane` adds, in list order, the buffered terms of its triangle
// (entries [seg_first[lane], seg_end[lane]) of the list; they are contiguous
// because the list is in batch order), then clears its segment.
__device__ __forceinline__ void fold_list(const float* __restrict__ lterm, int32_t* seg_first, int32_t* seg_end,
                                          uint32_t lane, float s[9]) {
    wave_lds_sync();
    const int32_t first = seg_first[lane];
    if (first >= 0) {
        const int32_t end = seg_end[lane];
        int32_t e = first;
        // entries are loaded four at a time (independent reads) and added in order
        for (; e + 4 <= end; e += 4) {
            float v[4][9];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int k = 0; k < 9; ++k) v[j][k] = lterm[(e + j) * 9 + k];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int k = 0; k < 9; ++k) s[k] += v[j][k];
        }
        for (; e < end; ++e)
#pragma unroll
            for (int k = 0; k < 9; ++k) s[k] += lterm[e * 9 + k];
    }
    seg_first[lane] = -1;
    wave_lds_sync();
}

// Per-triangle corner gradients summed over the reference's bounding box in
// its row-major order.  Each wave takes a batch of up to 64 triangles.  Triangles
// with more than flat_work estimated pixels (boxes up to kWaveBox pixels) are
// walked by the whole wave, one at a time, adding the contributing pixels'
// terms in ascending pixel order through ballot + readlane; larger boxes are
// queued for k_corner_large.  The rows of the other triangles are
// concatenated and walked 64 pixels at a time (as the forward); pixels naming
// their triangle compute their terms and append them, in walk order, to a
// per-wave list, which each triangle's lane then adds in order (the walk
// order restricted to one triangle is the reference's row-major order).
__global__ __launch_bounds__(kBlock, 4) void k_corner_small(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, const uint32_t* __restrict__ cid,
                                                         const float2* __restrict__ cgrad,
                                                         const uint32_t* __restrict__ view_flags,
                                                         const uint8_t* __restrict__ won,
                                                         const uint32_t* __restrict__ slot_of,
                                                         float4* __restrict__ corner, uint4* __restrict__ queue,
                                                         uint32_t* __restrict__ counters, uint32_t bsz,
                                                         float flat_work) {
    __shared__ BwdRecs R;
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t w0 = wave * 64;
    int32_t* slot = R.slot + w0;
    float* lterm = R.term[wave];
    const bool full = view_flags[view] != 0u;
    const float alpha = span_alpha(vd, 0), beta = span_beta(vd, 0);
    const uint32_t* cv = cid + vd.key_base;
    const float2* gv = cgrad + vd.key_base;
    float4* cview = corner + static_cast<size_t>(view) * T * 3;
    int32_t* seg_first = R.seg_first + w0;
    int32_t* seg_end = R.seg_end + w0;
    const uint32_t stride = gridDim.x * (kBlock / 64) * bsz;
    for (uint32_t batch = (blockIdx.x * (kBlock / 64) + wave) * bsz; batch < T; batch += stride) {
        const uint32_t t = batch + lane;
        const bool mine = lane < bsz && t < T;
        int32_t rows = 0;
        {
            BwdRegion g;
            // a triangle no pixel names adds nothing: its corners are zero without a walk
            const bool ok = mine && won[static_cast<size_t>(view) * T + t] && setup_backward(vd, idx, t, V, full, g);
            bool deferred = false;
            if (mine && !ok) {
#pragma unroll
                for (int c = 0; c < 3; ++c) cview[slot_of[t * 3 + c]] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            } else if (ok && g.box > kWaveBox) {
                queue[atomicAdd(&counters[0], 1u)] = make_uint4(view, t, 0u, 0u);
            } else if (ok && g.work > flat_work) {
                deferred = true;
            } else if (ok) {
                rows = g.yb - g.ya + 1;
                const uint32_t i = w0 + lane;
                R.p0[i] = g.p[0];
                R.p1[i] = g.p[1];
                R.p2[i] = g.p[2];
                R.su[i] = make_float4(g.sp.u[0], g.sp.u[1], g.sp.u[2], __uint_as_float(g.sp.lower));
                R.sv[i] = make_float4(g.sp.v[0], g.sp.v[1], g.sp.v[2], __uint_as_float(g.sp.upper));
                R.rg[i] = make_int4(g.xa, g.xb, g.ya, 0);
            }
            unsigned long long pending = __ballot(deferred);
            while (pending) {
                const int j = __ffsll(pending) - 1;
                pending &= pending - 1;
                BwdRegion q;
                q.p[0] = bcf4(g.p[0], j);
                q.p[1] = bcf4(g.p[1], j);
                q.p[2] = bcf4(g.p[2], j);
                bc_spans(g.sp, j, q.sp);
                q.ya = bci(g.ya, j);
                q.yb = bci(g.yb, j);
                q.xa = bci(g.xa, j);
                q.xb = bci(g.xb, j);
                const uint32_t qt = batch + static_cast<uint32_t>(j);
                float w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int32_t y0 = q.ya; y0 <= q.yb; y0 += 64) {
                    const int32_t y = y0 + static_cast<int32_t>(lane);
                    int32_t lo = 0, hi = -1;
                    float sy = 0.0f;
                    if (y <= q.yb) {
                        sy = sample_y(vd, y);
                        line_span(q.sp, sy, alpha, beta, q.xa, q.xb, &lo, &hi);
                    }
                    const int32_t len = hi >= lo ? hi - lo + 1 : 0;
                    const int32_t incl = wave_incl_scan(len);
                    const int32_t total = __builtin_amdgcn_readlane(incl, 63);
                    const int32_t xoff = lo - (incl - len);
                    for (int32_t base = 0; base < total; base += 64) {
                        const int32_t m = base + static_cast<int32_t>(lane);
                        const int32_t rl = chunk_owner(incl, len, base, lane, slot);
                        const int32_t x = __shfl(xoff, rl, 64) + m;
                        const float syr = __shfl(sy, rl, 64);
                        float t9[9];
                        bool hit = false;
                        if (m < total) {
                            // id and gradient loaded together: no dependent second load
                            const uint32_t pix =
                                key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y0 + rl), vd.key_tx);
                            const uint32_t c = cv[pix];
                            const float2 g = gv[pix];
                            if (c == qt + 1u) hit = pixel_terms(q.p, sample_x(vd, x), syr, g, t9);
                        }
                        unsigned long long mk = __ballot(hit);
                        while (mk) {
                            const int src = __ffsll(mk) - 1;
                            mk &= mk - 1;
#pragma unroll
                            for (int kk = 0; kk < 9; ++kk) w[kk] += bcf(t9[kk], src);
                        }
                    }
                }
                if (lane == 0)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        cview[slot_of[qt * 3 + c]] = make_float4(w[c * 3], w[c * 3 + 1], 0.0f, w[c * 3 + 2]);
            }
        }
        const int32_t incl_l = wave_incl_scan(rows);
        const int32_t nrows = __builtin_amdgcn_readlane(incl_l, 63);
        if (rows > 0) R.rg[w0 + lane].z -= incl_l - rows;  // row = base + list position
        float s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t nlist = 0;
        seg_first[lane] = -1;
        wave_lds_sync();
        for (int32_t lc = 0; lc < nrows; lc += 64) {
            const int32_t f = lc + static_cast<int32_t>(lane);
            const int32_t o = chunk_owner(incl_l, rows, lc, lane, slot);
            int32_t lo = 0, hi = -1, y = 0;
            float sy = 0.0f;
            if (f < nrows) {
                const uint32_t i = w0 + static_cast<uint32_t>(o);
                const int4 rg = R.rg[i];
                const float4 su = R.su[i], sv = R.sv[i];
                LineSpans sp;
                sp.u[0] = su.x;
                sp.u[1] = su.y;
                sp.u[2] = su.z;
                sp.lower = __float_as_uint(su.w);
                sp.v[0] = sv.x;
                sp.v[1] = sv.y;
                sp.v[2] = sv.z;
                sp.upper = __float_as_uint(sv.w);
                sp.cols = 0u;
                y = rg.z + f;
                sy = sample_y(vd, y);
                line_span(sp, sy, alpha, beta, rg.x, rg.y, &lo, &hi);
            }
            const int32_t len = hi >= lo ? hi - lo + 1 : 0;
            const int32_t incl_p = wave_incl_scan(len);
            const int32_t npix = __builtin_amdgcn_readlane(incl_p, 63);
            const int32_t xoff = lo - (incl_p - len);
            for (int32_t pc = 0; pc < npix; pc += 64) {
                const int32_t m = pc + static_cast<int32_t>(lane);
                const int32_t k = chunk_owner(incl_p, len, pc, lane, slot);
                const int32_t x = __shfl(xoff, k, 64) + m;
                const int32_t yy = __shfl(y, k, 64);
                const int32_t oo = __shfl(o, k, 64);
                const float syy = __shfl(sy, k, 64);
                float t9[9];
                bool hit = false;
                if (m < npix) {
                    // id and gradient loaded together: no dependent second load
                    const uint32_t pix = key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(yy), vd.key_tx);
                    const uint32_t c = cv[pix];
                    const float2 g = gv[pix];
                    if (c == batch + static_cast<uint32_t>(oo) + 1u) {  // oo < bsz: a flat-class lane
                        const uint32_t i = w0 + static_cast<uint32_t>(oo);
                        const float4 pp[3] = {R.p0[i], R.p1[i], R.p2[i]};
                        hit = pixel_terms(pp, sample_x(vd, x), syy, g, t9);
                    }
                }
                const unsigned long long bal = __ballot(hit);
                const uint32_t cnt = static_cast<uint32_t>(__popcll(bal));
                if (nlist + cnt > kList) {
                    fold_list(lterm, seg_first, seg_end, lane, s);
                    nlist = 0;
                }
                // neighbouring hits in lane order: a hit whose predecessor belongs to
                // another triangle starts its triangle's run, one whose successor does ends it
                const unsigned long long below = bal & ((1ull << lane) - 1ull);
                const unsigned long long above = bal & ~((2ull << lane) - 1ull);
                const int32_t prev = below ? 63 - __builtin_clzll(below) : static_cast<int32_t>(lane);
                const int32_t next = above ? __builtin_ctzll(above) : static_cast<int32_t>(lane);
                const int32_t prev_o = __shfl(oo, prev, 64), next_o = __shfl(oo, next, 64);
                if (hit) {
                    const uint32_t pos = nlist + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                                                           __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
#pragma unroll
                    for (int kk = 0; kk < 9; ++kk) lterm[pos * 9 + kk] = t9[kk];
                    if ((!below || prev_o != oo) && seg_first[oo] < 0) seg_first[oo] = static_cast<int32_t>(pos);
                    if (!above || next_o != oo) seg_end[oo] = static_cast<int32_t>(pos) + 1;
                }
                nlist += cnt;
            }
        }
        fold_list(lterm, seg_first, seg_end, lane, s);
        if (rows > 0)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                cview[slot_of[t * 3 + c]] = make_float4(s[c * 3], s[c * 3 + 1], 0.0f, s[c * 3 + 2]);
    }
}

// Boxes over kWaveBox pixels: each lane sums a strided subset of the box, then
// a fixed-order tree over the workgroup (deterministic; not the reference's
// summation order: the same terms, within 1e-4 of the largest gradient plus
// 1e-4 relative; k_corner_large_ordered keeps the order).
__global__ __launch_bounds__(kBlock) void k_corner_large(ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T,
                                                         uint32_t V, const uint32_t* __restrict__ cid,
                                                         const float2* __restrict__ cgrad,
                                                         const uint32_t* __restrict__ view_flags,
                                                         const uint32_t* __restrict__ slot_of,
                                                         float4* __restrict__ corner, const uint4* __restrict__ queue,
                                                         uint32_t* __restrict__ counters) {
    __shared__ uint32_t entry;
    __shared__ float red[9][kBlock];
    const uint32_t count = counters[0];
    if (blockIdx.x >= count) return;  // more workgroups than queued boxes
    while (true) {
        if (threadIdx.x == 0) entry = atomicAdd(&counters[1], 1u);
        __syncthreads();
        const uint32_t e = entry;
        __syncthreads();
        if (e >= count) return;
        const uint4 q = queue[e];
        const ViewDesc& vd = vt.v[q.x];
        BwdRegion g;
        const bool ok = setup_backward(vd, idx, q.y, V, view_flags[q.x] != 0u, g);
        const float alpha = span_alpha(vd, 0), beta = span_beta(vd, 0);
        const uint32_t* cv = cid + vd.key_base;
        const float2* gv = cgrad + vd.key_base;
        float s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int32_t y = g.ya + static_cast<int32_t>(threadIdx.x >> 6); ok && y <= g.yb; y += kBlock / 64) {
            const float sy = sample_y(vd, y);
            int32_t lo, hi;
            line_span(g.sp, sy, alpha, beta, g.xa, g.xb, &lo, &hi);
            for (int32_t x = lo + static_cast<int32_t>(threadIdx.x & 63u); x <= hi; x += 64) {
                const uint32_t pix = key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y), vd.key_tx);
                float t9[9];
                if (cv[pix] == q.y + 1u && pixel_terms(g.p, sample_x(vd, x), sy, gv[pix], t9))
#pragma unroll
                    for (int k = 0; k < 9; ++k) s[k] += t9[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = s[k];
        __syncthreads();
        for (uint32_t h = kBlock / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h)
#pragma unroll
                for (int k = 0; k < 9; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            float4* out = corner + static_cast<size_t>(q.x) * T * 3;
            for (int c = 0; c < 3; ++c)
                out[slot_of[q.y * 3 + c]] = make_float4(red[c * 3][0], red[c * 3 + 1][0], 0.0f, red[c * 3 + 2][0]);
        }
        __syncthreads();
    }
}

// Boxes over kWaveBox pixels in the reference's order (PDR_FLAG_DETERMINISTIC):
// one wave per queued box walks its rows in order, 64 pixels at a time; the
// chunk's contributing pixels put their 9 terms in LDS in pixel order (ballot
// rank), and lanes 0-8 each add one term component over that list in order,
// so every component is the left fold of rasterize_backward.hlsl:88-151 bit for bit.
__global__ __launch_bounds__(kBlock) void k_corner_large_ordered(
    ViewTable vt, const uint32_t* __restrict__ idx, uint32_t T, uint32_t V, const uint32_t* __restrict__ cid,
    const float2* __restrict__ cgrad, const uint32_t* __restrict__ view_flags,
    const uint32_t* __restrict__ slot_of, float4* __restrict__ corner, const uint4* __restrict__ queue,
    uint32_t* __restrict__ counters) {
    __shared__ float terms[kBlock / 64][64 * 9];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    float* lt = terms[wave];
    const uint32_t count = counters[0];
    while (true) {
        uint32_t e = 0;
        if (lane == 0) e = atomicAdd(&counters[1], 1u);
        e = __builtin_amdgcn_readfirstlane(e);
        if (e >= count) return;  // wave-uniform
        const uint4 q = queue[e];
        const ViewDesc& vd = vt.v[q.x];
        BwdRegion g;
        const bool ok = setup_backward(vd, idx, q.y, V, view_flags[q.x] != 0u, g);
        const float alpha = span_alpha(vd, 0), beta = span_beta(vd, 0);
        const uint32_t* cv = cid + vd.key_base;
        const float2* gv = cgrad + vd.key_base;
        float acc = 0.0f;  // lane k < 9: term component k
        for (int32_t y = g.ya; ok && y <= g.yb; ++y) {
            const float sy = sample_y(vd, y);
            int32_t lo, hi;
            line_span(g.sp, sy, alpha, beta, g.xa, g.xb, &lo, &hi);
            for (int32_t x0 = lo; x0 <= hi; x0 += 64) {
                const int32_t x = x0 + static_cast<int32_t>(lane);
                float t9[9];
                bool hit = false;
                if (x <= hi) {
                    const uint32_t pix = key_index(static_cast<uint32_t>(x), static_cast<uint32_t>(y), vd.key_tx);
                    hit = cv[pix] == q.y + 1u && pixel_terms(g.p, sample_x(vd, x), sy, gv[pix], t9);
                }
                const unsigned long long bal = __ballot(hit);
                if (!bal) continue;
                if (hit) {
                    const uint32_t pos = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
#pragma unroll
                    for (int k = 0; k < 9; ++k) lt[pos * 9 + k] = t9[k];
                }
                wave_lds_sync();
                const uint32_t cnt = static_cast<uint32_t>(__popcll(bal));
                if (lane < 9)
                    for (uint32_t j = 0; j < cnt; ++j) acc += lt[j * 9 + lane];
                wave_lds_sync();
            }
        }
        float w[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = bcf(acc, k);
        if (lane == 0) {
            float4* out = corner + static_cast<size_t>(q.x) * T * 3;
            for (int c = 0; c < 3; ++c)
                out[slot_of[q.y * 3 + c]] = make_float4(w[c * 3], w[c * 3 + 1], 0.0f, w[c * 3 + 2]);
        }
    }
}

// ---- Vertex -> corner adjacency, cached across calls.
//
// The adjacency depends on the index buffer alone, so the context keeps it
// with a copy of the indices it was built from.  Every call compares the
// indices with that copy on the device, and the build kernels exit at once
// unless a word differs, T or V changed, or nothing was built yet: a sequence
// of calls over one mesh builds it once, with no host readback (the calls stay
// capturable in a graph).  State words (kCsr*): the compare sets kCsrStale, the
// build's last block clears it and records T and V.  `counts` is all zeros
// between calls (the placement consumes each vertex's count, the last block
// clears the invalid-corner cursor); a newly allocated `counts` or `state` is
// zero-filled by the host.
enum : uint32_t { kCsrStale = 0, kCsrBuilt = 1, kCsrT = 2, kCsrV = 3, kCsrDone = 4 };
constexpr uint32_t kCsrLdsVerts = 16384;  // one-workgroup build: V + 1 counts in 64 KB of LDS
constexpr uint32_t kCsrScanTile = 2048;   // scan: 256 threads x 8 counts

// Ascending Shell sort (Ciura's gaps) of one vertex's corner list.
__device__ __forceinline__ void sort_corner_list(uint32_t* a, uint32_t n) {
    const uint32_t gaps[8] = {701u, 301u, 132u, 57u, 23u, 10u, 4u, 1u};
    for (int g = 0; g < 8; ++g) {
        const uint32_t gap = gaps[g];
        for (uint32_t i = gap; i < n; ++i) {
            const uint32_t x = a[i];
            uint32_t j = i;
            for (; j >= gap && a[j - gap] > x; j -= gap) a[j] = a[j - gap];
            a[j] = x;
        }
    }
}

// Compares this block's share of the indices with the saved copy; block 0
// also checks that an adjacency of this T and V was built.
__device__ __forceinline__ bool csr_differs(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ saved,
                                            uint32_t n, uint32_t T, uint32_t V, const uint32_t* state) {
    bool diff = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        diff |= idx[i] != saved[i];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        diff |= !(state[kCsrBuilt] == 1u && state[kCsrT] == T && state[kCsrV] == V);
    return diff;
}

// Meshes with V < kCsrLdsVerts: compare and build in one launch.  Every block
// compares its share; the last block to finish builds the whole adjacency in
// LDS when any block found a difference (counts, exclusive scan, placement
// with the indices saved, per-vertex sort).
__global__ __launch_bounds__(1024) void k_csr_small(const uint32_t* __restrict__ idx, uint32_t n, uint32_t T,
                                                    uint32_t V, uint32_t* state, uint32_t* __restrict__ saved,
                                                    uint32_t* __restrict__ offsets, uint32_t* __restrict__ corners,
                                                    uint32_t* __restrict__ slot_of) {
    __shared__ uint32_t cnt[kCsrLdsVerts];
    __shared__ uint32_t last, stale;
    using Scan = hipcub::BlockScan<uint32_t, 1024>;
    __shared__ typename Scan::TempStorage scan_tmp;
    const uint32_t tid = threadIdx.x;
    if (__syncthreads_or(csr_differs(idx, saved, n, T, V, state)) && tid == 0) atomicOr(&state[kCsrStale], 1u);
    if (tid == 0) {
        __threadfence();
        last = atomicAdd(&state[kCsrDone], 1u) == gridDim.x - 1;
        if (last) {
            __threadfence();
            stale = atomicOr(&state[kCsrStale], 0u);
            state[kCsrDone] = 0u;
        }
    }
    __syncthreads();
    if (!last || !stale) return;
    for (uint32_t v = tid; v <= V; v += 1024) cnt[v] = 0u;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += 1024) {
        const uint32_t v = idx[i];
        if (v < V) atomicAdd(&cnt[v], 1u);
    }
    __syncthreads();
    // exclusive scan of cnt[0..V]: consecutive runs per thread, then the runs' totals
    const uint32_t per = (V + 1 + 1023) / 1024, lo = tid * per, hi = min(lo + per, V + 1);
    uint32_t sum = 0;
    for (uint32_t v = lo; v < hi; ++v) sum += cnt[v];
    uint32_t run;
    Scan(scan_tmp).ExclusiveSum(sum, run);
    for (uint32_t v = lo; v < hi; ++v) {
        const uint32_t c = cnt[v];
        cnt[v] = run;  // now the list's cursor
        offsets[v] = run;
        run += c;
    }
    __syncthreads();
    // corners with an invalid vertex go past offsets[V], each with its slot
    for (uint32_t i = tid; i < n; i += 1024) {
        const uint32_t v = idx[i];
        const uint32_t j = atomicAdd(&cnt[v < V ? v : V], 1u);
        corners[j] = i;
        if (v >= V) slot_of[i] = j;
        saved[i] = v;
    }
    __syncthreads();
    for (uint32_t v = tid; v < V; v += 1024) {
        const uint32_t b = offsets[v], m = offsets[v + 1] - b;
        uint32_t* a = corners + b;
        sort_corner_list(a, m);
        for (uint32_t i = 0; i < m; ++i) slot_of[a[i]] = b + i;
    }
    if (tid == 0) {
        state[kCsrStale] = 0u;
        state[kCsrBuilt] = 1u;
        state[kCsrT] = T;
        state[kCsrV] = V;
    }
}

// Larger meshes: the compare, then count / scan / place / sort kernels that
// each exit at once while kCsrStale is clear.
__global__ void k_csr_compare(const uint32_t* __restrict__ idx, uint32_t n, uint32_t T, uint32_t V, uint32_t* state,
                              const uint32_t* __restrict__ saved) {
    if (__syncthreads_or(csr_differs(idx, saved, n, T, V, state)) && threadIdx.x == 0)
        atomicOr(&state[kCsrStale], 1u);
}

__global__ void k_count_corners(const uint32_t* __restrict__ idx, uint32_t n, uint32_t V, const uint32_t* state,
                                uint32_t* __restrict__ counts) {
    if (!state[kCsrStale]) return;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        if (idx[i] < V) atomicAdd(&counts[idx[i]], 1u);
}

// Exclusive scan of counts[0..m) in tiles of kCsrScanTile: each tile's total,
// then each tile's scan offset by the totals of the tiles before it.
__global__ __launch_bounds__(256) void k_csr_scan_totals(const uint32_t* __restrict__ counts, uint32_t m,
                                                         const uint32_t* state, uint32_t* __restrict__ totals) {
    if (!state[kCsrStale]) return;
    using Reduce = hipcub::BlockReduce<uint32_t, 256>;
    __shared__ typename Reduce::TempStorage tmp;
    const uint32_t base = blockIdx.x * kCsrScanTile + threadIdx.x * 8;
    uint32_t s = 0;
    for (uint32_t k = 0; k < 8; ++k)
        if (base + k < m) s += counts[base + k];
    s = Reduce(tmp).Sum(s);
    if (threadIdx.x == 0) totals[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_csr_scan_out(const uint32_t* __restrict__ counts, uint32_t m,
                                                      const uint32_t* state, const uint32_t* __restrict__ totals,
                                                      uint32_t* __restrict__ offsets) {
    if (!state[kCsrStale]) return;
    using Reduce = hipcub::BlockReduce<uint32_t, 256>;
    using Scan = hipcub::BlockScan<uint32_t, 256>;
    __shared__ typename Reduce::TempStorage rtmp;
    __shared__ typename Scan::TempStorage stmp;
    __shared__ uint32_t prefix;
    uint32_t s = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 256) s += totals[b];
    s = Reduce(rtmp).Sum(s);
    if (threadIdx.x == 0) prefix = s;
    const uint32_t base = blockIdx.x * kCsrScanTile + threadIdx.x * 8;
    uint32_t c[8], t = 0;
    for (uint32_t k = 0; k < 8; ++k) {
        c[k] = base + k < m ? counts[base + k] : 0u;
        t += c[k];
    }
    uint32_t run;
    Scan(stmp).ExclusiveSum(t, run);
    __syncthreads();
    run += prefix;
    for (uint32_t k = 0; k < 8; ++k) {
        if (base + k < m) offsets[base + k] = run;
        run += c[k];
    }
}

// Places every corner in its vertex's list (order within a list fixed by
// k_sort_corner_lists) and saves the indices; corners with an invalid vertex go
// past offsets[V].  counts (after the scan) is consumed as a per-vertex cursor.
__global__ void k_place_corners(const uint32_t* __restrict__ idx, uint32_t n, uint32_t V, const uint32_t* state,
                                const uint32_t* __restrict__ offsets, uint32_t* __restrict__ counts,
                                uint32_t* __restrict__ corners, uint32_t* __restrict__ slot_of,
                                uint32_t* __restrict__ saved) {
    if (!state[kCsrStale]) return;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t v = idx[i];
        uint32_t j;
        if (v < V) {
            j = offsets[v] + atomicSub(&counts[v], 1u) - 1u;
        } else {
            j = offsets[V] + atomicAdd(&counts[V], 1u);
            slot_of[i] = j;
        }
        corners[j] = i;
        saved[i] = v;
    }
}

// One lane per vertex: its list sorted ascending, then the inverse map corner
// -> list position.  The last block to finish records the build.
__global__ void k_sort_corner_lists(uint32_t T, uint32_t V, uint32_t* state, const uint32_t* __restrict__ offsets,
                                    uint32_t* __restrict__ corners, uint32_t* __restrict__ slot_of,
                                    uint32_t* __restrict__ counts) {
    if (!state[kCsrStale]) return;
    __shared__ bool last;
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock) {
        const uint32_t b = offsets[v], n = offsets[v + 1] - b;
        uint32_t* a = corners + b;
        sort_corner_list(a, n);
        for (uint32_t i = 0; i < n; ++i) slot_of[a[i]] = b + i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&state[kCsrDone], 1u) == gridDim.x - 1;
        if (last) {
            state[kCsrDone] = 0u;
            counts[V] = 0u;
            state[kCsrBuilt] = 1u;
            state[kCsrT] = T;
            state[kCsrV] = V;
            __threadfence();
            state[kCsrStale] = 0u;
        }
    }
}

// reduce_vertex_gradients.hlsl:12-29 for every view: the corner gradients sit
// in the adjacency's order, so each vertex sums a contiguous run, and a block's
// 256 vertices sum one contiguous range.  The block streams that range through
// LDS in chunks of kReduceChunk corners (coalesced float4 loads); each lane adds
// the part of its run inside the chunk, in corner order, chunk after chunk.
constexpr uint32_t kReduceChunk = 2048;

// The backward's last kernel: it also clears the view's won marks and whole-box
// flag (read by the corner kernels before it) and the large-box queue's
// counters, which are zero between calls.
__global__ __launch_bounds__(kBlock) void k_vertex_reduce(ViewTable vt, uint32_t T, uint32_t V,
                                                          const uint32_t* __restrict__ offsets,
                                                          const float4* __restrict__ corner,
                                                          uint8_t* __restrict__ won, uint32_t* __restrict__ view_flags,
                                                          uint32_t* __restrict__ counters) {
    __shared__ float4 chunk[kReduceChunk];
    const uint32_t view = blockIdx.y;
    const ViewDesc& vd = vt.v[view];
    clear_queue_counters(counters);  // k_corner_large, before this pass, was the queue's consumer
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < T; t += gridDim.x * kBlock)
        won[static_cast<size_t>(view) * T + t] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) view_flags[view] = 0u;
    const float4* cg = corner + static_cast<size_t>(view) * T * 3;
    for (uint32_t v0 = blockIdx.x * kBlock; v0 < V; v0 += gridDim.x * kBlock) {
        const uint32_t v = v0 + threadIdx.x;
        const uint32_t last = v0 + kBlock < V ? v0 + kBlock : V;
        const uint32_t lo = offsets[v0], hi = offsets[last];  // the block's corner range
        const uint32_t a0 = v < V ? offsets[v] : hi, a1 = v < V ? offsets[v + 1] : hi;
        float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (uint32_t c = lo; c < hi; c += kReduceChunk) {
            const uint32_t n = hi - c < kReduceChunk ? hi - c : kReduceChunk;
            __syncthreads();  // the previous chunk's reads are done
            for (uint32_t i = threadIdx.x; i < n; i += kBlock) chunk[i] = cg[c + i];
            __syncthreads();
            const uint32_t b = a0 > c ? a0 : c, e = a1 < c + n ? a1 : c + n;
            for (uint32_t a = b; a < e; ++a) {
                const float4 q = chunk[a - c];
                s.x += q.x;
                s.y += q.y;
                s.z += q.z;
                s.w += q.w;
            }
        }
        if (v < V) reinterpret_cast<float4*>(vd.grad_pos)[v] = s;
    }
}

__global__ void k_mark_if_differ(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, size_t n,
                                 uint32_t* flag) {
    bool diff = false;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        diff |= a[i] != b[i];
    if (__any(diff) && (threadIdx.x & 63u) == 0) atomicOr(flag, 1u);
}

__global__ void k_check_indices(const uint32_t* idx, size_t n, uint32_t V, uint32_t* flags) {
    bool bad = false;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        bad |= idx[i] >= V;
    if (__any(bad) && (threadIdx.x & 63u) == 0) atomicOr(flags, 1u);
}

// Triangles per wave batch: 64, or fewer for small meshes so that the launch
// still has about `waves` waves to overlap their latencies.
uint32_t tri_batch(uint32_t T, uint32_t views, uint32_t waves) {
    const uint64_t per = (static_cast<uint64_t>(T) * views + waves - 1) / waves;
    return per >= 64 ? 64u : (per <= 1 ? 1u : static_cast<uint32_t>(per));
}

// Workgroups per view: one per four batches, at most 8192 per view, or with
// `views` > 0 at most 8192 over all views (the candidate pass: each wave then
// loops over batches, its next batch's corners in flight).
uint32_t tri_grid(uint32_t T, uint32_t bsz, uint32_t views = 0) {
    const uint32_t g = (T + (kBlock / 64) * bsz - 1) / ((kBlock / 64) * bsz);
    const uint32_t cap = views ? (8192u + views - 1u) / views : 8192u;
    return g < cap ? (g ? g : 1u) : cap;
}

}  // namespace

hipError_t fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_u32<<<grid_for(n), kBlock, 0, st>>>(p, value, n);
    return hipGetLastError();
}

hipError_t fill_f32(float* p, float value, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0)
        k_fill_f32x4<<<grid_for(n / 4 ? n / 4 : 1), kBlock, 0, st>>>(p, value, n);
    else
        k_fill_f32<<<grid_for(n), kBlock, 0, st>>>(p, value, n);
    return hipGetLastError();
}

hipError_t fill_u64(unsigned long long* p, unsigned long long value, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_u64<<<grid_for(n), kBlock, 0, st>>>(p, value, n);
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

hipError_t raster_candidates(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, uint32_t cull,
                             const RasterScratch& s, const TuneParams& tp, hipStream_t st) {
    hipError_t e;
    const uint32_t bsz = tri_batch(T, vt.n, tp.fwd_waves);
    k_raster_small<<<dim3(tri_grid(T, bsz, vt.n), vt.n), kBlock, 0, st>>>(vt, idx, T, V, cull, s.keys, s.queue,
                                                                    s.counters, bsz, tp.wave_work, nullptr, nullptr,
                                                                    tp.order);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_raster_large<<<kLargeBlocks, kBlock, 0, st>>>(vt, idx, T, V, cull, s.keys, s.queue, s.counters);
    return hipGetLastError();
}

size_t tri_record_bytes() { return sizeof(TriRecord); }

size_t bin_scan_temp_bytes(uint32_t n) {
    size_t a = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), static_cast<int>(n));
    return a;
}

hipError_t raster_tiled(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, uint32_t cull,
                        int want_deriv, const RasterScratch& s, const TuneParams& tp, hipStream_t st) {
    if (tp.bin_tiles == 0 || tp.bin_tiles > kMaxBinTiles) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = fill_u32(s.bin_counts, 0u, static_cast<size_t>(vt.total_rtiles) + 1, st)) != hipSuccess) return e;
    if ((e = fill_u32(reinterpret_cast<uint32_t*>(s.bin_seeded), 0u, (vt.total_rtiles + 3u) / 4u, st)) != hipSuccess)
        return e;
    if ((e = fill_u32(s.medium_count, 0u, kMaxViews, st)) != hipSuccess) return e;
    const dim3 grid(grid_for(T) < 1024 ? grid_for(T) : 1024, vt.n);
    k_bin_count<<<grid, kBlock, 0, st>>>(vt, idx, T, V, cull, tp.bin_tiles, s.bin_counts, s.bin_seeded,
                                         s.medium_list, s.medium_count, static_cast<TriRecord*>(s.tri_records),
                                         s.tri_spans);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the triangles too large to bin (global keys; the largest of them queued as slabs)
    k_raster_small<<<dim3(64, vt.n), kBlock, 0, st>>>(vt, idx, T, V, cull, s.keys, s.queue, s.counters,
                                                     64u, tp.wave_work, s.medium_list, s.medium_count, 0u);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t bytes = s.bin_temp_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(s.bin_temp, bytes, s.bin_counts, s.bin_offsets,
                                              static_cast<int>(vt.total_rtiles + 1), st)) != hipSuccess)
        return e;
    k_bin_fill<<<grid, kBlock, 0, st>>>(vt, T, s.bin_counts, s.bin_offsets, s.bin_list, s.tri_spans);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_raster_large<<<kLargeBlocks, kBlock, 0, st>>>(vt, idx, T, V, cull, s.keys, s.queue, s.counters);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_tile_raster<<<vt.total_rtiles, kBlock, 0, st>>>(vt, idx, T, want_deriv, s.bin_offsets, s.bin_list,
                                                      s.bin_seeded, static_cast<const TriRecord*>(s.tri_records),
                                                      s.keys, s.counters);
    return hipGetLastError();
}

hipError_t raster_resolve(const ViewTable& vt, const uint32_t* idx, int want_deriv, const RasterScratch& s,
                          hipStream_t st) {
    k_raster_resolve<<<(vt.total_keys + kBlock - 1) / kBlock, kBlock, 0, st>>>(vt, idx, want_deriv, s.keys,
                                                                                s.counters);
    return hipGetLastError();
}

hipError_t backward_prepass(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, const RasterScratch& s,
                            hipStream_t st) {
    k_bwd_prepass<<<(vt.total_keys + kBlock - 1) / kBlock, kBlock, 0, st>>>(vt, idx, T, V, s.cid, s.cgrad,
                                                                           s.view_flags, s.won);
    return hipGetLastError();
}

size_t csr_total_words(uint32_t V) { return (static_cast<size_t>(V) + 1 + kCsrScanTile - 1) / kCsrScanTile; }

// Vertex -> corner adjacency with corners ascending per vertex, and each
// corner's position in it; rebuilt only when the indices differ from the
// ones it was last built from.
hipError_t build_vertex_csr(const uint32_t* idx, uint32_t T, uint32_t V, const VertexCsr& c, hipStream_t st) {
    const uint32_t n = T * 3;
    hipError_t e;
    if (V < kCsrLdsVerts) {
        const uint32_t blocks = (n + 8191) / 8192;
        k_csr_small<<<blocks < 1u ? 1u : (blocks > 64u ? 64u : blocks), 1024, 0, st>>>(idx, n, T, V, c.state, c.saved, c.offsets, c.corners,
                                                          c.slot_of);
        return hipGetLastError();
    }
    k_csr_compare<<<grid_for(n), kBlock, 0, st>>>(idx, n, T, V, c.state, c.saved);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_count_corners<<<grid_for(n), kBlock, 0, st>>>(idx, n, V, c.state, c.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t m = V + 1, tiles = static_cast<uint32_t>(csr_total_words(V));
    k_csr_scan_totals<<<tiles, 256, 0, st>>>(c.counts, m, c.state, c.totals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_csr_scan_out<<<tiles, 256, 0, st>>>(c.counts, m, c.state, c.totals, c.offsets);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_place_corners<<<grid_for(n), kBlock, 0, st>>>(idx, n, V, c.state, c.offsets, c.counts, c.corners, c.slot_of,
                                                    c.saved);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_sort_corner_lists<<<grid_for(V), kBlock, 0, st>>>(T, V, c.state, c.offsets, c.corners, c.slot_of, c.counts);
    return hipGetLastError();
}

hipError_t mark_if_differ(const uint32_t* a, const uint32_t* b, size_t n, uint32_t* flag, hipStream_t st) {
    if (!n) return hipSuccess;
    k_mark_if_differ<<<grid_for(n), kBlock, 0, st>>>(a, b, n, flag);
    return hipGetLastError();
}

hipError_t corner_gradients(const ViewTable& vt, const uint32_t* idx, uint32_t T, uint32_t V, const RasterScratch& s,
                            float4* corner, const TuneParams& tp, bool ordered, hipStream_t st) {
    hipError_t e;
    const uint32_t bsz = tri_batch(T, vt.n, tp.bwd_waves);
    k_corner_small<<<dim3(tri_grid(T, bsz), vt.n), kBlock, 0, st>>>(vt, idx, T, V, s.cid, s.cgrad, s.view_flags,
                                                                    s.won, s.slot_of, corner, s.queue, s.counters,
                                                                    bsz, tp.flat_work);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ordered)
        k_corner_large_ordered<<<kLargeBlocks, kBlock, 0, st>>>(vt, idx, T, V, s.cid, s.cgrad, s.view_flags,
                                                                s.slot_of, corner, s.queue, s.counters);
    else
        k_corner_large<<<kLargeBlocks, kBlock, 0, st>>>(vt, idx, T, V, s.cid, s.cgrad, s.view_flags, s.slot_of,
                                                        corner, s.queue, s.counters);
    return hipGetLastError();
}

hipError_t vertex_reduce(const ViewTable& vt, uint32_t T, uint32_t V, const VertexCsr& csr, const float4* corner,
                         uint8_t* won, uint32_t* view_flags, uint32_t* counters, hipStream_t st) {
    k_vertex_reduce<<<dim3(grid_for(V) < 4096 ? grid_for(V) : 4096, vt.n), kBlock, 0, st>>>(
        vt, T, V, csr.offsets, corner, won, view_flags, counters);
    return hipGetLastError();
}

}  // namespace pdr

// Self-test hook (host only, no device work): the candidate region of one
// triangle given by its clip positions (12 floats) in a W x H view with the
// resolved viewport (vx, vy, vw, vh), as the forward (allow_cols = 1) or the
// backward (allow_cols = 0) walks it.  Writes the candidate rectangle
// (x0, x1, y0, y1) and mask[y * W + x] = 1 for every candidate pixel.  Returns
// 1 when the triangle has candidates, 0 otherwise.
extern "C" int pdr_testing_candidate_region(const float* clip12, uint32_t width, uint32_t height,
                                            const float* viewport4, uint32_t cull, int allow_cols, int32_t* rect4,
                                            uint8_t* mask) {
    pdr::ViewDesc vd{};
    vd.width = width;
    vd.height = height;
    vd.vx = viewport4[0];
    vd.vy = viewport4[1];
    vd.vw = viewport4[2];
    vd.vh = viewport4[3];
    const float4 p0 = make_float4(clip12[0], clip12[1], clip12[2], clip12[3]);
    const float4 p1 = make_float4(clip12[4], clip12[5], clip12[6], clip12[7]);
    const float4 p2 = make_float4(clip12[8], clip12[9], clip12[10], clip12[11]);
    pdr::ScreenTri r;
    pdr::LineSpans sp;
    for (size_t i = 0; i < static_cast<size_t>(width) * height; ++i) mask[i] = 0;
    if (!pdr::setup_screen(vd, p0, p1, p2, cull, allow_cols != 0, r, sp)) return 0;
    rect4[0] = r.x0;
    rect4[1] = r.x1;
    rect4[2] = r.y0;
    rect4[3] = r.y1;
    const float alpha = pdr::span_alpha(vd, sp.cols), beta = pdr::span_beta(vd, sp.cols);
    const int32_t l0 = sp.cols ? r.x0 : r.y0, l1 = sp.cols ? r.x1 : r.y1;
    for (int32_t l = l0; l <= l1; ++l) {
        // the line's pixel-centre sample (rasterize_forward.hlsl:55-57)
        const float s = sp.cols ? (static_cast<float>(l) - vd.vx + 0.5f) * 2.0f / vd.vw - 1.0f
                                : (static_cast<float>(l) - vd.vy + 0.5f) * 2.0f / vd.vh - 1.0f;
        int32_t lo, hi;
        pdr::line_span(sp, s, alpha, beta, sp.cols ? r.y0 : r.x0, sp.cols ? r.y1 : r.x1, &lo, &hi);
        for (int32_t a = lo; a <= hi; ++a) {
            const int32_t x = sp.cols ? l : a, y = sp.cols ? a : l;
            mask[static_cast<size_t>(y) * width + x] = 1;
        }
    }
    return 1;
}
