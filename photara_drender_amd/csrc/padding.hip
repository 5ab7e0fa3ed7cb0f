// pad_texture_atlas (src/texture_padding.cpp:226-351) on gfx950.
//
// The nearest-source field is Danielsson's 8SSEDT (src/texture_padding.cpp:
// 37-97): a forward raster sweep relaxing each texel from its left, upper,
// upper-left and upper-right neighbours, then the mirrored backward sweep.
// A texel's forward value depends on its left neighbour (same sweep) and on
// three texels of the row above, so texel (x, y) can be settled at step
// t = x + 2 y: the sweep runs as a wavefront of W + 2 (H - 1) steps in one
// workgroup, one row per lane and step (rows y, y + R, y + 2 R, ... share a
// lane, R = ceil(W / 2), and are never active together).  Each row keeps its
// last four settled texels in an LDS ring, which is everything the next row
// reads.  The backward sweep is the forward sweep on the mirrored grid with
// negated offsets (same relaxation order, same squared lengths), so results
// equal the sequential sweeps exactly.  Preparation and the fill are one
// texel per lane over the whole grid.
#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr int32_t kFar = 12000;   // offset of "no source yet"
constexpr int32_t kSide = 16384;  // offsets are clamped to +-kSide (int16 storage)
constexpr uint32_t kSweepThreads = 1024;
constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint32_t pack(int32_t x, int32_t y) {
    return (static_cast<uint32_t>(x) & 0xFFFFu) | (static_cast<uint32_t>(y) << 16);
}
__device__ __forceinline__ int32_t off_x(uint32_t v) { return static_cast<int16_t>(v & 0xFFFFu); }
__device__ __forceinline__ int32_t off_y(uint32_t v) { return static_cast<int16_t>(v >> 16); }

__device__ __forceinline__ int32_t clamp_side(int32_t v) { return v < -kSide ? -kSide : (v > kSide ? kSide : v); }

// An offset (x, y) as two int16 lanes: the same bits as pack(x, y).
typedef int16_t Offset __attribute__((ext_vector_type(2)));

__device__ __forceinline__ Offset as_offset(uint32_t v) { return __builtin_bit_cast(Offset, v); }
__device__ __forceinline__ uint32_t as_word(Offset v) { return __builtin_bit_cast(uint32_t, v); }
// x^2 + y^2 (at most 2 * 30001^2 here, so the int32 dot product is exact)
__device__ __forceinline__ int32_t len2(Offset v) { return __builtin_amdgcn_sdot2(v, v, 0, false); }

// A neighbour that does not exist: a candidate built from it is never closer
// than any offset the sweep holds (|components| <= kSide).
constexpr int16_t kNone16 = 30000;

// cur <- neighbour's offset + d when strictly closer (squared lengths), clamped
// to +-kSide; cur2 = |cur|^2.  Components stay within int16 (|nb| <= kNone16).
__device__ __forceinline__ void relax(Offset& cur, int32_t& cur2, Offset nb, Offset d) {
    const Offset cand = nb + d;
    const int32_t c2 = len2(cand);
    const Offset side = {static_cast<int16_t>(kSide), static_cast<int16_t>(kSide)};
    const Offset clamped = __builtin_elementwise_max(__builtin_elementwise_min(cand, side), -side);
    const bool take = c2 < cur2;
    cur = take ? clamped : cur;
    cur2 = take ? len2(clamped) : cur2;
}

// Adds a lane count to *counter once per wave (callers loop over a bounded grid,
// so few atomics meet on the counter).
__device__ __forceinline__ void count_total(unsigned long long* counter, uint32_t n) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) n += __shfl_xor(n, m, 64);
    if ((threadIdx.x & 63u) == 0 && n) atomicAdd(counter, static_cast<unsigned long long>(n));
}

// Step barrier of the sweep: LDS writes complete, then the workgroup barrier.
// Global stores of earlier steps may stay in flight (no step reads them).
__device__ __forceinline__ void lds_step_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Sources (:259-276): valid texels and earlier covered fills; other texels lose their fill flag.
// counters: valid, gutter, unobserved, remaining, filled.
__global__ void k_pad_prepare(uint32_t P, const float* __restrict__ valid, const float* __restrict__ coverage,
                              float* __restrict__ filled, float threshold, uint8_t* __restrict__ source,
                              unsigned long long* __restrict__ counters) {
    uint32_t nvalid = 0, nrefill = 0;
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < P; p += gridDim.x * kBlock) {
        const bool is_valid = valid[p] >= threshold;
        const bool refill = !is_valid && filled && filled[p] > 0.5f && coverage && coverage[p] > 0.5f;
        if (!is_valid && !refill && filled) filled[p] = 0.0f;
        source[p] = is_valid || refill ? 1 : 0;
        nvalid += is_valid ? 1u : 0u;
        nrefill += refill ? 1u : 0u;
    }
    count_total(counters + 0, nvalid);
    count_total(counters + 2, nrefill);
    count_total(counters + 4, nrefill);
}

// One 8SSEDT sweep as a wavefront; kMirror = the backward sweep.  field
// holds packed int16 offsets: read as the initial state of the backward
// sweep, written by both; field[W * H] is a discard slot.  ring: R x 4
// settled offsets (dynamic LDS).  Each lane walks kC row classes
// c = lane + j * blockDim.x; a class settles (x, y) at step 2 c + x + 2 R k
// for its k-th row y = c + k R.  A step's only global load is the input of a
// texel, which does not depend on the sweep: each class loads it kD steps
// ahead into a register ring indexed by the step (static slots: the step loop
// is unrolled by kD).  Every lane issues the same global loads and stores each
// step (clamped addresses, the discard slot), so the waits for a step's
// input cover only that load and not the ones issued after it.
// Word i of an array below 4 GiB, addressed by a 32-bit byte offset.
__device__ __forceinline__ uint32_t* word_at(uint32_t* base, uint32_t i) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + (i << 2));
}

template <bool kMirror>
__device__ __forceinline__ uint32_t sweep_input(const uint8_t* __restrict__ source, const uint32_t* __restrict__ field,
                                                uint32_t W, uint32_t H, uint32_t x, uint32_t y) {
    x = x < W ? x : W - 1u;
    y = y < H ? y : H - 1u;
    const uint32_t rx = kMirror ? W - 1u - x : x, ry = kMirror ? H - 1u - y : y;
    const uint32_t idx = ry * W + rx;
    if constexpr (kMirror)
        return *word_at(const_cast<uint32_t*>(field), idx);
    else
        return static_cast<uint32_t>(source[idx]);
}

// Slot k & 3 of a ring row (k static after unrolling).
__device__ __forceinline__ uint32_t ring_slot(const uint4& row, int k) {
    switch (k & 3) {
        case 0: return row.x;
        case 1: return row.y;
        case 2: return row.z;
        default: return row.w;
    }
}

// Next texel of a class walk: along the row, then R rows down.
__device__ __forceinline__ void class_advance(uint32_t& x, uint32_t& y, uint32_t R) {
    const bool wrap = x + 1u == 2u * R;
    x = wrap ? 0u : x + 1u;
    y = wrap ? y + R : y;
}

template <int kC, int kD, bool kMirror>
__global__ __launch_bounds__(kSweepThreads) void k_pad_sweep(uint32_t W, uint32_t H, const uint8_t* __restrict__ source,
                                                              uint32_t* __restrict__ field) {
    extern __shared__ uint4 ring_rows[];  // R rows of 4 slots, then a discard word
    uint32_t* ring = reinterpret_cast<uint32_t*>(ring_rows);
    const uint32_t R = (W + 1u) / 2u;
    const uint32_t steps = W + 2u * (H - 1u);
    const uint32_t discard = W * H;
    const Offset none = {kNone16, kNone16};
    uint32_t px[kC], py[kC];  // texel the class settles next
    uint32_t fx[kC], fy[kC];  // texel whose input the class loads next
    uint32_t in[kC][kD];      // in[j][t % kD]: input of class j's texel of step t
#pragma unroll
    for (int j = 0; j < kC; ++j) {
        const uint32_t c = threadIdx.x + static_cast<uint32_t>(j) * blockDim.x;
        px[j] = fx[j] = 0;
        py[j] = fy[j] = c;
#pragma unroll
        for (int d = 0; d < kD; ++d) {
            in[j][d] = sweep_input<kMirror>(source, field, W, H, fx[j], fy[j]);
            if (c < R && static_cast<uint32_t>(d) >= 2u * c) class_advance(fx[j], fy[j], R);
        }
    }
    static_assert(kD % 4 == 0, "ring slots are static only for kD a multiple of 4");
    for (uint32_t t0 = 0; t0 < steps; t0 += kD) {
#pragma unroll
        for (int s = 0; s < kD; ++s) {
            // steps past the last one (up to kD - 1) settle nothing: stores go to the discard slots
            const uint32_t t = t0 + static_cast<uint32_t>(s);
#pragma unroll
            for (int j = 0; j < kC; ++j) {
                const uint32_t c = threadIdx.x + static_cast<uint32_t>(j) * blockDim.x;
                const bool live = c < R && t >= 2u * c;
                const uint32_t x = px[j], y = py[j];
                const bool settle = live && x < W && y < H;
                // ring rows of this class and of the row above (clamped for idle lanes),
                // slot k = step k mod 4 (t0 is a multiple of 4, so the slots are static)
                const uint4 own = ring_rows[c < R ? c : R - 1u];
                const uint4 up = ring_rows[c == 0 ? R - 1u : (c - 1u < R ? c - 1u : R - 1u)];
                const Offset left = x > 0 ? as_offset(ring_slot(own, s + 3)) : none;
                const Offset above = y > 0 ? as_offset(ring_slot(up, s + 2)) : none;
                const Offset above_left = y > 0 && x > 0 ? as_offset(ring_slot(up, s + 1)) : none;
                const Offset above_right = y > 0 && x + 1u < W ? as_offset(ring_slot(up, s + 3)) : none;
                Offset cur;
                if constexpr (kMirror) {
                    cur = -as_offset(in[j][s]);
                } else {
                    const Offset far = {static_cast<int16_t>(kFar), static_cast<int16_t>(kFar)};
                    const Offset zero = {0, 0};
                    cur = in[j][s] ? zero : far;
                }
                int32_t cur2 = len2(cur);
                relax(cur, cur2, left, Offset{-1, 0});
                relax(cur, cur2, above, Offset{0, -1});
                relax(cur, cur2, above_left, Offset{-1, -1});
                relax(cur, cur2, above_right, Offset{1, -1});
                ring[settle ? c * 4u + static_cast<uint32_t>(s & 3) : R * 4u] = as_word(cur);
                const uint32_t rx = kMirror ? W - 1u - x : x, ry = kMirror ? H - 1u - y : y;
                *word_at(field, settle ? ry * W + rx : discard) = as_word(kMirror ? -cur : cur);
                if (live) class_advance(px[j], py[j], R);
                // the input of the class's texel at step t + kD
                in[j][s] = sweep_input<kMirror>(source, field, W, H, fx[j], fy[j]);
                if (c < R && t + static_cast<uint32_t>(kD) >= 2u * c) class_advance(fx[j], fy[j], R);
            }
            lds_step_barrier();
        }
    }
}

// Row classes per lane and lanes of the sweep for a W-wide grid.
void sweep_shape(uint32_t W, int& classes, uint32_t& threads) {
    const uint32_t R = (W + 1u) / 2u;
    classes = R <= kSweepThreads ? 1 : (R <= 2 * kSweepThreads ? 2 : (R <= 4 * kSweepThreads ? 4 : 8));
    threads = classes == 1 ? (R + 63u) / 64u * 64u : kSweepThreads;
}

// Fills (:289-348).  Sources are never written; colours are read from `src_color`
// (the atlas itself, or a copy of it when a far sentinel offset could point at
// a non-source texel inside the atlas).
__global__ void k_pad_fill(uint32_t W, uint32_t H, uint32_t C, const uint32_t* __restrict__ field,
                           const uint8_t* __restrict__ source, const float* __restrict__ coverage,
                           unsigned long long margin2, unsigned long long limit2, int fill_unobserved,
                           const float* src_color, float* color, float* __restrict__ filled,
                           unsigned long long* __restrict__ counters) {
    uint32_t ngutter = 0, nunobserved = 0, nremaining = 0;
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < W * H; p += gridDim.x * kBlock) {
        if (source[p]) continue;
        const uint32_t v = field[p];
        const int64_t ox = off_x(v), oy = off_y(v);
        const unsigned long long d2 = static_cast<unsigned long long>(ox * ox + oy * oy);
        const int64_t sx = static_cast<int64_t>(p % W) + ox, sy = static_cast<int64_t>(p / W) + oy;
        if (d2 >= 2ull * kFar * kFar || sx < 0 || sy < 0 || sx >= W || sy >= H) continue;
        const bool covered = coverage && coverage[p] > 0.5f;
        const bool fill = covered ? (fill_unobserved && d2 <= limit2) : (margin2 > 0 && d2 <= margin2);
        if (!fill) {
            nremaining += covered ? 1u : 0u;
            continue;
        }
        const size_t s = static_cast<size_t>(sy) * W + static_cast<size_t>(sx);
        for (uint32_t c = 0; c < C; ++c) color[static_cast<size_t>(p) * C + c] = src_color[s * C + c];
        if (C == 4) color[static_cast<size_t>(p) * C + 3] = 1.0f;
        if (filled) filled[p] = 1.0f;
        ngutter += covered ? 0u : 1u;
        nunobserved += covered ? 1u : 0u;
    }
    count_total(counters + 1, ngutter);
    count_total(counters + 2, nunobserved);
    count_total(counters + 3, nremaining);
    count_total(counters + 4, ngutter + nunobserved);
}

uint32_t blocks_of(uint64_t n) {
    const uint64_t b = (n + kBlock - 1) / kBlock;
    return static_cast<uint32_t>(b < 2048 ? (b ? b : 1) : 2048);
}

}  // namespace

hipError_t pad_prepare(uint32_t P, const float* valid, const float* coverage, float* filled, float threshold,
                       uint8_t* source, unsigned long long* counters, hipStream_t st) {
    k_pad_prepare<<<blocks_of(P), kBlock, 0, st>>>(P, valid, coverage, filled, threshold, source, counters);
    return hipGetLastError();
}

template <int kC, int kD>
static hipError_t launch_sweeps(uint32_t W, uint32_t H, const uint8_t* source, uint32_t* field, uint32_t threads,
                                size_t lds, hipStream_t st) {
    if (lds > 64u * 1024u) {
        for (const void* k : {reinterpret_cast<const void*>(&k_pad_sweep<kC, kD, false>),
                              reinterpret_cast<const void*>(&k_pad_sweep<kC, kD, true>)}) {
            const hipError_t e =
                hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
            if (e != hipSuccess) return e;
        }
    }
    k_pad_sweep<kC, kD, false><<<1, threads, lds, st>>>(W, H, source, field);
    k_pad_sweep<kC, kD, true><<<1, threads, lds, st>>>(W, H, source, field);
    return hipGetLastError();
}

hipError_t pad_field(uint32_t W, uint32_t H, const uint8_t* source, uint32_t* field, hipStream_t st) {
    const size_t lds = (static_cast<size_t>((W + 1u) / 2u) * 4u + 1u) * sizeof(uint32_t);
    int classes = 1;
    uint32_t threads = 64;
    sweep_shape(W, classes, threads);
    if (classes == 1) return launch_sweeps<1, 8>(W, H, source, field, threads, lds, st);
    if (classes == 2) return launch_sweeps<2, 8>(W, H, source, field, threads, lds, st);
    if (classes == 4) return launch_sweeps<4, 8>(W, H, source, field, threads, lds, st);
    return launch_sweeps<8, 4>(W, H, source, field, threads, lds, st);
}

hipError_t pad_fill(uint32_t W, uint32_t H, uint32_t C, const uint32_t* field, const uint8_t* source,
                    const float* coverage, uint32_t margin, int fill_unobserved, uint32_t max_fill,
                    const float* src_color, float* color, float* filled, unsigned long long* counters,
                    hipStream_t st) {
    const unsigned long long margin2 = static_cast<unsigned long long>(margin) * margin;
    const unsigned long long limit2 = max_fill == 0 ? ~0ull : static_cast<unsigned long long>(max_fill) * max_fill;
    k_pad_fill<<<blocks_of(static_cast<uint64_t>(W) * H), kBlock, 0, st>>>(
        W, H, C, field, source, coverage, margin2, limit2, fill_unobserved, src_color, color, filled, counters);
    return hipGetLastError();
}

}  // namespace pdr
