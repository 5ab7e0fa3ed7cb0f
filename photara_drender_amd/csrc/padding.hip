// pad_texture_atlas (src/texture_padding.cpp:226-351) on gfx950.
//
// The nearest-source field is Danielsson's 8SSEDT (src/texture_padding.cpp:
// 37-97): a forward raster sweep relaxing each texel from its left, upper,
// upper-left and upper-right neighbours, then the mirrored backward sweep.
// A texel's forward value depends on its left neighbour (same sweep) and on
// three texels of the row above, so texel (x, y) can be settled at step
// t = x + 2 y: the sweep runs as a wavefront.  Rows go in bands of 64 to
// single-wave workgroups (lane = row), each band pipelined behind the band
// above.  The backward sweep is the forward sweep on the mirrored grid with
// negated offsets (same relaxation order, same squared lengths), so results
// equal the sequential sweeps exactly.  Preparation and the fill are one
// texel per lane over the whole grid.
#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr int32_t kFar = 12000;   // offset of "no source yet"
constexpr int32_t kSide = 16384;  // offsets are clamped to +-kSide (int16 storage)
constexpr uint32_t kBlock = 256;

__device__ __forceinline__ int32_t off_x(uint32_t v) { return static_cast<int16_t>(v & 0xFFFFu); }
__device__ __forceinline__ int32_t off_y(uint32_t v) { return static_cast<int16_t>(v >> 16); }

// An offset (x, y) as two int16 lanes: x in the low half of the word, y in the high half.
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

// Word i of an array below 4 GiB, addressed by a 32-bit byte offset.
__device__ __forceinline__ uint32_t* word_at(uint32_t* base, uint32_t i) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + (i << 2));
}

// One 8SSEDT sweep; kMirror = the backward sweep.  field holds packed int16
// offsets: read as the initial state of the backward sweep, written by both;
// field[W * H] is a discard slot.  Rows in bands of 64, one single-wave
// workgroup per band, lane i = row 64 b + i settling column x = t - 2 i at its
// step t.  Inside a band the row above is the neighbouring lane: its last
// three results come through one DPP shift by a lane per step (no barrier or
// LDS).  Lane 0 reads the row above the band from `edge`: the band above stores each
// result of its last row there as one 64-bit word, `tag` in the high half
// (relaxed agent-scope atomic stores, which reach every XCD).  Before each
// chunk of kChunk steps the band below polls the words it needs until all
// carry the tag: a word is written whole, so no fence orders data and flag,
// and no L2 is written back or invalidated.  The caller clears `edge` (tag 0)
// before the sweeps and gives each sweep its own non-zero tag.  A wait gives
// up after max_polls polls (status[0] = 1, and every later wait returns at
// once), so a launch always ends; the field is then recomputed by launching
// one band at a time in order (band0 = the band, one workgroup), where every
// wait finds the band above complete.  Inputs are loaded kChunk steps
// ahead into a register ring; every lane issues the same load and store each
// step (clamped addresses, the discard slot).
constexpr uint32_t kBandRows = 64;
constexpr uint32_t kChunk = 32;
constexpr uint32_t kMaxPolls = 1u << 20;

template <bool kMirror>
__global__ __launch_bounds__(64) void k_pad_band_sweep(uint32_t W, uint32_t H, const uint8_t* __restrict__ source,
                                                        uint32_t* __restrict__ field, unsigned long long* edge,
                                                        uint32_t tag, uint32_t* status, uint32_t band0,
                                                        uint32_t max_polls) {
    const uint32_t band = band0 + blockIdx.x, lane = threadIdx.x;
    const uint32_t y = band * kBandRows + lane;  // row in sweep (possibly mirrored) coordinates
    const uint32_t rows = H - band * kBandRows < kBandRows ? H - band * kBandRows : kBandRows;
    const uint32_t steps = W + 2u * (rows - 1u);
    const bool publish = band * kBandRows + kBandRows < H;  // a band below reads this band's last row
    const uint32_t ry = kMirror ? H - 1u - (y < H ? y : H - 1u) : (y < H ? y : H - 1u);
    const Offset none = {kNone16, kNone16};
    unsigned long long* my_edge = edge + static_cast<size_t>(band) * W;
    const unsigned long long* up_edge = edge + static_cast<size_t>(band ? band - 1u : 0u) * W;
    const unsigned long long tag_hi = static_cast<unsigned long long>(tag) << 32;
    // inputs of this lane's texels at the steps of the current and the next chunk
    uint32_t in[kChunk];
    auto load_input = [&](uint32_t x) {
        const uint32_t cx = x < W ? x : W - 1u;
        const uint32_t rx = kMirror ? W - 1u - cx : cx;
        const uint32_t idx = ry * W + rx;
        return kMirror ? *word_at(field, idx) : static_cast<uint32_t>(source[idx]);
    };
#pragma unroll
    for (uint32_t k = 0; k < kChunk; ++k) in[k] = load_input(k - 2u * lane);  // x of step k (wraps: clamped)
    Offset prev = none;                      // this row's previous result (left neighbour)
    Offset up1 = none, up2 = none, up3 = none;  // row above at columns x + 1, x, x - 1
    uint32_t edge_vals = 0u;                 // lanes 0..kChunk+1: row above the band, columns t0 - 1 + lane
    for (uint32_t t0 = 0; t0 < steps; t0 += kChunk) {
        if (band > 0) {
            // columns t0 - 1 .. t0 + kChunk of the row above are needed in this chunk
            const int32_t col = static_cast<int32_t>(t0) - 1 + static_cast<int32_t>(lane);
            const bool want = col >= 0 && col < static_cast<int32_t>(W) && lane < kChunk + 2u;
            unsigned long long v = 0ull;
            uint32_t polls = 0;
            while (true) {
                if (want) v = __hip_atomic_load(&up_edge[col], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__all(!want || (v >> 32) == tag)) break;
                const uint32_t failed = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (++polls >= max_polls || failed != 0u) {
                    if (lane == 0) __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            edge_vals = want ? static_cast<uint32_t>(v) : as_word(none);
            // lane 0: the row above the band at columns t0 + 1, t0, t0 - 1
            const uint32_t e0 = __builtin_amdgcn_readlane(edge_vals, 0), e1 = __builtin_amdgcn_readlane(edge_vals, 1),
                           e2 = __builtin_amdgcn_readlane(edge_vals, 2);
            up1 = lane == 0 ? as_offset(e2) : up1;
            up2 = lane == 0 ? as_offset(e1) : up2;
            up3 = lane == 0 ? as_offset(e0) : up3;
        }
#pragma unroll
        for (uint32_t s = 0; s < kChunk; ++s) {
            const uint32_t t = t0 + s;
            const int32_t x = static_cast<int32_t>(t) - 2 * static_cast<int32_t>(lane);
            const bool settle = x >= 0 && x < static_cast<int32_t>(W) && y < H;
            Offset cur;
            if constexpr (kMirror) {
                cur = -as_offset(in[s]);
            } else {
                const Offset far = {static_cast<int16_t>(kFar), static_cast<int16_t>(kFar)};
                const Offset zero = {0, 0};
                cur = in[s] ? zero : far;
            }
            const bool has_up = y > 0;
            int32_t cur2 = len2(cur);
            relax(cur, cur2, x > 0 ? prev : none, Offset{-1, 0});
            relax(cur, cur2, has_up ? up2 : none, Offset{0, -1});
            relax(cur, cur2, has_up && x > 0 ? up3 : none, Offset{-1, -1});
            relax(cur, cur2, has_up && x + 1 < static_cast<int32_t>(W) ? up1 : none, Offset{1, -1});
            const Offset res = settle ? cur : none;
            const uint32_t rx = kMirror ? W - 1u - static_cast<uint32_t>(x) : static_cast<uint32_t>(x);
            *word_at(field, settle ? ry * W + rx : W * H) = as_word(kMirror ? -cur : cur);
            if (publish && lane == kBandRows - 1u && settle)
                __hip_atomic_store(&my_edge[x], tag_hi | as_word(cur), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the next step's neighbours: own result on the left, the lane above's
            // results above (a DPP shift of the wave by one lane, wave_shr:1), and for
            // lane 0 the row above the band at column x + 2 (edge_vals lane s + 3)
            prev = res;
            up3 = up2;
            up2 = up1;
            const uint32_t feed = __builtin_amdgcn_readlane(edge_vals, static_cast<int>(s + 3));
            up1 = as_offset(static_cast<uint32_t>(
                __builtin_amdgcn_update_dpp(static_cast<int>(feed), static_cast<int>(as_word(res)), 0x138, 0xf, 0xf, false)));
            in[s] = load_input(static_cast<uint32_t>(x + static_cast<int32_t>(kChunk)));
        }
    }
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

uint32_t pad_band_count(uint32_t H) { return (H + kBandRows - 1u) / kBandRows; }

hipError_t pad_field_bands(uint32_t W, uint32_t H, const uint8_t* source, uint32_t* field, unsigned long long* edge,
                           uint32_t* status, uint32_t max_polls, hipStream_t st) {
    const uint32_t bands = pad_band_count(H);
    const uint32_t polls = max_polls ? max_polls : kMaxPolls;
    hipError_t e = fill_u32(status, 0u, 1, st);
    if (e != hipSuccess) return e;
    if ((e = fill_u64(edge, 0ull, static_cast<size_t>(bands) * W, st)) != hipSuccess) return e;  // tag 0: nothing yet
    k_pad_band_sweep<false><<<bands, 64, 0, st>>>(W, H, source, field, edge, 1u, status, 0u, polls);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_pad_band_sweep<true><<<bands, 64, 0, st>>>(W, H, source, field, edge, 2u, status, 0u, polls);
    return hipGetLastError();
}

hipError_t pad_field_bands_ordered(uint32_t W, uint32_t H, const uint8_t* source, uint32_t* field,
                                   unsigned long long* edge, uint32_t* status, hipStream_t st) {
    const uint32_t bands = pad_band_count(H);
    hipError_t e = fill_u32(status, 0u, 1, st);
    if (e != hipSuccess) return e;
    // fresh tags (3, 4): words left by an abandoned parallel sweep never pass for this one's
    for (uint32_t b = 0; b < bands; ++b) {
        k_pad_band_sweep<false><<<1, 64, 0, st>>>(W, H, source, field, edge, 3u, status, b, kMaxPolls);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (uint32_t b = 0; b < bands; ++b) {
        k_pad_band_sweep<true><<<1, 64, 0, st>>>(W, H, source, field, edge, 4u, status, b, kMaxPolls);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
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
