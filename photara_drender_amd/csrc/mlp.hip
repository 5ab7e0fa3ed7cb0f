// Warm-start MLP on gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Per texel: 40 features (8 views x [R, G, B, confidence, cos]) -> 64 -> 64 -> 3,
// GELU (tanh form) on both hidden layers; the output is added to the
// confidence-weighted mean colour of the 8 samples.  Parameters: one float32
// vector W1 [40,64] | b1 [64] | W2 [64,64] | b2 [64] | W3 [64,3] | b3 [3].
// Matrix-core operands are bfloat16 (round to nearest even), accumulation and
// everything else float32.
//
// Layout.  A 16-texel tile is the N axis of every product: layer l computes
// Z_l^T = W_l^T H_{l-1}^T (hidden units on M, texels on N).  With the operand
// layout of the 16x16x32 instruction (lane = 16 q + r: A[m = r][k = 8 q + j],
// B[k = 8 q + j][n = r], D[m = 4 q + i][n = r]), the four D blocks of a 64-unit
// layer leave lane (q, r) holding units {16 mb + 4 q + i} of texel r; blocks
// 2 ks and 2 ks + 1 are exactly a B operand of K step ks when K position
// 8 q + j is read as unit(ks, q, j) = 16 (2 ks + j / 4) + 4 q + j % 4.  Weight
// operands are built once per wave with that K permutation, so activations stay
// in registers from layer to layer.  The backward runs the same chain with W3
// and W2 as A operands; weight gradients sum over texels (K = 32 texels) from
// activations staged texel-contiguous in LDS.
#include "pdr_internal.hpp"

namespace pdr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 40, kHid = 64, kOut = 3;
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid * kOut, kParams = kB3 + kOut;
static_assert(kParams == static_cast<int>(kMlpParams), "parameter layout");
constexpr int kPart = kParams + 2;  // per-wave partials: unscaled gradient, loss sum, valid count
constexpr int kRow = 40;            // LDS row stride (bf16) of the texel-contiguous activation tiles
constexpr uint32_t kInferBlock = 512;

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)

// GELU, tanh form: 0.5 z (1 + tanh u) with u = c (z + 0.044715 z^3), evaluated as
// z * s with s = (1 + tanh u) / 2 = 1 / (1 + e^(-2u)) (hardware exp and
// reciprocal, no cancellation); its derivative with 1 - tanh^2 u = 4 s (1 - s).
__device__ __forceinline__ float gelu_sigmoid(float z, float& one_minus_s) {
    const float u = kGeluC * (z + 0.044715f * z * z * z);
    const float e = __expf(-2.0f * u);
    const float s = __builtin_amdgcn_rcpf(1.0f + e);
    one_minus_s = s < 0.5f ? 1.0f - s : e * s;  // e * s = e / (1 + e), exact where s is near 1
    return s;
}

// GELU for the forward pass: z / (1 + 2^(z (k0 + k1 z^2))) with k0 = -2 c log2 e
// and k1 = 0.044715 k0, the same function as z * gelu_sigmoid(z) in five
// full-rate and two quarter-rate instructions.
__device__ __forceinline__ float gelu(float z) {
    constexpr float k0 = -2.0f * kGeluC * 1.4426950408889634f, k1 = 0.044715f * k0;
    const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(k1, z * z, k0));
    return z * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float gelu_grad(float z) {
    float oms;
    const float s = gelu_sigmoid(z, oms);
    return s + 2.0f * z * s * oms * kGeluC * (1.0f + 3.0f * 0.044715f * z * z);
}

__device__ __forceinline__ int unit(int ks, int q, int j) { return 16 * (2 * ks + j / 4) + 4 * q + (j % 4); }

__device__ __forceinline__ __bf16 bf(float v) { return static_cast<__bf16>(v); }

// Forward weight operands and biases of lane (q, r).
struct FwdOps {
    bf16x8 a1[4][2], a2[4][2], a3[2];
    float b1[4][4], b2[4][4];
};

__device__ void load_forward_ops(const float* __restrict__ w, int q, int r, FwdOps& o) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int in = 32 * ks + 8 * q + j;
                o.a1[mb][ks][j] = bf(in < kIn ? w[kW1 + in * kHid + 16 * mb + r] : 0.0f);
                o.a2[mb][ks][j] = bf(w[kW2 + unit(ks, q, j) * kHid + 16 * mb + r]);
            }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) o.a3[ks][j] = bf(r < kOut ? w[kW3 + unit(ks, q, j) * kOut + r] : 0.0f);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o.b1[mb][i] = w[kB1 + 16 * mb + 4 * q + i];
            o.b2[mb][i] = w[kB2 + 16 * mb + 4 * q + i];
        }
}

// Stages the features of texels t0 .. t0 + 15 in LDS (zero beyond n).
__device__ __forceinline__ void stage_features(const float* __restrict__ x, uint32_t n, uint32_t t0, int lane,
                                               float (*feat)[kIn]) {
    for (int e = lane; e < 16 * (kIn / 4); e += 64) {
        const int tt = e / (kIn / 4), c = e % (kIn / 4);
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (t0 + static_cast<uint32_t>(tt) < n)
            v = reinterpret_cast<const float4*>(x + static_cast<size_t>(t0 + tt) * kIn)[c];
        reinterpret_cast<float4*>(feat[tt])[c] = v;
    }
}

// Feature B operands of texel r of the staged tile (feature index 32 ks + 8 q + j).
__device__ __forceinline__ void feature_operands(const float (*feat)[kIn], int q, int r, bf16x8 b[2]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        b[0][j] = bf(feat[r][8 * q + j]);
        b[1][j] = bf(q == 0 ? feat[r][32 + j] : 0.0f);
    }
}

// Confidence-weighted mean colour of texel r's samples (0 without samples);
// returns the confidence sum.  Lane group q adds views 2 q and 2 q + 1.
__device__ __forceinline__ float base_color(const float (*feat)[kIn], int q, int r, float out[3]) {
    float s = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk) {
        const int k = 2 * q + dk;
        const float w = feat[r][5 * k + 3];
        s += w;
        c0 += feat[r][5 * k] * w;
        c1 += feat[r][5 * k + 1] * w;
        c2 += feat[r][5 * k + 2] * w;
    }
#pragma unroll
    for (int m = 16; m <= 32; m <<= 1) {
        s += __shfl_xor(s, m, 64);
        c0 += __shfl_xor(c0, m, 64);
        c1 += __shfl_xor(c1, m, 64);
        c2 += __shfl_xor(c2, m, 64);
    }
    const bool any = s > 0.0f;
    out[0] = any ? c0 / s : 0.0f;
    out[1] = any ? c1 / s : 0.0f;
    out[2] = any ? c2 / s : 0.0f;
    return s;
}

// The forward operands of every lane of a wave in LDS, lane-contiguous (one
// 16-byte read per lane and operand, no bank conflicts): the inference kernel
// reads them there rather than holding 104 registers of weights per lane,
// which leaves room for more waves to keep feature loads in flight.
struct FwdOpsLds {
    bf16x8 a1[4][2][64], a2[4][2][64], a3[2][64];
    f32x4 b1[4][64], b2[4][64];
};

// The values load_forward_ops gives lane `lane` (= 16 q + r), stored one
// operand at a time so that they never all occupy registers.
__device__ void store_forward_ops(const float* __restrict__ w, int q, int r, int lane, FwdOpsLds& ops) {
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x8 a1, a2;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int in = 32 * ks + 8 * q + j;
                a1[j] = bf(in < kIn ? w[kW1 + in * kHid + 16 * mb + r] : 0.0f);
                a2[j] = bf(w[kW2 + unit(ks, q, j) * kHid + 16 * mb + r]);
            }
            ops.a1[mb][ks][lane] = a1;
            ops.a2[mb][ks][lane] = a2;
        }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        bf16x8 a3;
#pragma unroll
        for (int j = 0; j < 8; ++j) a3[j] = bf(r < kOut ? w[kW3 + unit(ks, q, j) * kOut + r] : 0.0f);
        ops.a3[ks][lane] = a3;
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const int c = 16 * mb + 4 * q;
        ops.b1[mb][lane] = f32x4{w[kB1 + c], w[kB1 + c + 1], w[kB1 + c + 2], w[kB1 + c + 3]};
        ops.b2[mb][lane] = f32x4{w[kB2 + c], w[kB2 + c + 1], w[kB2 + c + 2], w[kB2 + c + 3]};
    }
}

// Operand accessors of forward_tile: registers (FwdOps) or LDS (FwdOpsLds, lane `l`).
struct RegOps {
    const FwdOps& o;
    __device__ bf16x8 a1(int mb, int ks) const { return o.a1[mb][ks]; }
    __device__ bf16x8 a2(int mb, int ks) const { return o.a2[mb][ks]; }
    __device__ bf16x8 a3(int ks) const { return o.a3[ks]; }
    __device__ f32x4 b1(int mb) const { return f32x4{o.b1[mb][0], o.b1[mb][1], o.b1[mb][2], o.b1[mb][3]}; }
    __device__ f32x4 b2(int mb) const { return f32x4{o.b2[mb][0], o.b2[mb][1], o.b2[mb][2], o.b2[mb][3]}; }
};
struct LdsOps {
    const FwdOpsLds& o;
    int l;
    __device__ bf16x8 a1(int mb, int ks) const { return o.a1[mb][ks][l]; }
    __device__ bf16x8 a2(int mb, int ks) const { return o.a2[mb][ks][l]; }
    __device__ bf16x8 a3(int ks) const { return o.a3[ks][l]; }
    __device__ f32x4 b1(int mb) const { return o.b1[mb][l]; }
    __device__ f32x4 b2(int mb) const { return o.b2[mb][l]; }
};

// Forward of one tile: pre-activations z1, z2 (D layout; the biases enter as
// the first accumulator) and the network output y (lanes q = 0 hold outputs
// i < 3 of texel r, without b3 and base colour).
template <class Ops>
__device__ __forceinline__ void forward_tile(const Ops& o, const bf16x8 xb[2], f32x4 z1[4], f32x4 z2[4], f32x4& y) {
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) z1[mb] = mfma(o.a1(mb, 1), xb[1], mfma(o.a1(mb, 0), xb[0], o.b1(mb)));
    bf16x8 h1[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) h1[ks][j] = bf(gelu(z1[2 * ks + j / 4][j % 4]));
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) z2[mb] = mfma(o.a2(mb, 1), h1[1], mfma(o.a2(mb, 0), h1[0], o.b2(mb)));
    bf16x8 h2[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) h2[ks][j] = bf(gelu(z2[2 * ks + j / 4][j % 4]));
    y = mfma(o.a3(1), h2[1], mfma(o.a3(0), h2[0], zero));
}

// Inference: out [n, 3] = network + base colour for texels with samples, 0 for
// texels without.  One 16-texel tile per wave per iteration.
__global__ __launch_bounds__(kInferBlock) void k_mlp_infer(const float* __restrict__ x, uint32_t n,
                                                           const float* __restrict__ w, float* __restrict__ out) {
    __shared__ __align__(16) float feat_all[kInferBlock / 64][16][kIn];
    __shared__ FwdOpsLds ops;
    const int lane = static_cast<int>(threadIdx.x & 63u), q = lane >> 4, r = lane & 15;
    float (*feat)[kIn] = feat_all[threadIdx.x >> 6];
    if (threadIdx.x < 64) store_forward_ops(w, q, r, lane, ops);  // the lane layout is the same in every wave
    __syncthreads();
    const LdsOps o{ops, lane};
    const float b3[3] = {w[kB3], w[kB3 + 1], w[kB3 + 2]};
    const uint32_t waves = gridDim.x * (kInferBlock / 64);
    const uint32_t tiles = (n + 15u) / 16u;
    // the lane's share of a tile's features (16 texels x kIn / 4 float4), loaded a tile ahead
    constexpr int kShare = (16 * (kIn / 4) + 63) / 64;
    auto load_share = [&](uint32_t t0, float4 (&v)[kShare]) {
#pragma unroll
        for (int j = 0; j < kShare; ++j) {
            const int e = lane + 64 * j, tt = e / (kIn / 4), c = e % (kIn / 4);
            v[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (e < 16 * (kIn / 4) && t0 + static_cast<uint32_t>(tt) < n)
                v[j] = reinterpret_cast<const float4*>(x + static_cast<size_t>(t0 + tt) * kIn)[c];
        }
    };
    float4 next[kShare];
    uint32_t tile = blockIdx.x * (kInferBlock / 64) + (threadIdx.x >> 6);
    if (tile < tiles) load_share(tile * 16u, next);
    for (; tile < tiles; tile += waves) {
        const uint32_t t0 = tile * 16u, t = t0 + static_cast<uint32_t>(r);
        wave_sync();  // the previous tile's LDS reads are done
#pragma unroll
        for (int j = 0; j < kShare; ++j) {
            const int e = lane + 64 * j;
            if (e < 16 * (kIn / 4)) reinterpret_cast<float4*>(feat[e / (kIn / 4)])[e % (kIn / 4)] = next[j];
        }
        wave_sync();
        if (tile + waves < tiles) load_share((tile + waves) * 16u, next);
        bf16x8 xb[2];
        feature_operands(feat, q, r, xb);
        float base[3];
        const float s = base_color(feat, q, r, base);
        f32x4 z1[4], z2[4], y;
        forward_tile(o, xb, z1, z2, y);
        if (q == 0 && t < n) {
#pragma unroll
            for (int i = 0; i < 3; ++i) out[static_cast<size_t>(t) * 3 + i] = s > 0.0f ? y[i] + b3[i] + base[i] : 0.0f;
        }
    }
}

// Training gradient: per wave, partial sums over its 32-texel chunks of the
// gradient of sum_t valid_t * |y_t - target_t|^2 (the output gradient 2 d is
// rounded to bfloat16 where it feeds matrix cores), the squared error and the
// valid count, written to part [wave][kPart].  Block = one wave.
struct TrainLds {
    float feat[16][kIn];
    __bf16 x[48][kRow];  // features, rows 40..47 zero
    __bf16 h1[kHid][kRow], h2[kHid][kRow], dz1[kHid][kRow], dz2[kHid][kRow];
    __bf16 dy[16][kRow];  // rows 3..15 zero
};

__device__ __forceinline__ bf16x8 lds_row8(const __bf16 (*t)[kRow], int row, int q) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = t[row][8 * q + j];
    return v;
}

__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ x, const float* __restrict__ target,
                                                 const float* __restrict__ valid, uint32_t n,
                                                 const float* __restrict__ w, float* __restrict__ part) {
    __shared__ __align__(16) TrainLds L;
    const int lane = static_cast<int>(threadIdx.x), q = lane >> 4, r = lane & 15;
    FwdOps o;
    load_forward_ops(w, q, r, o);
    const float b3[3] = {w[kB3], w[kB3 + 1], w[kB3 + 2]};
    // backward A operands: W3 [hidden x output (K)], W2 [input unit x output unit (K, permuted)]
    bf16x8 g3[4], g2[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = 8 * q + j;
            g3[mb][j] = bf(c < kOut ? w[kW3 + (16 * mb + r) * kOut + c] : 0.0f);
            g2[mb][0][j] = bf(w[kW2 + (16 * mb + r) * kHid + unit(0, q, j)]);
            g2[mb][1][j] = bf(w[kW2 + (16 * mb + r) * kHid + unit(1, q, j)]);
        }
    }
    for (int i = lane; i < 8 * kRow; i += 64) L.x[40 + i / kRow][i % kRow] = bf(0.0f);
    for (int i = lane; i < 13 * kRow; i += 64) L.dy[3 + i / kRow][i % kRow] = bf(0.0f);
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 gw1[3][4], gw2[4][4], gw3[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw3[a] = zero;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gw2[a][b] = zero;
            if (a < 3) gw1[a][b] = zero;
        }
    }
    float gb1[4][4], gb2[4][4], gb3[3] = {0.0f, 0.0f, 0.0f}, loss = 0.0f, count = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) gb1[a][i] = gb2[a][i] = 0.0f;
    const uint32_t chunks = (n + 31u) / 32u;
    for (uint32_t chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {
        for (int half = 0; half < 2; ++half) {
            const int col = 16 * half + r;  // texel column in the chunk
            const uint32_t t0 = chunk * 32u + 16u * static_cast<uint32_t>(half);
            const uint32_t t = t0 + static_cast<uint32_t>(r);
            wave_sync();  // earlier reads of the staged tiles are done
            stage_features(x, n, t0, lane, L.feat);
            wave_sync();
            bf16x8 xb[2];
            feature_operands(L.feat, q, r, xb);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                L.x[8 * q + j][col] = xb[0][j];
                if (q == 0) L.x[32 + j][col] = xb[1][j];
            }
            float base[3];
            base_color(L.feat, q, r, base);
            f32x4 z1[4], z2[4], y;
            forward_tile(RegOps{o}, xb, z1, z2, y);
            // output gradient 2 d (lanes q = 0 hold outputs 0..2 of texel t)
            float dyv[3] = {0.0f, 0.0f, 0.0f};
            if (q == 0 && t < n) {
                const float v = valid ? valid[t] : 1.0f;
                count += v;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float d = (y[i] + b3[i] + base[i] - target[static_cast<size_t>(t) * 3 + i]) * v;
                    loss += d * d;
                    dyv[i] = 2.0f * d;
                    gb3[i] += dyv[i];
                }
            }
            bf16x8 dyb;  // dY^T as a B operand: B[k = output][n = texel], nonzero in lanes q = 0
#pragma unroll
            for (int j = 0; j < 8; ++j) dyb[j] = bf(j < 3 ? dyv[j % 3] : 0.0f);
            if (q == 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) L.dy[c][col] = dyb[c];
            // dZ2 = (W3 dY) * gelu'(z2)
            f32x4 dz2[4];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const f32x4 d = mfma(g3[mb], dyb, zero);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int u = 16 * mb + 4 * q + i;
                    dz2[mb][i] = d[i] * gelu_grad(z2[mb][i]);
                    gb2[mb][i] += dz2[mb][i];
                    L.h1[u][col] = bf(gelu(z1[mb][i]));
                    L.h2[u][col] = bf(gelu(z2[mb][i]));
                    L.dz2[u][col] = bf(dz2[mb][i]);
                }
            }
            bf16x8 dz2b[2];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int j = 0; j < 8; ++j) dz2b[ks][j] = bf(dz2[2 * ks + j / 4][j % 4]);
            // dZ1 = (W2 dZ2) * gelu'(z1)
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const f32x4 d = mfma(g2[mb][1], dz2b[1], mfma(g2[mb][0], dz2b[0], zero));
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float v = d[i] * gelu_grad(z1[mb][i]);
                    gb1[mb][i] += v;
                    L.dz1[16 * mb + 4 * q + i][col] = bf(v);
                }
            }
        }
        wave_sync();
        // weight gradients over the chunk's 32 texels: dW[a][b] = sum_t act[t][a] grad[t][b]
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 gz1 = lds_row8(L.dz1, 16 * nb + r, q), gz2 = lds_row8(L.dz2, 16 * nb + r, q);
#pragma unroll
            for (int mb = 0; mb < 3; ++mb) gw1[mb][nb] = mfma(lds_row8(L.x, 16 * mb + r, q), gz1, gw1[mb][nb]);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) gw2[mb][nb] = mfma(lds_row8(L.h1, 16 * mb + r, q), gz2, gw2[mb][nb]);
        }
        const bf16x8 gy = lds_row8(L.dy, r, q);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) gw3[mb] = mfma(lds_row8(L.h2, 16 * mb + r, q), gy, gw3[mb]);
    }
    // this wave's partials; weight gradients from the D layout (lane holds [m = 16 mb + 4 q + i][n = 16 nb + r])
    float* p = part + static_cast<size_t>(blockIdx.x) * kPart;
#pragma unroll
    for (int mb = 0; mb < 3; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int in = 16 * mb + 4 * q + i;
                if (in < kIn) p[kW1 + in * kHid + 16 * nb + r] = gw1[mb][nb][i];
            }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i) p[kW2 + (16 * mb + 4 * q + i) * kHid + 16 * nb + r] = gw2[mb][nb][i];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (r < kOut) p[kW3 + (16 * mb + 4 * q + i) * kOut + r] = gw3[mb][i];
    // bias gradients: sums over the 16 texel lanes of each lane group
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float a = gb1[mb][i], b = gb2[mb][i];
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) {
                a += __shfl_xor(a, m, 64);
                b += __shfl_xor(b, m, 64);
            }
            if (r == 0) {
                p[kB1 + 16 * mb + 4 * q + i] = a;
                p[kB2 + 16 * mb + 4 * q + i] = b;
            }
        }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        gb3[0] += __shfl_xor(gb3[0], m, 64);
        gb3[1] += __shfl_xor(gb3[1], m, 64);
        gb3[2] += __shfl_xor(gb3[2], m, 64);
        loss += __shfl_xor(loss, m, 64);
        count += __shfl_xor(count, m, 64);
    }
    if (lane == 0) {
        p[kB3] = gb3[0];
        p[kB3 + 1] = gb3[1];
        p[kB3 + 2] = gb3[2];
        p[kParams] = loss;
        p[kParams + 1] = count;
    }
}

// Sums the per-wave partials in wave order and scales by 1 / (3 max(count, 1)):
// grad [kParams] and loss_count [2] = (loss, valid count).  With params set,
// also one Adam step on the float32 parameters (bias corrections
// b1c = 1 - beta1^t, b2c = 1 - beta2^t).
// Sums the per-wave partials (fixed order: kReduceSegments contiguous wave
// ranges, each in order, then the ranges in order), then Adam.  A block takes
// 64 consecutive entries of the partial rows.
constexpr uint32_t kReduceSegments = 16;
__global__ __launch_bounds__(64 * kReduceSegments) void k_mlp_reduce(
    const float* __restrict__ part, uint32_t waves, float* __restrict__ grad, float* __restrict__ loss_count,
    float* __restrict__ params, float* __restrict__ m1, float* __restrict__ m2, float lr, float beta1, float beta2,
    float eps, float b1c, float b2c) {
    __shared__ float seg_s[kReduceSegments][64], seg_c[kReduceSegments][64];
    const uint32_t lane = threadIdx.x & 63u, seg = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 64u + lane;
    const uint32_t per = (waves + kReduceSegments - 1) / kReduceSegments;
    const uint32_t k0 = seg * per, k1 = k0 + per < waves ? k0 + per : waves;
    float s = 0.0f, c = 0.0f;
    if (i < static_cast<uint32_t>(kPart))
        for (uint32_t k = k0; k < k1; ++k) {
            s += part[static_cast<size_t>(k) * kPart + i];
            c += part[static_cast<size_t>(k) * kPart + kParams + 1];
        }
    seg_s[seg][lane] = s;
    seg_c[seg][lane] = c;
    __syncthreads();
    if (seg != 0 || i >= static_cast<uint32_t>(kPart)) return;
    s = 0.0f;
    c = 0.0f;
    for (uint32_t g = 0; g < kReduceSegments; ++g) {
        s += seg_s[g][lane];
        c += seg_c[g][lane];
    }
    const float inv = 1.0f / (3.0f * (c > 1.0f ? c : 1.0f));
    if (i >= static_cast<uint32_t>(kParams)) {
        if (loss_count) loss_count[i - kParams] = i == static_cast<uint32_t>(kParams) ? s * inv : c;
        return;
    }
    const float g = s * inv;
    if (grad) grad[i] = g;
    if (params) {
        const float a = beta1 * m1[i] + (1.0f - beta1) * g;
        const float b = beta2 * m2[i] + (1.0f - beta2) * g * g;
        m1[i] = a;
        m2[i] = b;
        params[i] -= lr * (a / b1c) / (sqrtf(b / b2c) + eps);
    }
}

}  // namespace

uint32_t mlp_grad_waves(uint64_t n) {
    const uint64_t chunks = (n + 31u) / 32u;
    return static_cast<uint32_t>(chunks < kMlpMaxWaves ? (chunks ? chunks : 1u) : kMlpMaxWaves);
}

hipError_t mlp_infer(const float* x, uint32_t n, const float* w, float* out, hipStream_t st) {
    const uint32_t tiles = (n + 15u) / 16u;
    uint32_t blocks = (tiles + kInferBlock / 64 - 1) / (kInferBlock / 64);
    // two resident 8-wave blocks per CU (93 VGPRs, 46 KB of LDS each): the
    // operands are staged in LDS once per block, the tiles strided over the grid
    blocks = blocks < 512u ? (blocks ? blocks : 1u) : 512u;
    k_mlp_infer<<<blocks, kInferBlock, 0, st>>>(x, n, w, out);
    return hipGetLastError();
}

hipError_t mlp_grad(const float* x, const float* target, const float* valid, uint32_t n, const float* w, float* part,
                    uint32_t waves, hipStream_t st) {
    k_mlp_grad<<<waves, 64, 0, st>>>(x, target, valid, n, w, part);
    return hipGetLastError();
}

hipError_t mlp_reduce(const float* part, uint32_t waves, float* grad, float* loss_count, float* params, float* m1,
                      float* m2, float lr, float beta1, float beta2, float eps, float b1c, float b2c, hipStream_t st) {
    k_mlp_reduce<<<(kPart + 63) / 64, 64 * kReduceSegments, 0, st>>>(part, waves, grad, loss_count, params, m1, m2, lr,
                                                                     beta1, beta2, eps, b1c, b2c);
    return hipGetLastError();
}

}  // namespace pdr
