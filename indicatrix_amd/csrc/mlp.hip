// mlp.hip -- the metric surrogate: a BF16 MLP trained and evaluated with
// hand-written MFMA kernels on CDNA4 (gfx950).
//
// Network: in (padded to 64) -> 4 x [256, GELU] -> out (padded to 64).
// Weights and activations are BF16, every product accumulates in FP32 on the
// matrix cores (v_mfma_f32_32x32x16_bf16, gfx950's double-K BF16 MFMA),
// master weights and Adam moments are FP32.  The reference has no
// surrogate; this is new work whose judge is the path tracer's own metric
// evaluator.
//
// A training step is four launches:
//   forward   one block per 16 batch rows (eight waves) runs all five layers
//             (v_mfma_f32_16x16x32_bf16, two 16 x 16 tiles per wave), the
//             activations ping-ponging in LDS, the weights streamed from L2
//             as 16-byte MFMA operands; bias, GELU and GELU' (one Gaussian
//             CDF for both) fused in the epilogue, which also stores the
//             activations batch-transposed; the output layer's epilogue
//             forms the MSE loss and dY (BF16);
//   backward  one block per 16 rows runs the dZ chain dZ_l = (dZ_{l+1} W_{l+1})
//             * GELU'(Z_l) through LDS against a transposed BF16 copy of the
//             weights (contiguous MFMA operands), GELU' read back, not
//             recomputed;
//   dW        every layer's dW = dZ^T A and db = dZ^T 1 as 64 x 64 tiles
//             (v_mfma_f32_32x32x16_bf16) split over the batch (split-K, one
//             launch for all five layers, 73 tiles x <= 16 splits), both
//             operands batch-contiguous 16-byte loads (no LDS);
//   reduce    the split partials added in a fixed order (deterministic),
//             then Adam, writing the FP32 master, the BF16 working copy and
//             its transpose.
// Memory layout: weights [out x in] (and [in x out] transposed); the
// training tensors the dW GEMM reads are batch-transposed [width x batch];
// inputs, outputs and dY row-major [batch x width].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_dev.h"

namespace ixg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 16, kInP = 64, kH = 256, kL = 4, kOut = 3, kOutP = 64;
constexpr int kRows = 16;       // batch rows per forward / backward block
constexpr int kPitch = kH + 8;  // LDS row pitch (bf16), 16-byte multiple, staggers banks
constexpr size_t kW0 = (size_t)kH * kInP, kWH = (size_t)kH * kH, kW4 = (size_t)kOutP * kH;
constexpr size_t kNWeights = kW0 + 3 * kWH + kW4;
constexpr size_t kNBias = (size_t)kH * kL + kOutP;
constexpr size_t kNParams = kNWeights + kNBias;
constexpr int kMaxSplits = 16;  // dW split-K partial sets

__host__ __device__ constexpr size_t w_off(int l) { return l == 0 ? 0 : kW0 + (size_t)(l - 1) * kWH; }
__host__ __device__ constexpr int w_rows(int l) { return l == kL ? kOutP : kH; }
__host__ __device__ constexpr int w_cols(int l) { return l == 0 ? kInP : kH; }
__host__ __device__ constexpr size_t b_off(int l) { return kNWeights + (size_t)l * kH; }

// FP32 -> BF16, round to nearest even: gfx950's v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// The Gaussian CDF Phi(v) = (1 + erf(v / sqrt 2)) / 2 with erf from
// Abramowitz & Stegun 7.1.28 (Hastings): erf(x) = 1 - (1 + a1 x + ... + a6 x^6)^-16
// for x >= 0, |error| <= 3e-7 in exact arithmetic and about 2e-6 in f32 (the
// 16th power scales the polynomial's rounding), far below the BF16 rounding
// the activations receive.
__device__ __forceinline__ float gauss_cdf(float v) {
    const float x = fabsf(v) * 0.70710678118654752f;
    float p = fmaf(x, 0.0000430638f, 0.0002765672f);
    p = fmaf(x, p, 0.0001520143f);
    p = fmaf(x, p, 0.0092705272f);
    p = fmaf(x, p, 0.0422820123f);
    p = fmaf(x, p, 0.0705230784f);
    p = fmaf(x, p, 1.0f);
    p = p * p;
    p = p * p;
    p = p * p;
    p = p * p;  // the 16th power; inf beyond x ~ 25 gives erf = 1
    const float erf_abs = 1.0f - __builtin_amdgcn_rcpf(p);
    return 0.5f + copysignf(0.5f * erf_abs, v);
}
// GELU(v) = v Phi(v) and GELU'(v) = Phi(v) + v phi(v), phi the Gaussian density.
__device__ __forceinline__ float gelu_of(float v, float cdf) { return v * cdf; }
__device__ __forceinline__ float gelu_grad_of(float v, float cdf) {
    return fmaf(v, 0.39894228040143268f * __expf(-0.5f * v * v), cdf);
}

// Eight consecutive bf16 (16-byte aligned) as an MFMA operand.
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

// v_mfma_f32_16x16x32_bf16: lane l supplies A row l % 16 and B column l % 16,
// k = 8 (l / 16) + 0..7, and holds D column l % 16, rows 4 (l / 16) + 0..3.
__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// v_mfma_f32_32x32x16_bf16: lane l supplies A row l % 32 and B column l % 32,
// k = 8 (l / 32) + 0..7, and holds D column l % 32, rows 8 (i / 4) + 4 (l / 32) + i % 4.
__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int acc32_row(int i, int lane) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }

// A layer's weight operands for one wave, loaded ahead of use: NT column
// tiles of 16 (from n0) x KS k steps of 32 of W row-major [N x K].
template <int KS, int NT>
__device__ __forceinline__ void load_w(const uint16_t* __restrict__ w, int K, int n0, int lane,
                                       bf16x8 (&wr)[2][8]) {
    const uint16_t* p = w + (size_t)(n0 + (lane & 15)) * K + 8 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int s = 0; s < KS; ++s) wr[t][s] = ld8(p + (size_t)16 * t * K + 32 * s);
}

// RG x NT tiles of 16 x 16: acc[g][t] = A_lds[16 g + .., :32 KS] . wr[t][:]:
// each weight operand feeds the RG row groups, each activation operand the
// NT column tiles.
template <int RG, int KS, int NT>
__device__ __forceinline__ void mma_w(const uint16_t* __restrict__ a_lds, int lane, const bf16x8 (&wr)[2][8],
                                      f32x4 (&acc)[RG][2]) {
    const uint16_t* ar = a_lds + (lane & 15) * kPitch + 8 * (lane >> 4);
#pragma unroll
    for (int g = 0; g < RG; ++g)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const bf16x8 a = ld8(ar + g * 16 * kPitch + 32 * s);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[g][t] = mfma16(a, wr[t][s], acc[g][t]);
        }
}

// Streaming variant for the inference blocks: RG x NT tiles, the weight
// operands loaded inside the k loop (RG row groups share each of them).
template <int RG, int NT>
__device__ __forceinline__ void tiles16(const uint16_t* __restrict__ a_lds, const uint16_t* __restrict__ w, int K,
                                        int n0, int lane, f32x4 (&acc)[RG][2]) {
    const uint16_t* ar = a_lds + (lane & 15) * kPitch + 8 * (lane >> 4);
    const uint16_t* wr = w + (size_t)(n0 + (lane & 15)) * K + 8 * (lane >> 4);
#pragma unroll
    for (int g = 0; g < RG; ++g)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for (int k = 0; k < K; k += 32) {
        bf16x8 b[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) b[t] = ld8(wr + (size_t)16 * t * K + k);
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const bf16x8 a = ld8(ar + g * 16 * kPitch + k);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[g][t] = mfma16(a, b[t], acc[g][t]);
        }
    }
}

// Four consecutive bf16 / f32 of one transposed row (the 16 x 16 accumulator
// rows a lane holds are consecutive batch rows).
__device__ __forceinline__ void st4(uint16_t* p, const uint16_t (&h)[4]) {
    uint2 v;
    v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    *reinterpret_cast<uint2*>(p) = v;
}

// Hidden layer l's epilogue for one wave's two 16-column tiles: bias, GELU
// into the next LDS activation buffer and, when training, the BF16
// activation and GELU'(Z) stored batch-transposed.
template <bool TRAIN, int RG>
__device__ __forceinline__ void hidden_epilogue(int l, const f32x4 (&acc)[RG][2], const float* __restrict__ bias,
                                                uint16_t* __restrict__ out, int wave, int lane, int B, int r0,
                                                uint16_t* __restrict__ actT, float* __restrict__ dgeluT) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = wave * 32 + 16 * j + (lane & 15);
        const float b = bias[l * kH + n];
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const int m0 = 16 * g + 4 * (lane >> 4);
            uint16_t a4[4];
            float d4[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float v = acc[g][j][i] + b;
                const float cdf = gauss_cdf(v);
                a4[i] = f2bf(gelu_of(v, cdf));
                out[(m0 + i) * kPitch + n] = a4[i];
                if (TRAIN) d4[i] = gelu_grad_of(v, cdf);
            }
            if (TRAIN) {
                const size_t o = ((size_t)l * kH + n) * B + r0 + m0;
                st4(actT + o, a4);
                *reinterpret_cast<float4*>(dgeluT + o) = make_float4(d4[0], d4[1], d4[2], d4[3]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// forward (+ loss and dY when training): one block per 16 RG batch rows,
// eight waves, wave w owns 32 of the 256 hidden columns (waves 0..3 one
// 16-column tile of the 64 output columns).  Training also stores, batch-
// transposed for the dW GEMM, the BF16 inputs and activations and dY, and
// GELU'(Z) in FP32 for the backward pass.
// ---------------------------------------------------------------------------
template <bool TRAIN, int RG>
__global__ __launch_bounds__(512) void mlp_forward_kernel(
    const float* __restrict__ x, const float* __restrict__ t, const uint16_t* __restrict__ w_bf,
    const float* __restrict__ params, int B, float* __restrict__ y, float* __restrict__ y3,
    uint16_t* __restrict__ act_inT, uint16_t* __restrict__ actT, float* __restrict__ dgeluT,
    uint16_t* __restrict__ dy, uint16_t* __restrict__ dyT, float* __restrict__ loss_part) {
    constexpr int R = kRows * RG;
    __shared__ __align__(16) uint16_t sA[2][R * kPitch];
    __shared__ float s_loss[8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * R;
    const float* bias = params + kNWeights;
    // inputs: f32 [B x 16] -> bf16 [R x 64], zero-padded
    for (int e = tid; e < R * kInP; e += 512) {
        const int m = e / kInP, c = e % kInP;
        const uint16_t v = c < kIn ? f2bf(x[(size_t)(r0 + m) * kIn + c]) : (uint16_t)0;
        sA[0][m * kPitch + c] = v;
        if (TRAIN) act_inT[(size_t)c * B + r0 + m] = v;
    }
    f32x4 acc[RG][2];
    if constexpr (!TRAIN) {
        // inference: the weights stream through the k loops, each operand used by RG row groups
        __syncthreads();
        for (int l = 0; l < kL; ++l) {
            tiles16<RG, 2>(sA[l & 1], w_bf + w_off(l), w_cols(l), wave * 32, lane, acc);
            hidden_epilogue<false, RG>(l, acc, bias, sA[(l + 1) & 1], wave, lane, B, r0, actT, dgeluT);
            __syncthreads();
        }
        if (wave < 4) {
            tiles16<RG, 1>(sA[kL & 1], w_bf + w_off(kL), kH, wave * 16, lane, acc);
            const int n = wave * 16 + (lane & 15);
            const float b = bias[kL * kH + n];
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const size_t row = (size_t)(r0 + 16 * g + 4 * (lane >> 4) + i);
                    const float v = acc[g][0][i] + b;
                    if (!y3) y[row * kOutP + n] = v;
                    else if (n < kOut) y3[row * kOut + n] = v;
                }
        }
        return;
    }
    // training: layer l's MFMAs run on operands loaded during layer l-1's epilogue
    bf16x8 wa[2][8], wb[2][8];
    load_w<2, 2>(w_bf + w_off(0), kInP, wave * 32, lane, wa);
    __syncthreads();
    mma_w<RG, 2, 2>(sA[0], lane, wa, acc);
    load_w<8, 2>(w_bf + w_off(1), kH, wave * 32, lane, wb);
    hidden_epilogue<TRAIN, RG>(0, acc, bias, sA[1], wave, lane, B, r0, actT, dgeluT);
    __syncthreads();
    mma_w<RG, 8, 2>(sA[1], lane, wb, acc);
    load_w<8, 2>(w_bf + w_off(2), kH, wave * 32, lane, wa);
    hidden_epilogue<TRAIN, RG>(1, acc, bias, sA[0], wave, lane, B, r0, actT, dgeluT);
    __syncthreads();
    mma_w<RG, 8, 2>(sA[0], lane, wa, acc);
    load_w<8, 2>(w_bf + w_off(3), kH, wave * 32, lane, wb);
    hidden_epilogue<TRAIN, RG>(2, acc, bias, sA[1], wave, lane, B, r0, actT, dgeluT);
    __syncthreads();
    mma_w<RG, 8, 2>(sA[1], lane, wb, acc);
    if (wave < 4) load_w<8, 1>(w_bf + w_off(kL), kH, wave * 16, lane, wa);
    hidden_epilogue<TRAIN, RG>(3, acc, bias, sA[0], wave, lane, B, r0, actT, dgeluT);
    __syncthreads();
    // output layer: 64 columns, one 16-column tile for each of waves 0..3
    float lsum = 0.0f;
    if (wave < 4) {
        mma_w<RG, 8, 1>(sA[0], lane, wa, acc);
        const int n = wave * 16 + (lane & 15);
        const float b = bias[kL * kH + n];
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const int m0 = 16 * g + 4 * (lane >> 4);
            uint16_t g4[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const size_t row = (size_t)(r0 + m0 + i);
                const float v = acc[g][0][i] + b;
                if (TRAIN || !y3) y[row * kOutP + n] = v;
                if (!TRAIN && y3 && n < kOut) y3[row * kOut + n] = v;
                if (TRAIN) {
                    float gr = 0.0f;
                    if (n < kOut) {
                        const float d = v - t[row * kOut + n];
                        lsum += d * d;
                        gr = 2.0f * d / (float)(B * kOut);
                    }
                    g4[i] = f2bf(gr);
                    dy[row * kOutP + n] = g4[i];
                }
            }
            if (TRAIN) st4(dyT + (size_t)n * B + r0 + m0, g4);
        }
    }
    if (TRAIN) {
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (lane == 0) s_loss[wave] = lsum;
        __syncthreads();
        if (tid == 0) loss_part[blockIdx.x] = (s_loss[0] + s_loss[1]) + (s_loss[2] + s_loss[3]);
    }
}

// Backward step j's epilogue for one wave's two tiles: dZ_j = G * GELU'(Z_j)
// in BF16, into the next LDS buffer and batch-transposed to memory.
__device__ __forceinline__ void backward_epilogue(int j, const f32x4 (&acc)[1][2], const float* __restrict__ dgeluT,
                                                  uint16_t* __restrict__ out, int wave, int lane, int B, int r0,
                                                  uint16_t* __restrict__ dzT) {
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
        const int k = wave * 32 + 16 * jj + (lane & 15);
        const int m0 = 4 * (lane >> 4);
        const size_t o = ((size_t)j * kH + k) * B + r0 + m0;
        const float4 dg = *reinterpret_cast<const float4*>(dgeluT + o);
        const float dgv[4] = {dg.x, dg.y, dg.z, dg.w};
        uint16_t d4[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            d4[i] = f2bf(acc[0][jj][i] * dgv[i]);
            out[(m0 + i) * kPitch + k] = d4[i];
        }
        st4(dzT + o, d4);
    }
}

// ---------------------------------------------------------------------------
// backward: the dZ chain dZ_j = (dZ_{j+1} W_{j+1}) * GELU'(Z_j), one block
// per 16 rows, against the transposed weights (G[m][k] = dZ[m][:] . WT[k][:]),
// the BF16 gradients stored batch-transposed for the dW GEMM
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void mlp_backward_kernel(const uint16_t* __restrict__ dy,
                                                           const uint16_t* __restrict__ w_bfT,
                                                           const float* __restrict__ dgeluT, int B,
                                                           uint16_t* __restrict__ dzT) {
    __shared__ __align__(16) uint16_t sG[2][kRows * kPitch];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * kRows;
    if (tid < kRows * kOutP / 8) {  // dY [16 x 64], 16 B per thread
        const int m = tid / (kOutP / 8), c8 = (tid % (kOutP / 8)) * 8;
        *reinterpret_cast<u32x4*>(&sG[0][m * kPitch + c8]) =
            *reinterpret_cast<const u32x4*>(dy + (size_t)(r0 + m) * kOutP + c8);
    }
    // step j's MFMAs run on operands loaded during step j+1's epilogue
    bf16x8 wa[2][8], wb[2][8];
    f32x4 acc[1][2];
    load_w<2, 2>(w_bfT + w_off(kL), kOutP, wave * 32, lane, wa);  // WT_4 [256 x 64]
    __syncthreads();
    mma_w<1, 2, 2>(sG[0], lane, wa, acc);
    load_w<8, 2>(w_bfT + w_off(3), kH, wave * 32, lane, wb);
    backward_epilogue(3, acc, dgeluT, sG[1], wave, lane, B, r0, dzT);
    __syncthreads();
    mma_w<1, 8, 2>(sG[1], lane, wb, acc);
    load_w<8, 2>(w_bfT + w_off(2), kH, wave * 32, lane, wa);
    backward_epilogue(2, acc, dgeluT, sG[0], wave, lane, B, r0, dzT);
    __syncthreads();
    mma_w<1, 8, 2>(sG[0], lane, wa, acc);
    load_w<8, 2>(w_bfT + w_off(1), kH, wave * 32, lane, wb);
    backward_epilogue(1, acc, dgeluT, sG[1], wave, lane, B, r0, dzT);
    __syncthreads();
    mma_w<1, 8, 2>(sG[1], lane, wb, acc);
    backward_epilogue(0, acc, dgeluT, sG[0], wave, lane, B, r0, dzT);
}

// ---------------------------------------------------------------------------
// dW = dZ^T A and db = dZ^T 1 for every layer: 64 x 64 tiles split over the
// batch, both operands batch-contiguous (the transposed stores above), read
// straight from L2 as 16-byte MFMA operands; the bias tiles multiply by ones
// ---------------------------------------------------------------------------
constexpr int kDwTile = 64, kDwStep = 64;

__host__ __device__ constexpr int dw_weight_tiles(int l) { return (w_rows(l) / kDwTile) * (w_cols(l) / kDwTile); }
__host__ __device__ constexpr int dw_tiles() {
    int t = 0;
    for (int q = 0; q <= kL; ++q) t += dw_weight_tiles(q) + w_rows(q) / kDwTile;
    return t;
}

// tile t: the weight tiles of layers 0..4, then the bias tiles of layers 0..4
__device__ __forceinline__ void dw_tile_of(int t, int& l, int& tn, int& tk, bool& bias) {
    bias = false;
    l = 0;
    tn = tk = 0;
#pragma unroll
    for (int q = 0; q <= kL; ++q) {
        const int nt = dw_weight_tiles(q);
        if (t < nt) {
            l = q;
            tn = t / (w_cols(q) / kDwTile);
            tk = t % (w_cols(q) / kDwTile);
            return;
        }
        t -= nt;
    }
    bias = true;
#pragma unroll
    for (int q = 0; q <= kL; ++q) {
        const int nt = w_rows(q) / kDwTile;
        if (t < nt) {
            l = q;
            tn = t;
            return;
        }
        t -= nt;
    }
}

__global__ __launch_bounds__(256) void mlp_dw_kernel(const uint16_t* __restrict__ dyT,
                                                     const uint16_t* __restrict__ dzT,
                                                     const uint16_t* __restrict__ act_inT,
                                                     const uint16_t* __restrict__ actT, int B, int split_rows,
                                                     float* __restrict__ part) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int l, tn, tk;
    bool bias;
    dw_tile_of(blockIdx.x, l, tn, tk, bias);
    const int K = w_cols(l);
    const uint16_t* DT = l == kL ? dyT : dzT + (size_t)l * kH * B;          // [N x B]
    const uint16_t* XT = l == 0 ? act_inT : actT + (size_t)(l - 1) * kH * B;  // [K x B]
    const int b_begin = blockIdx.y * split_rows;
    const int b_end = min(B, b_begin + split_rows);
    const int wn = (wave >> 1) * 32, wk = (wave & 1) * 32;
    const uint16_t* ar = DT + (size_t)(tn * kDwTile + wn + (lane & 31)) * B + 8 * (lane >> 5);
    const uint16_t* br = XT + (size_t)(tk * kDwTile + wk + (lane & 31)) * B + 8 * (lane >> 5);
    const bf16x8 ones = __builtin_bit_cast(bf16x8, u32x4{0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u});
    // two accumulators over alternating k steps shorten the dependency chain
    f32x16 acc0 = {}, acc1 = {};
#pragma unroll 2
    for (int b0 = b_begin; b0 < b_end; b0 += kDwStep) {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = ld8(ar + b0 + 16 * q);
            b[q] = bias ? ones : ld8(br + b0 + 16 * q);
        }
        acc0 = mfma32(a[0], b[0], acc0);
        acc1 = mfma32(a[1], b[1], acc1);
        acc0 = mfma32(a[2], b[2], acc0);
        acc1 = mfma32(a[3], b[3], acc1);
    }
    float* P = part + (size_t)blockIdx.y * kNParams;
    if (bias) {  // every column holds the row sums: keep column 0
        if (wk == 0 && (lane & 31) == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) P[b_off(l) + tn * kDwTile + wn + acc32_row(i, lane)] = acc0[i] + acc1[i];
        }
        return;
    }
    const int k = tk * kDwTile + wk + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int n = tn * kDwTile + wn + acc32_row(i, lane);
        P[w_off(l) + (size_t)n * K + k] = acc0[i] + acc1[i];
    }
}

// ---------------------------------------------------------------------------
// reduce the partial gradients (fixed order), then write them or take an
// Adam step; block 0 also sums the loss partials
// ---------------------------------------------------------------------------
// index of weight i (row-major [rows x cols] of its layer) in the transposed copy
__device__ __forceinline__ size_t transposed_slot(size_t i) {
    int l = 0;
#pragma unroll
    for (int q = kL; q > 0; --q)
        if (i >= w_off(q)) {
            l = q;
            break;
        }
    const size_t r = i - w_off(l);
    const size_t n = r / w_cols(l), k = r % w_cols(l);
    return w_off(l) + k * w_rows(l) + n;
}

template <bool ADAM>
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ part, int splits,
                                                         const float* __restrict__ loss_part, int row_blocks,
                                                         float* __restrict__ loss, float* __restrict__ g_out,
                                                         float* __restrict__ w, float* __restrict__ m,
                                                         float* __restrict__ v, uint16_t* __restrict__ w_bf,
                                                         uint16_t* __restrict__ w_bfT, float lr, float b1, float b2,
                                                         float eps, float c1, float c2) {
    if (blockIdx.x == 0 && loss) {
        // strided per-thread sums, then a fixed tree: the same order every call
        __shared__ float red[256];
        float s = 0.0f;
        for (int q = threadIdx.x; q < row_blocks; q += 256) s += loss_part[q];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w2 = 128; w2 > 0; w2 >>= 1) {
            if ((int)threadIdx.x < w2) red[threadIdx.x] += red[threadIdx.x + w2];
            __syncthreads();
        }
        if (threadIdx.x == 0) *loss = red[0];
    }
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= kNParams) return;
    float p[kMaxSplits];
#pragma unroll
    for (int s = 0; s < kMaxSplits; ++s) p[s] = s < splits ? part[(size_t)s * kNParams + i] : 0.0f;
    float g = 0.0f;
#pragma unroll
    for (int s = 0; s < kMaxSplits; ++s) g += p[s];
    if (!ADAM) {
        g_out[i] = g;
        return;
    }
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (i < kNWeights) {
        const uint16_t h = f2bf(wi);
        w_bf[i] = h;
        w_bfT[transposed_slot(i)] = h;
    }
}

// BF16 working copies (and the transpose) of the FP32 master weights.
__global__ void mlp_scatter_bf16_kernel(const float* __restrict__ w, uint16_t* __restrict__ w_bf,
                                        uint16_t* __restrict__ w_bfT) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNWeights) return;
    const uint16_t h = f2bf(w[i]);
    w_bf[i] = h;
    w_bfT[transposed_slot(i)] = h;
}

// Fill with zero (a kernel rather than a memset so that callers may capture
// the training step in a graph).
__global__ void mlp_zero_kernel(float* __restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace ixg

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
extern "C" uint32_t ixg_mlp_splits(uint32_t B, uint32_t* split_rows) {
    // at most kMaxSplits row ranges of >= 256 rows, each a multiple of the
    // 64-row k step
    uint32_t s = B / 256 > 0 ? B / 256 : 1;
    if (s > (uint32_t)ixg::kMaxSplits) s = ixg::kMaxSplits;
    uint32_t rows = (B + s - 1) / s;
    rows = (rows + ixg::kDwStep - 1) / ixg::kDwStep * ixg::kDwStep;
    *split_rows = rows;
    return (B + rows - 1) / rows;
}

extern "C" hipError_t ixg_mlp_forward(const ixg_mlp_dev* d, const float* x, const float* t, uint32_t B, int train,
                                      float* y3, hipStream_t stream) {
    if (train)
        hipLaunchKernelGGL((ixg::mlp_forward_kernel<true, 1>), dim3(B / ixg::kRows), dim3(512), 0, stream, x, t,
                           d->w_bf, d->w, (int)B, d->y, nullptr, d->act_in, d->act, d->dgelu, d->dy, d->dyT,
                           d->loss_part);
    else  // inference: 64 rows per block, every weight operand used four times
        hipLaunchKernelGGL((ixg::mlp_forward_kernel<false, 4>), dim3(B / (4 * ixg::kRows)), dim3(512), 0, stream, x,
                           nullptr, d->w_bf, d->w, (int)B, d->y, y3, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr);
    return hipGetLastError();
}

// backward + dW + reduce: adam = 0 writes the reduced gradient to d->g.
extern "C" hipError_t ixg_mlp_backward(const ixg_mlp_dev* d, uint32_t B, int adam, int want_loss, float lr, float c1,
                                       float c2, hipStream_t stream) {
    const uint32_t row_blocks = B / ixg::kRows;
    hipLaunchKernelGGL(ixg::mlp_backward_kernel, dim3(row_blocks), dim3(512), 0, stream, d->dy, d->w_bfT, d->dgelu,
                       (int)B, d->dz);
    uint32_t split_rows = 0;
    const uint32_t splits = ixg_mlp_splits(B, &split_rows);
    hipLaunchKernelGGL(ixg::mlp_dw_kernel, dim3(ixg::dw_tiles(), splits), dim3(256), 0, stream, d->dyT, d->dz,
                       d->act_in, d->act, (int)B, (int)split_rows, d->part);
    const dim3 grid((unsigned)((ixg::kNParams + 255) / 256));
    float* loss = want_loss ? d->loss : nullptr;
    if (adam)
        hipLaunchKernelGGL(ixg::mlp_reduce_kernel<true>, grid, dim3(256), 0, stream, d->part, (int)splits,
                           d->loss_part, (int)row_blocks, loss, nullptr, d->w, d->m, d->v, d->w_bf, d->w_bfT, lr,
                           0.9f, 0.999f, 1e-8f, c1, c2);
    else
        hipLaunchKernelGGL(ixg::mlp_reduce_kernel<false>, grid, dim3(256), 0, stream, d->part, (int)splits,
                           d->loss_part, (int)row_blocks, loss, d->g, nullptr, nullptr, nullptr, nullptr, nullptr,
                           0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_scatter(const ixg_mlp_dev* d, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_scatter_bf16_kernel, dim3((unsigned)((ixg::kNWeights + 255) / 256)), dim3(256), 0,
                       stream, d->w, d->w_bf, d->w_bfT);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_zero(float* p, int n, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}
