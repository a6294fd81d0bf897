// mlp.hip -- the metric surrogate: a BF16 MLP trained and evaluated with
// hand-written MFMA kernels on CDNA4 (gfx950).
//
// Network: in (padded to 64) -> 4 x [256, GELU] -> out (padded to 64).
// Weights and activations are BF16, every product accumulates in FP32 on the
// matrix cores (v_mfma_f32_32x32x8_bf16), master weights and Adam moments are
// FP32.  The reference has no surrogate; this is new work whose judge is the
// path tracer's own metric evaluator.
//
// One GEMM kernel serves all three products of a layer through operand
// transposes, with the layer epilogues fused into its store:
//   forward   Z = A W^T + b      (store pre-activation f32, GELU(Z) bf16)
//   dA        G = dZ W           (times GELU'(Z_prev), store bf16)
//   dW        dW = dZ^T A        (f32)
// Block tile 64x64 (four waves, 32x32 each), K tile 32 staged in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ixg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU with the exact Gaussian CDF, and its derivative.
__device__ __forceinline__ float gelu(float v) { return 0.5f * v * (1.0f + erff(v * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float v) {
    float cdf = 0.5f * (1.0f + erff(v * 0.70710678118654752f));
    float pdf = 0.39894228040143268f * expf(-0.5f * v * v);
    return cdf + v * pdf;
}

enum Epilogue : int { EPI_GELU = 0, EPI_LINEAR = 1, EPI_DGELU = 2, EPI_F32 = 3 };

constexpr int TM = 64, TN = 64, TK = 32;

// C[M x N] = op(A)[M x K] * op(B)[K x N]
//   op(A)(m,k) = TA ? A[k*lda + m] : A[m*lda + k]
//   op(B)(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]
// M, N multiples of 64 and K a multiple of 32 (callers pad).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const uint16_t* __restrict__ A, int lda,
                                                       const uint16_t* __restrict__ B, int ldb, int M,
                                                       int N, int K, const float* __restrict__ bias,
                                                       const float* __restrict__ aux,
                                                       float* __restrict__ out_f32,
                                                       uint16_t* __restrict__ out_bf16, int ldc) {
    __shared__ uint16_t As[TM][TK + 4];  // [m][k]
    __shared__ uint16_t Bs[TK][TN + 4];  // [k][n]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += TK) {
        // stage A tile: 64 x 32 = 2048 values, 8 per thread
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int e = tid + r * 256;
            int mm = e / TK, kk = e % TK;
            if (TA) {  // contiguous along m
                mm = e % TM;
                kk = e / TM;
            }
            int gm = m0 + mm, gk = k0 + kk;
            As[mm][kk] = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int e = tid + r * 256;
            int kk = e / TN, nn = e % TN;
            if (TB) {  // contiguous along k
                kk = e % TK;
                nn = e / TK;
            }
            int gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK; ks += 8) {
            bf16x4 a, b;
            const int kb = ks + (lane >> 5) * 4;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = (short)As[wm + (lane & 31)][kb + i];
                b[i] = (short)Bs[kb + i][wn + (lane & 31)];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: lane holds column (lane % 32), rows 8*(i/4) + 4*(lane/32) + (i%4)
    const int gn = n0 + wn + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int gm = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const size_t o = (size_t)gm * ldc + gn;
        float v = acc[i];
        if (EPI == EPI_GELU) {
            v += bias[gn];
            out_f32[o] = v;
            out_bf16[o] = f2bf(gelu(v));
        } else if (EPI == EPI_LINEAR) {
            out_f32[o] = v + bias[gn];
        } else if (EPI == EPI_DGELU) {
            out_bf16[o] = f2bf(v * gelu_grad(aux[o]));
        } else {
            out_f32[o] = v;
        }
    }
}

// X f32 [B x n_in] -> bf16 [B x ld], zero-padded columns.
__global__ void mlp_input_kernel(const float* __restrict__ x, int n_in, int B, int ld, uint16_t* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * ld) return;
    int b = i / ld, c = i % ld;
    out[i] = c < n_in ? f2bf(x[(size_t)b * n_in + c]) : (uint16_t)0;
}

// MSE loss on the first n_out columns of y [B x ld] and its gradient (bf16, zero-padded).
__global__ void mlp_loss_kernel(const float* __restrict__ y, const float* __restrict__ t, int n_out, int B, int ld,
                                uint16_t* __restrict__ dy, float* __restrict__ loss_sum) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < B * ld) {
        int b = i / ld, c = i % ld;
        float g = 0.0f;
        if (c < n_out) {
            float d = y[(size_t)b * ld + c] - t[(size_t)b * n_out + c];
            l = d * d;
            g = 2.0f * d / (float)(B * n_out);
        }
        dy[i] = f2bf(g);
    }
    // wave reduction of the squared errors, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0 && loss_sum) atomicAdd(loss_sum, l);
}

// Column sums of a bf16 [B x N] gradient (the bias gradient), one block per column group.
__global__ void mlp_colsum_kernel(const uint16_t* __restrict__ g, int B, int N, float* __restrict__ out) {
    int n = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < B; b += blockDim.x) s += bf2f(g[(size_t)b * N + n]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = red[0];
}

// Adam on the FP32 master parameters; writes the BF16 working copy through
// an index map (master index -> padded bf16 slot).
__global__ void mlp_adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ g, const int32_t* __restrict__ bf_slot,
                                uint16_t* __restrict__ w_bf, int n, float lr, float b1, float b2, float eps,
                                float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    int s = bf_slot[i];
    if (s >= 0) w_bf[s] = f2bf(wi);
}

__global__ void mlp_scatter_bf16_kernel(const float* __restrict__ w, const int32_t* __restrict__ bf_slot,
                                        uint16_t* __restrict__ w_bf, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = bf_slot[i];
    if (s >= 0) w_bf[s] = f2bf(w[i]);
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
#define IXG_GEMM_LAUNCH(TA, TB, EPI)                                                                         \
    hipLaunchKernelGGL((ixg::mlp_gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, stream, A, lda, B, ldb, M, N, \
                       K, bias, aux, out_f32, out_bf16, ldc)

extern "C" hipError_t ixg_mlp_gemm(int mode, const uint16_t* A, int lda, const uint16_t* B, int ldb, int M,
                                   int N, int K, const float* bias, const float* aux, float* out_f32,
                                   uint16_t* out_bf16, int ldc, hipStream_t stream) {
    if (M % ixg::TM || N % ixg::TN || K % ixg::TK) return hipErrorInvalidValue;
    dim3 grid(N / ixg::TN, M / ixg::TM);
    switch (mode) {
    case 0: IXG_GEMM_LAUNCH(false, true, ixg::EPI_GELU); break;     // forward hidden
    case 1: IXG_GEMM_LAUNCH(false, true, ixg::EPI_LINEAR); break;   // forward output
    case 2: IXG_GEMM_LAUNCH(false, false, ixg::EPI_DGELU); break;   // dA * GELU'
    case 3: IXG_GEMM_LAUNCH(true, false, ixg::EPI_F32); break;      // dW
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_input(const float* x, int n_in, int B, int ld, uint16_t* out, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_input_kernel, dim3((B * ld + 255) / 256), dim3(256), 0, stream, x, n_in, B, ld, out);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_loss(const float* y, const float* t, int n_out, int B, int ld, uint16_t* dy,
                                   float* loss, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_loss_kernel, dim3((B * ld + 255) / 256), dim3(256), 0, stream, y, t, n_out, B, ld,
                       dy, loss);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_colsum(const uint16_t* g, int B, int N, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_colsum_kernel, dim3(N), dim3(256), 0, stream, g, B, N, out);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_adam(float* w, float* m, float* v, const float* g, const int32_t* slot,
                                   uint16_t* w_bf, int n, float lr, float b1, float b2, float eps, float c1,
                                   float c2, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, w, m, v, g, slot, w_bf,
                       n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_scatter(const float* w, const int32_t* slot, uint16_t* w_bf, int n,
                                      hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_scatter_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, w, slot, w_bf, n);
    return hipGetLastError();
}

extern "C" hipError_t ixg_mlp_zero(float* p, int n, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::mlp_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}
