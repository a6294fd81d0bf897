// mlp_api.cpp -- host side of the BF16 metric surrogate (see mlp.hip).
//
// Parameter storage (FP32 master, padded so every GEMM is tile-aligned):
//   weights  W0 [256 x 64] | W1, W2, W3 [256 x 256] | W4 [64 x 256]
//   biases   b0..b3 [256]  | b4 [64]
// The padding (input columns 16..63, output rows 3..63) starts at zero and
// receives zero gradient, so it stays zero; the logical network is
// 16 -> 4 x 256 (GELU) -> 3 with 202,499 parameters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/ixg.h"

extern "C" hipError_t ixg_mlp_gemm(int mode, const uint16_t* A, int lda, const uint16_t* B, int ldb, int M,
                                   int N, int K, const float* bias, const float* aux, float* out_f32,
                                   uint16_t* out_bf16, int ldc, hipStream_t stream);
extern "C" hipError_t ixg_mlp_input(const float* x, int n_in, int B, int ld, uint16_t* out, hipStream_t stream);
extern "C" hipError_t ixg_mlp_loss(const float* y, const float* t, int n_out, int B, int ld, uint16_t* dy,
                                   float* loss, hipStream_t stream);
extern "C" hipError_t ixg_mlp_colsum(const uint16_t* g, int B, int N, float* out, hipStream_t stream);
extern "C" hipError_t ixg_mlp_adam(float* w, float* m, float* v, const float* g, const int32_t* slot,
                                   uint16_t* w_bf, int n, float lr, float b1, float b2, float eps, float c1,
                                   float c2, hipStream_t stream);
extern "C" hipError_t ixg_mlp_scatter(const float* w, const int32_t* slot, uint16_t* w_bf, int n,
                                      hipStream_t stream);
extern "C" hipError_t ixg_mlp_zero(float* p, int n, hipStream_t stream);
extern "C" hipStream_t ixg_internal_stream(ixg_ctx* ctx);
extern "C" int ixg_internal_device(ixg_ctx* ctx);

namespace {
constexpr int IN = 16, IN_P = 64, H = 256, L = 4, OUT = 3, OUT_P = 64;
constexpr size_t W0 = (size_t)H * IN_P, WH = (size_t)H * H, W4 = (size_t)OUT_P * H;
constexpr size_t N_WEIGHTS = W0 + 3 * WH + W4;
constexpr size_t N_BIAS = (size_t)H * L + OUT_P;
constexpr size_t N_PARAMS = N_WEIGHTS + N_BIAS;

size_t w_off(int l) { return l == 0 ? 0 : W0 + (size_t)(l - 1) * WH; }
size_t b_off(int l) { return N_WEIGHTS + (size_t)l * H; }
}  // namespace

struct ixg_mlp {
    ixg_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t max_batch = 0;
    uint64_t step = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 master / Adam / grads
    uint16_t* w_bf = nullptr;                                      // BF16 working weights
    int32_t* slot = nullptr;
    uint16_t* act_in = nullptr;
    float* pre[L] = {};
    uint16_t* act[L] = {};
    uint16_t* dz[L] = {};
    float* y = nullptr;
    uint16_t* dy = nullptr;
    float* loss = nullptr;
};

namespace {
ixg_status fin(hipError_t e) { return e == hipSuccess ? IXG_OK : IXG_DEVICE_LOST; }

#define MLP_TRY(expr)                              \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return IXG_DEVICE_LOST; \
    } while (0)

hipError_t forward(ixg_mlp* n, const float* d_x, uint32_t B) {
    hipError_t e = ixg_mlp_input(d_x, IN, (int)B, IN_P, n->act_in, n->stream);
    const uint16_t* a = n->act_in;
    int k = IN_P;
    for (int l = 0; l < L && e == hipSuccess; ++l) {
        e = ixg_mlp_gemm(0, a, k, n->w_bf + w_off(l), k, (int)B, H, k, n->w + b_off(l), nullptr, n->pre[l],
                         n->act[l], H, n->stream);
        a = n->act[l];
        k = H;
    }
    if (e == hipSuccess)
        e = ixg_mlp_gemm(1, a, H, n->w_bf + w_off(4), H, (int)B, OUT_P, H, n->w + b_off(4), nullptr, n->y,
                         nullptr, OUT_P, n->stream);
    return e;
}

hipError_t backward(ixg_mlp* n, uint32_t B) {
    const int Bi = (int)B;
    // output layer: dW4 = dY^T A3, db4 = colsum(dY), dZ3 = (dY W4) * GELU'(Z3)
    hipError_t e = ixg_mlp_gemm(3, n->dy, OUT_P, n->act[L - 1], H, OUT_P, H, Bi, nullptr, nullptr,
                                n->g + w_off(4), nullptr, H, n->stream);
    if (e == hipSuccess) e = ixg_mlp_colsum(n->dy, Bi, OUT_P, n->g + b_off(4), n->stream);
    if (e == hipSuccess)
        e = ixg_mlp_gemm(2, n->dy, OUT_P, n->w_bf + w_off(4), H, Bi, H, OUT_P, nullptr, n->pre[L - 1], nullptr,
                         n->dz[L - 1], H, n->stream);
    for (int l = L - 1; l >= 0 && e == hipSuccess; --l) {
        const uint16_t* a_prev = l == 0 ? n->act_in : n->act[l - 1];
        const int k = l == 0 ? IN_P : H;
        e = ixg_mlp_gemm(3, n->dz[l], H, a_prev, k, H, k, Bi, nullptr, nullptr, n->g + w_off(l), nullptr, k,
                         n->stream);
        if (e == hipSuccess) e = ixg_mlp_colsum(n->dz[l], Bi, H, n->g + b_off(l), n->stream);
        if (e == hipSuccess && l > 0)
            e = ixg_mlp_gemm(2, n->dz[l], H, n->w_bf + w_off(l), H, Bi, H, H, nullptr, n->pre[l - 1], nullptr,
                             n->dz[l - 1], H, n->stream);
    }
    return e;
}

ixg_status check_batch(ixg_mlp* n, uint32_t B) {
    if (!n || B == 0 || B > n->max_batch || B % 64) return IXG_INVALID;
    return IXG_OK;
}
}  // namespace

extern "C" {

uint32_t ixg_mlp_param_count(void) { return (uint32_t)N_PARAMS; }

ixg_status ixg_mlp_create(ixg_ctx* ctx, uint32_t max_batch, uint64_t seed, ixg_mlp** out) {
    if (!ctx || !out || max_batch == 0 || max_batch % 64) return IXG_INVALID;
    *out = nullptr;
    if (hipSetDevice(ixg_internal_device(ctx)) != hipSuccess) return IXG_DEVICE_LOST;
    ixg_mlp* n = new ixg_mlp();
    n->ctx = ctx;
    n->stream = ixg_internal_stream(ctx);
    n->max_batch = max_batch;
    const size_t B = max_batch;
    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    alloc((void**)&n->w, N_PARAMS * 4);
    alloc((void**)&n->m, N_PARAMS * 4);
    alloc((void**)&n->v, N_PARAMS * 4);
    alloc((void**)&n->g, N_PARAMS * 4);
    alloc((void**)&n->w_bf, N_WEIGHTS * 2);
    alloc((void**)&n->slot, N_PARAMS * 4);
    alloc((void**)&n->act_in, B * IN_P * 2);
    for (int l = 0; l < L; ++l) {
        alloc((void**)&n->pre[l], B * H * 4);
        alloc((void**)&n->act[l], B * H * 2);
        alloc((void**)&n->dz[l], B * H * 2);
    }
    alloc((void**)&n->y, B * OUT_P * 4);
    alloc((void**)&n->dy, B * OUT_P * 2);
    alloc((void**)&n->loss, 4);
    if (e != hipSuccess) {
        ixg_mlp_destroy(n);
        return IXG_DEVICE_LOST;
    }
    // He-uniform weights on the logical fan-in, zero biases, zero padding
    std::vector<float> host(N_PARAMS, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l <= L; ++l) {
        const int rows = l == L ? OUT : H, cols = l == 0 ? IN : H, ld = l == 0 ? IN_P : H;
        const float lim = std::sqrt(6.0f / (float)cols) * (l == L ? 0.5f : 1.0f);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) host[w_off(l) + (size_t)r * ld + c] = U(rng);
    }
    std::vector<int32_t> slot(N_PARAMS);
    for (size_t i = 0; i < N_PARAMS; ++i) slot[i] = i < N_WEIGHTS ? (int32_t)i : -1;
    e = hipMemcpy(n->slot, slot.data(), N_PARAMS * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ixg_mlp_destroy(n);
        return IXG_DEVICE_LOST;
    }
    ixg_status st = ixg_mlp_set_params(n, host.data(), N_PARAMS);
    if (st != IXG_OK) {
        ixg_mlp_destroy(n);
        return st;
    }
    *out = n;
    return IXG_OK;
}

void ixg_mlp_destroy(ixg_mlp* n) {
    if (!n) return;
    (void)hipStreamSynchronize(n->stream);
    void* ptrs[] = {n->w, n->m, n->v, n->g, n->w_bf, n->slot, n->act_in, n->y, n->dy, n->loss};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int l = 0; l < L; ++l) {
        if (n->pre[l]) (void)hipFree(n->pre[l]);
        if (n->act[l]) (void)hipFree(n->act[l]);
        if (n->dz[l]) (void)hipFree(n->dz[l]);
    }
    delete n;
}

ixg_status ixg_mlp_set_params(ixg_mlp* n, const float* host, size_t count) {
    if (!n || !host || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(hipMemcpyAsync(n->w, host, N_PARAMS * 4, hipMemcpyHostToDevice, n->stream));
    MLP_TRY(ixg_mlp_scatter(n->w, n->slot, n->w_bf, (int)N_PARAMS, n->stream));
    n->step = 0;
    MLP_TRY(ixg_mlp_zero(n->m, (int)N_PARAMS, n->stream));
    MLP_TRY(ixg_mlp_zero(n->v, (int)N_PARAMS, n->stream));
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_get_params(ixg_mlp* n, float* host, size_t count) {
    if (!n || !host || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(hipMemcpyAsync(host, n->w, N_PARAMS * 4, hipMemcpyDeviceToHost, n->stream));
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_infer(ixg_mlp* n, const float* d_x, uint32_t B, float* d_y) {
    ixg_status st = check_batch(n, B);
    if (st != IXG_OK || !d_x || !d_y) return IXG_INVALID;
    MLP_TRY(forward(n, d_x, B));
    MLP_TRY(hipMemcpy2DAsync(d_y, OUT * 4, n->y, OUT_P * 4, OUT * 4, B, hipMemcpyDeviceToDevice, n->stream));
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_gradients(ixg_mlp* n, const float* d_x, const float* d_t, uint32_t B, float* host_grad,
                             size_t count, float* loss) {
    ixg_status st = check_batch(n, B);
    if (st != IXG_OK || !d_x || !d_t || !host_grad || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(ixg_mlp_zero(n->loss, 1, n->stream));
    MLP_TRY(forward(n, d_x, B));
    MLP_TRY(ixg_mlp_loss(n->y, d_t, OUT, (int)B, OUT_P, n->dy, n->loss, n->stream));
    MLP_TRY(backward(n, B));
    MLP_TRY(hipMemcpyAsync(host_grad, n->g, N_PARAMS * 4, hipMemcpyDeviceToHost, n->stream));
    float l = 0.0f;
    MLP_TRY(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
    MLP_TRY(hipStreamSynchronize(n->stream));
    if (loss) *loss = l / (float)(B * OUT);
    return IXG_OK;
}

ixg_status ixg_mlp_train_step(ixg_mlp* n, const float* d_x, const float* d_t, uint32_t B, float lr,
                              float* loss) {
    ixg_status st = check_batch(n, B);
    if (st != IXG_OK || !d_x || !d_t) return IXG_INVALID;
    if (loss) MLP_TRY(ixg_mlp_zero(n->loss, 1, n->stream));
    MLP_TRY(forward(n, d_x, B));
    MLP_TRY(ixg_mlp_loss(n->y, d_t, OUT, (int)B, OUT_P, n->dy, loss ? n->loss : nullptr, n->stream));
    MLP_TRY(backward(n, B));
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
    MLP_TRY(ixg_mlp_adam(n->w, n->m, n->v, n->g, n->slot, n->w_bf, (int)N_PARAMS, lr, b1, b2, 1e-8f, c1, c2,
                         n->stream));
    if (loss) {
        float l = 0.0f;
        MLP_TRY(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
        MLP_TRY(hipStreamSynchronize(n->stream));
        *loss = l / (float)(B * OUT);
    }
    return IXG_OK;
}

ixg_status ixg_mlp_sync(ixg_mlp* n) { return n ? fin(hipStreamSynchronize(n->stream)) : IXG_INVALID; }

}  // extern "C"
