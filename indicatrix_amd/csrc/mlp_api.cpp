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
#include "mlp_dev.h"

extern "C" hipStream_t ixg_internal_stream(ixg_ctx* ctx);
extern "C" int ixg_internal_device(ixg_ctx* ctx);

namespace {
constexpr int IN = 16, IN_P = 64, H = 256, L = 4, OUT = 3, OUT_P = 64;
constexpr size_t W0 = (size_t)H * IN_P, WH = (size_t)H * H, W4 = (size_t)OUT_P * H;
constexpr size_t N_WEIGHTS = W0 + 3 * WH + W4;
constexpr size_t N_BIAS = (size_t)H * L + OUT_P;
constexpr size_t N_PARAMS = N_WEIGHTS + N_BIAS;
constexpr uint32_t MAX_SPLITS = 16;

size_t w_off(int l) { return l == 0 ? 0 : W0 + (size_t)(l - 1) * WH; }
}  // namespace

struct ixg_mlp {
    ixg_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t max_batch = 0;
    uint64_t step = 0;
    ixg_mlp_dev d{};
};

namespace {
ixg_status fin(hipError_t e) { return e == hipSuccess ? IXG_OK : IXG_DEVICE_LOST; }

#define MLP_TRY(expr)                                 \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return IXG_DEVICE_LOST; \
    } while (0)

ixg_status check_batch(ixg_mlp* n, uint32_t B) {
    if (!n || B == 0 || B > n->max_batch || B % 64) return IXG_INVALID;
    return IXG_OK;
}

// forward + backward + reduction (+ Adam when adam), loss to *loss if asked
ixg_status step(ixg_mlp* n, const float* d_x, const float* d_t, uint32_t B, bool adam, float lr, float* loss) {
    float c1 = 1.0f, c2 = 1.0f;
    if (adam) {
        n->step += 1;
        c1 = 1.0f - std::pow(0.9f, (float)n->step);
        c2 = 1.0f - std::pow(0.999f, (float)n->step);
    }
    MLP_TRY(ixg_mlp_forward(&n->d, d_x, d_t, B, 1, nullptr, n->stream));
    MLP_TRY(ixg_mlp_backward(&n->d, B, adam ? 1 : 0, loss ? 1 : 0, lr, c1, c2, n->stream));
    if (loss) {
        float l = 0.0f;
        MLP_TRY(hipMemcpyAsync(&l, n->d.loss, 4, hipMemcpyDeviceToHost, n->stream));
        MLP_TRY(hipStreamSynchronize(n->stream));
        *loss = l / (float)(B * OUT);
    }
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
    const size_t B = max_batch, blocks = max_batch / 16;
    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    ixg_mlp_dev& d = n->d;
    alloc((void**)&d.w, N_PARAMS * 4);
    alloc((void**)&d.m, N_PARAMS * 4);
    alloc((void**)&d.v, N_PARAMS * 4);
    alloc((void**)&d.g, N_PARAMS * 4);
    alloc((void**)&d.w_bf, N_WEIGHTS * 2);
    alloc((void**)&d.w_bfT, N_WEIGHTS * 2);
    alloc((void**)&d.act_in, B * IN_P * 2);
    alloc((void**)&d.act, L * B * H * 2);
    alloc((void**)&d.dgelu, L * B * H * 4);
    alloc((void**)&d.dz, L * B * H * 2);
    alloc((void**)&d.y, B * OUT_P * 4);
    alloc((void**)&d.dy, B * OUT_P * 2);
    alloc((void**)&d.dyT, B * OUT_P * 2);
    alloc((void**)&d.part, MAX_SPLITS * N_PARAMS * 4);
    alloc((void**)&d.loss_part, blocks * 4);
    alloc((void**)&d.loss, 4);
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
    ixg_mlp_dev& d = n->d;
    void* ptrs[] = {d.w, d.m, d.v, d.g, d.w_bf, d.w_bfT, d.act_in, d.act, d.dgelu, d.dz, d.y, d.dy, d.dyT,
                    d.part, d.loss_part, d.loss};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete n;
}

ixg_status ixg_mlp_set_params(ixg_mlp* n, const float* host, size_t count) {
    if (!n || !host || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(hipMemcpyAsync(n->d.w, host, N_PARAMS * 4, hipMemcpyHostToDevice, n->stream));
    MLP_TRY(ixg_mlp_scatter(&n->d, n->stream));
    n->step = 0;
    MLP_TRY(ixg_mlp_zero(n->d.m, (int)N_PARAMS, n->stream));
    MLP_TRY(ixg_mlp_zero(n->d.v, (int)N_PARAMS, n->stream));
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_get_params(ixg_mlp* n, float* host, size_t count) {
    if (!n || !host || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(hipMemcpyAsync(host, n->d.w, N_PARAMS * 4, hipMemcpyDeviceToHost, n->stream));
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_get_state(ixg_mlp* n, float* m, float* v, size_t count, uint64_t* step) {
    if (!n || !m || !v || !step || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(hipMemcpyAsync(m, n->d.m, N_PARAMS * 4, hipMemcpyDeviceToHost, n->stream));
    MLP_TRY(hipMemcpyAsync(v, n->d.v, N_PARAMS * 4, hipMemcpyDeviceToHost, n->stream));
    *step = n->step;
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_set_state(ixg_mlp* n, const float* m, const float* v, size_t count, uint64_t step) {
    if (!n || !m || !v || count != N_PARAMS) return IXG_INVALID;
    MLP_TRY(hipMemcpyAsync(n->d.m, m, N_PARAMS * 4, hipMemcpyHostToDevice, n->stream));
    MLP_TRY(hipMemcpyAsync(n->d.v, v, N_PARAMS * 4, hipMemcpyHostToDevice, n->stream));
    n->step = step;
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_infer(ixg_mlp* n, const float* d_x, uint32_t B, float* d_y) {
    ixg_status st = check_batch(n, B);
    if (st != IXG_OK || !d_x || !d_y) return IXG_INVALID;
    MLP_TRY(ixg_mlp_forward(&n->d, d_x, nullptr, B, 0, d_y, n->stream));
    return fin(hipStreamSynchronize(n->stream));
}

ixg_status ixg_mlp_gradients(ixg_mlp* n, const float* d_x, const float* d_t, uint32_t B, float* host_grad,
                             size_t count, float* loss) {
    ixg_status st = check_batch(n, B);
    if (st != IXG_OK || !d_x || !d_t || !host_grad || count != N_PARAMS) return IXG_INVALID;
    float l = 0.0f;
    st = step(n, d_x, d_t, B, false, 0.0f, &l);
    if (st != IXG_OK) return st;
    MLP_TRY(hipMemcpyAsync(host_grad, n->d.g, N_PARAMS * 4, hipMemcpyDeviceToHost, n->stream));
    MLP_TRY(hipStreamSynchronize(n->stream));
    if (loss) *loss = l;
    return IXG_OK;
}

ixg_status ixg_mlp_train_step(ixg_mlp* n, const float* d_x, const float* d_t, uint32_t B, float lr,
                              float* loss) {
    ixg_status st = check_batch(n, B);
    if (st != IXG_OK || !d_x || !d_t) return IXG_INVALID;
    return step(n, d_x, d_t, B, true, lr, loss);
}

ixg_status ixg_mlp_sync(ixg_mlp* n) { return n ? fin(hipStreamSynchronize(n->stream)) : IXG_INVALID; }

}  // extern "C"
