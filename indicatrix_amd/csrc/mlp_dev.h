// mlp_dev.h -- the device buffers of one surrogate network and the
// launchers of its kernels (mlp.hip), shared with the host side (mlp_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct ixg_mlp_dev {
    float* w;          // FP32 master parameters [ixg_mlp_param_count()]
    float* m;          // Adam first moments
    float* v;          // Adam second moments
    float* g;          // reduced gradient (the gradients entry point)
    uint16_t* w_bf;    // BF16 weights, [out x in] per layer
    uint16_t* w_bfT;   // BF16 weights, each layer transposed [in x out]
    uint16_t* act_in;  // [64 x B] BF16 inputs, batch-transposed
    uint16_t* act;     // [4][256 x B] BF16 activations, batch-transposed
    float* dgelu;      // [4][256 x B] GELU'(Z), batch-transposed
    uint16_t* dz;      // [4][256 x B] BF16 back-propagated gradients, batch-transposed
    float* y;          // [B x 64] outputs
    uint16_t* dy;      // [B x 64] BF16 output gradient
    uint16_t* dyT;     // [64 x B] the same, batch-transposed
    float* part;       // [16][params] dW / db split-K partials
    float* loss_part;  // [B/16] squared-error partials per row block
    float* loss;       // [1] summed squared error
};

// dW split-K plan for a batch of B rows: the number of row ranges (<= 16)
// and their length (a multiple of 64).
extern "C" uint32_t ixg_mlp_splits(uint32_t B, uint32_t* split_rows);
// The forward pass of B rows (B a multiple of 64); train = 1 also stores
// what the backward pass reads and forms the loss and dY against t;
// train = 0 writes the three outputs per row to y3 [B x 3].
extern "C" hipError_t ixg_mlp_forward(const ixg_mlp_dev* d, const float* x, const float* t, uint32_t B, int train,
                                      float* y3, hipStream_t stream);
// backward + dW + the reduction: adam = 1 takes an Adam step (lr, bias
// corrections c1, c2), adam = 0 writes the reduced gradient to d->g;
// want_loss also sums the squared error into d->loss.
extern "C" hipError_t ixg_mlp_backward(const ixg_mlp_dev* d, uint32_t B, int adam, int want_loss, float lr, float c1,
                                       float c2, hipStream_t stream);
// The BF16 working copies (and their transpose) of the FP32 master weights.
extern "C" hipError_t ixg_mlp_scatter(const ixg_mlp_dev* d, hipStream_t stream);
extern "C" hipError_t ixg_mlp_zero(float* p, int n, hipStream_t stream);
