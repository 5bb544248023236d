// Learned troubled-cell indicator: a 2-hidden-layer MLP (in -> 64 -> 64 -> 1, ReLU, sigmoid output),
// trained in FP32 on the GPU (mlp.hip) and evaluated in BF16 on MFMA inside the limiter pass.
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

constexpr int kMlpH = 64;              // hidden width (both layers)
constexpr int kMlpInMax = 64;          // inputs of the largest supported order (N = 6: 2 Np + 8 = 64)
// input width padded to whole 32-wide blocks (two 16-deep bf16 MFMA K steps each): 32 up to N = 3, 64 beyond
__host__ __device__ constexpr int mlp_pad(int nin) { return nin <= 32 ? 32 : 64; }

// number of inputs for order N: modal coefficients of rho and E (2 Np), 3 neighbour-mean jumps of each,
// and the TVB scale of each (log10 of the largest edge-midpoint deviation over M h^2)
__host__ __device__ constexpr int mlp_inputs(int Np) { return 2 * Np + 8; }

// flat parameter vector: W1 [H][in], b1 [H], W2 [H][H], b2 [H], w3 [H], b3
__host__ __device__ constexpr int mlp_params(int nin) { return kMlpH * nin + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1; }

// BF16 copies of the layer-1/2 weights in the per-lane order of the 16x16x16 bf16 MFMA A operand
// (lane l: row 16 m + l % 16, k = 16 s + 4 (l / 16) + j, j = 0..3), plus FP32 biases and w3.
struct MlpInfer {
    const unsigned short* W1f = nullptr;   // [4 m][mlp_pad(nin) / 16 s][64 lanes][4]
    const unsigned short* W2f = nullptr;   // [4 m][4 s][64 lanes][4]
    const float* b1 = nullptr;
    const float* b2 = nullptr;
    const float* w3 = nullptr;
    float b3 = 0.0f;
    float zthr = 0.0f;                     // decision: logit >= zthr (zthr = logit(threshold))
    int nin = 0;
};

struct MlpTrainArgs {
    int nin = 0;
    int n = 0;                     // samples in this batch
    const float* X = nullptr;      // [n][nin] batch inputs (gathered)
    const float* y = nullptr;      // [n] labels (0 / 1)
    const float* p = nullptr;      // parameters
    float pos_weight = 1.0f;       // weight of the positive (troubled) class in the loss
    float* part = nullptr;         // [blocks][params + 1] partial gradients and loss
};

// batch of n samples: partial gradients per 64-sample block
hipError_t mlp_grad_partials(const MlpTrainArgs& a, hipStream_t s);
// g = sum of partials (fixed order); Adam step t (1-based) on p with moments m, v; loss_out[0] = batch loss
hipError_t mlp_reduce_adam(int nparams, int nblocks, const float* part, float* g, float* p, float* m, float* v, int t,
                           float lr, int apply, float* loss_out, hipStream_t s);
// gathers rows idx[i] of X / y into Xb / yb
hipError_t mlp_gather(int nin, int n, const int* idx, const float* X, const float* y, float* Xb, float* yb,
                      hipStream_t s);
// FP32 forward: logits z[n]
hipError_t mlp_forward32(int nin, int n, const float* X, const float* p, float* z, hipStream_t s);
// BF16 MFMA forward: logits z[n]
hipError_t mlp_forward_bf16(const MlpInfer& w, int n, const float* X, float* z, hipStream_t s);

}  // namespace dg
