// Training and standalone inference kernels of the troubled-cell MLP (in -> 64 -> 64 -> 1).
//
// Training is FP32: one block of 256 threads takes 64 samples through the forward pass and the
// backward pass of the weighted binary cross-entropy with logits
//     L = (1/n) sum_s [ pw y_s softplus(-z_s) + (1 - y_s) softplus(z_s) ],   dL/dz_s = (pw y_s (sig(z_s) - 1)
//                                                                             + (1 - y_s) sig(z_s)) / n,
// staging weights and activations in LDS, and writes its partial gradient; a second kernel sums the
// partials of all blocks in a fixed order (deterministic) and takes one Adam step.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mlp.hpp"
#include "mlp_infer.hpp"

namespace dg {

constexpr int kS = 64;   // samples per block

__device__ __forceinline__ float softplus(float z) { return z > 0.f ? z + log1pf(expf(-z)) : log1pf(expf(z)); }
__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + expf(-z)); }

// PAD: the input width padded to mlp_pad(nin) (sizes the LDS staging only; the loops run to nin)
template <int PAD>
__global__ __launch_bounds__(256) void mlp_grad_kernel(MlpTrainArgs a) {
    constexpr int H = kMlpH;
    __shared__ float sW1[H * PAD], sb1[H], sW2[H * H], sb2[H], sw3[H];
    __shared__ float sx[kS][PAD + 1], h1[kS][H + 1], h2[kS][H + 1], dz[kS];
    __shared__ float sb3, sloss[kS];
    const int nin = a.nin, t = threadIdx.x, s0 = blockIdx.x * kS;
    const float* P = a.p;
    const int oW1 = 0, ob1 = H * nin, oW2 = ob1 + H, ob2 = oW2 + H * H, ow3 = ob2 + H, ob3 = ow3 + H;
    const int NP = ob3 + 1;
    for (int i = t; i < H * nin; i += 256) sW1[i] = P[oW1 + i];
    for (int i = t; i < H * H; i += 256) sW2[i] = P[oW2 + i];
    if (t < H) {
        sb1[t] = P[ob1 + t];
        sb2[t] = P[ob2 + t];
        sw3[t] = P[ow3 + t];
    }
    if (t == 0) sb3 = P[ob3];
    for (int i = t; i < kS * nin; i += 256) {
        const int s = i / nin, k = i - s * nin;
        sx[s][k] = s0 + s < a.n ? a.X[size_t(s0 + s) * nin + k] : 0.f;
    }
    __syncthreads();
    const int j = t & 63, sg = t >> 6;   // unit j, samples sg, sg + 4, ...
    for (int s = sg; s < kS; s += 4) {
        float acc = sb1[j];
        for (int k = 0; k < nin; ++k) acc += sW1[j * nin + k] * sx[s][k];
        h1[s][j] = fmaxf(0.f, acc);
    }
    __syncthreads();
    for (int s = sg; s < kS; s += 4) {
        float acc = sb2[j];
        for (int k = 0; k < H; ++k) acc += sW2[j * H + k] * h1[s][k];
        h2[s][j] = fmaxf(0.f, acc);
    }
    __syncthreads();
    if (t < kS) {
        const bool live = s0 + t < a.n;
        float z = sb3;
        for (int k = 0; k < H; ++k) z += sw3[k] * h2[t][k];
        const float y = live ? a.y[s0 + t] : 0.f;
        const float inv_n = 1.f / float(a.n);
        dz[t] = live ? (a.pos_weight * y * (sigm(z) - 1.f) + (1.f - y) * sigm(z)) * inv_n : 0.f;
        sloss[t] = live ? (a.pos_weight * y * softplus(-z) + (1.f - y) * softplus(z)) * inv_n : 0.f;
    }
    __syncthreads();
    float* G = a.part + size_t(blockIdx.x) * (NP + 1);
    // layer 3
    if (t < H) {
        float acc = 0.f;
        for (int s = 0; s < kS; ++s) acc += dz[s] * h2[s][t];
        G[ow3 + t] = acc;
    } else if (t == H) {
        float acc = 0.f, l = 0.f;
        for (int s = 0; s < kS; ++s) {
            acc += dz[s];
            l += sloss[s];
        }
        G[ob3] = acc;
        G[NP] = l;
    }
    __syncthreads();
    // dh2 = dz w3 (h2 > 0), in place
    for (int s = sg; s < kS; s += 4) h2[s][j] = h2[s][j] > 0.f ? dz[s] * sw3[j] : 0.f;
    __syncthreads();
    // layer 2: db2[j] = sum_s dh2[s][j], dW2[j][k] = sum_s dh2[s][j] h1[s][k]
    if (t < H) {
        float acc = 0.f;
        for (int s = 0; s < kS; ++s) acc += h2[s][t];
        G[ob2 + t] = acc;
    }
    for (int e = t; e < H * H; e += 256) {
        const int jj = e / H, k = e - jj * H;
        float acc = 0.f;
        for (int s = 0; s < kS; ++s) acc += h2[s][jj] * h1[s][k];
        G[oW2 + e] = acc;
    }
    // dh1[s][k] = (h1 > 0) sum_j W2[j][k] dh2[s][j]
    float d1[kS / 4];
    for (int i = 0; i < kS / 4; ++i) {
        const int s = sg + 4 * i;
        float acc = 0.f;
        for (int jj = 0; jj < H; ++jj) acc += sW2[jj * H + j] * h2[s][jj];
        d1[i] = h1[s][j] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    for (int i = 0; i < kS / 4; ++i) h1[sg + 4 * i][j] = d1[i];
    __syncthreads();
    // layer 1
    if (t < H) {
        float acc = 0.f;
        for (int s = 0; s < kS; ++s) acc += h1[s][t];
        G[ob1 + t] = acc;
    }
    for (int e = t; e < H * nin; e += 256) {
        const int jj = e / nin, k = e - jj * nin;
        float acc = 0.f;
        for (int s = 0; s < kS; ++s) acc += h1[s][jj] * sx[s][k];
        G[oW1 + e] = acc;
    }
}

__global__ __launch_bounds__(256) void mlp_reduce_adam_kernel(int np, int nb, const float* __restrict__ part,
                                                              float* __restrict__ g, float* __restrict__ p,
                                                              float* __restrict__ m, float* __restrict__ v, int t,
                                                              float lr, int apply, float* __restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > np) return;
    float acc = 0.f;
    for (int b = 0; b < nb; ++b) acc += part[size_t(b) * (np + 1) + i];
    if (i == np) {
        if (loss) loss[0] = acc;
        return;
    }
    g[i] = acc;
    if (!apply) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.f - b1) * acc;
    const float vi = b2 * v[i] + (1.f - b2) * acc * acc;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, float(t))), vh = vi / (1.f - powf(b2, float(t)));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ __launch_bounds__(256) void mlp_gather_kernel(int nin, int n, const int* __restrict__ idx,
                                                         const float* __restrict__ X, const float* __restrict__ y,
                                                         float* __restrict__ Xb, float* __restrict__ yb) {
    const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
    if (i >= size_t(n) * nin) return;
    const int s = int(i / nin), k = int(i - size_t(s) * nin);
    const int src = idx[s];
    Xb[i] = X[size_t(src) * nin + k];
    if (k == 0) yb[s] = y[src];
}

template <int PAD>
__global__ __launch_bounds__(128) void mlp_forward32_kernel(int nin, int n, const float* __restrict__ X,
                                                            const float* __restrict__ P, float* __restrict__ z) {
    constexpr int H = kMlpH;
    __shared__ float sP[H * PAD + H + H * H + H + H + 1];
    const int np = mlp_params(nin);
    for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = P[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* W1 = sP;
    const float* b1 = W1 + H * nin;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* w3 = b2 + H;
    float x[PAD], h[H];
    for (int k = 0; k < nin; ++k) x[k] = X[size_t(s) * nin + k];
    for (int j = 0; j < H; ++j) {
        float acc = b1[j];
        for (int k = 0; k < nin; ++k) acc += W1[j * nin + k] * x[k];
        h[j] = fmaxf(0.f, acc);
    }
    float out = w3[H];
    for (int j = 0; j < H; ++j) {
        float acc = b2[j];
        for (int k = 0; k < H; ++k) acc += W2[j * H + k] * h[k];
        out += w3[j] * fmaxf(0.f, acc);
    }
    z[s] = out;
}

// one wavefront per 16 samples
template <int PAD>
__global__ __launch_bounds__(64) void mlp_forward_bf16_kernel(MlpInfer w, int n, const float* __restrict__ X,
                                                              float* __restrict__ z) {
    __shared__ float sx[PAD][16];
    const int lane = threadIdx.x, s0 = blockIdx.x * 16;
    for (int i = lane; i < PAD * 16; i += 64) {
        const int k = i >> 4, c = i & 15;
        sx[k][c] = (k < w.nin && s0 + c < n) ? X[size_t(s0 + c) * w.nin + k] : 0.f;
    }
    __syncthreads();
    const float zz = mlp_infer_wave<PAD / 16>(w, sx, lane);
    if (lane < 16 && s0 + lane < n) z[s0 + lane] = zz;
}

hipError_t mlp_grad_partials(const MlpTrainArgs& a, hipStream_t s) {
    if (a.nin > kMlpInMax || a.n < 1) return hipErrorInvalidValue;
    if (mlp_pad(a.nin) == 32)
        hipLaunchKernelGGL(mlp_grad_kernel<32>, dim3((a.n + kS - 1) / kS), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(mlp_grad_kernel<64>, dim3((a.n + kS - 1) / kS), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t mlp_reduce_adam(int np, int nb, const float* part, float* g, float* p, float* m, float* v, int t, float lr,
                           int apply, float* loss, hipStream_t s) {
    hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3((np + 1 + 255) / 256), dim3(256), 0, s, np, nb, part, g, p, m, v,
                       t, lr, apply, loss);
    return hipGetLastError();
}

hipError_t mlp_gather(int nin, int n, const int* idx, const float* X, const float* y, float* Xb, float* yb,
                      hipStream_t s) {
    const size_t tot = size_t(n) * nin;
    hipLaunchKernelGGL(mlp_gather_kernel, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, s, nin, n, idx, X, y, Xb,
                       yb);
    return hipGetLastError();
}

hipError_t mlp_forward32(int nin, int n, const float* X, const float* p, float* z, hipStream_t s) {
    if (nin > kMlpInMax) return hipErrorInvalidValue;
    if (mlp_pad(nin) == 32)
        hipLaunchKernelGGL(mlp_forward32_kernel<32>, dim3((n + 127) / 128), dim3(128), 0, s, nin, n, X, p, z);
    else
        hipLaunchKernelGGL(mlp_forward32_kernel<64>, dim3((n + 127) / 128), dim3(128), 0, s, nin, n, X, p, z);
    return hipGetLastError();
}

hipError_t mlp_forward_bf16(const MlpInfer& w, int n, const float* X, float* z, hipStream_t s) {
    if (w.nin > kMlpInMax) return hipErrorInvalidValue;
    if (mlp_pad(w.nin) == 32)
        hipLaunchKernelGGL(mlp_forward_bf16_kernel<32>, dim3((n + 15) / 16), dim3(64), 0, s, w, n, X, z);
    else
        hipLaunchKernelGGL(mlp_forward_bf16_kernel<64>, dim3((n + 15) / 16), dim3(64), 0, s, w, n, X, z);
    return hipGetLastError();
}

}  // namespace dg
