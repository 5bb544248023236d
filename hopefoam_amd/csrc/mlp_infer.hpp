// BF16 MFMA inference of the troubled-cell MLP for 16 cells by one wavefront (device code, included by
// mlp.hip and limiter.hip).
//
// v_mfma_f32_16x16x16_bf16 operands: A[i][k] in lane i + 16 (k / 4), element k % 4; B[k][j] in lane
// j + 16 (k / 4), element k % 4; D[i][j] in lane j + 16 (i / 4), register i % 4.  The D layout of a
// 16-row tile is exactly the B layout of one 16-deep K step, so the hidden activations of layer 1 feed
// layer 2 straight from the accumulator registers.
#pragma once
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace dg {

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline unsigned short f2bf(float x) {
    unsigned u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

// float -> bf16 bits on the device: the hardware conversion (v_cvt_pk_bf16_f32), round to nearest even as f2bf
__device__ inline unsigned short f2bf_dev(float x) {
    const __bf16 b = static_cast<__bf16>(x);
    unsigned short u;
    __builtin_memcpy(&u, &b, 2);
    return u;
}

// Inputs in registers: xb[s] element j = input 16 s + 4 (lane / 16) + j of cell lane % 16 (the bf16 B layout of
// a 16-deep K step).  Weights in the per-lane fragment order of MlpInfer (global memory or LDS); b1, b2 and w3
// 16-byte aligned (each lane reads its four consecutive rows as one vector).  Returns the logit of cell
// lane % 16 (the same value in all four 16-lane groups).
template <int KS>
__device__ inline float mlp_infer_regs(const mlp_s4* W1, const mlp_s4* W2, const float* b1, const float* b2,
                                       const float* w3, float b3, const mlp_s4 (&xb)[KS], int lane) {
    const int kg = lane >> 4;
    // K steps outer, row tiles inner: consecutive MFMAs feed different accumulators
    mlp_f4 acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = mlp_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m)
            acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W1[(m * KS + s) * 64 + lane], xb[s], acc[m], 0, 0, 0);
    mlp_s4 hb[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const float4 bb = *reinterpret_cast<const float4*>(b1 + 16 * m + 4 * kg);
        hb[m][0] = static_cast<short>(f2bf_dev(fmaxf(0.f, acc[m][0] + bb.x)));
        hb[m][1] = static_cast<short>(f2bf_dev(fmaxf(0.f, acc[m][1] + bb.y)));
        hb[m][2] = static_cast<short>(f2bf_dev(fmaxf(0.f, acc[m][2] + bb.z)));
        hb[m][3] = static_cast<short>(f2bf_dev(fmaxf(0.f, acc[m][3] + bb.w)));
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = mlp_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m)
            acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W2[(m * 4 + s) * 64 + lane], hb[s], acc[m], 0, 0, 0);
    float z = 0.f;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const float4 bb = *reinterpret_cast<const float4*>(b2 + 16 * m + 4 * kg);
        const float4 ww = *reinterpret_cast<const float4*>(w3 + 16 * m + 4 * kg);
        z += ww.x * fmaxf(0.f, acc[m][0] + bb.x);
        z += ww.y * fmaxf(0.f, acc[m][1] + bb.y);
        z += ww.z * fmaxf(0.f, acc[m][2] + bb.z);
        z += ww.w * fmaxf(0.f, acc[m][3] + bb.w);
    }
    z += __shfl_xor(z, 16, 64);
    z += __shfl_xor(z, 32, 64);
    return z + b3;
}

// sx: [16 KS][16] inputs (feature k of cell j), rows >= nin zero; KS = mlp_pad(nin) / 16 K steps.
template <int KS>
__device__ inline float mlp_infer_wave(const MlpInfer& w, const float (*sx)[16], int lane) {
    const int col = lane & 15, kg = lane >> 4;
    mlp_s4 xb[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) xb[s][j] = static_cast<short>(f2bf_dev(sx[16 * s + 4 * kg + j][col]));
    return mlp_infer_regs<KS>(reinterpret_cast<const mlp_s4*>(w.W1f), reinterpret_cast<const mlp_s4*>(w.W2f), w.b1,
                              w.b2, w.w3, w.b3, xb, lane);
}

}  // namespace dg
