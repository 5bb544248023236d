// Launch interface of the DG stage kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

struct StageArgs {
    int K;
    int flux;            // FLUX_* (Euler: LF or ROE)
    int mode;            // 0: q_out = rhs;  1: LSERK4 stage update
    int block_cap;       // > 0: at most this many persistent blocks (rounded down to a multiple of 8, at least 8)
    const double* q_in;  // [nvar][K][Np]
    double* q_out;       // [nvar][K][Np]
    double* res;         // [nvar][K][Np] (mode 1)
    const void* geo;     // [16][K] in the compute precision
    const int* conn;     // [3][K]: neighbour*4 + face, or -1 - bc tag
    const void* VqF;     // operator fragments in the compute precision (bf16x3: split fragments, 32-bit words)
    const void* PF;
    const void* LF;
    const void* Vf;      // [Nfq][Nfp]
    const int* fmask;    // [3][Nfp]
    const double* bcq;   // fixedValue nodal ghost values [nvar][K][3][Nfp] or null
    double ax, ay, gamma;
    double rk_a, rk_b, dt;
    // bf16x3 stage update only (else null): the limiter's moments of the new state (mean and the three
    // edge-midpoint values per variable, [K/16][nvar][4][16]) are written to mom_out, formed by the f64 moment
    // fragments p4f ([4 ceil(Np/16)][64], limiter.hpp) exactly as the limiter's moments kernel forms them
    double* mom_out = nullptr;
    const double* p4f = nullptr;
};

hipError_t launch_stage(int N, int phys, int prec, const StageArgs& a, int ncu, hipStream_t st);
hipError_t launch_field_diff(const double* a, const double* b, double* out, size_t n, int ncu, hipStream_t st);
hipError_t launch_mfma_layout_probe(double* out, hipStream_t st);
hipError_t launch_mfma_probe(const double* A, const double* B, double* D64, const float* Af, const float* Bf,
                             float* D32, hipStream_t st);
// A: 64 lanes x 4 words of bf16x3 fragments (hi pair, hi pair, lo pair, lo pair; lane i + 16 kq holds
// A[i][4 kq + e] in element e); B, D3, D1: 16 x 16 row-major
hipError_t launch_mfma_bf16x3_probe(const unsigned* A, const float* B, float* D3, float* D1, hipStream_t st);

}  // namespace dg
