// TVB minmod limiter kernels (limiter.hip) and their launch arguments.
#pragma once
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace dg {

// Per-element limiter geometry in the group layout [K/16][kLimGeo][16]:
//   2 f + c   neighbour-pair coefficient c (0, 1) of edge f: m_f - b_0 = a1 (b_f - b_0) + a2 (b_j - b_0)
//   6         h (longest edge)
//   7 + 2 f   outward normal (nx, ny) of face f
constexpr int kLimGeo = 13;

struct LimArgs {
    int K = 0;
    int nvar = 0;
    double M = 0.0, nu = 1.5;
    const double* lgeo = nullptr;    // [K/16][kLimGeo][16]
    const int* jsel = nullptr;       // [3][K] second face of the neighbour pair of edge f
    const int* conn = nullptr;       // [3][K] en * 4 + fn, or -1 - tag
    const double* p4 = nullptr;      // 4 x Np: mean and edge-midpoint P1 values
    const double* phi = nullptr;     // Np x 3
    const double* wbc = nullptr;     // Nfp edge-average weights
    const double* bcq = nullptr;     // nvar x K x 3 x Nfp fixedValue data, or null
    double* mom = nullptr;           // [K/16][nvar][4][16] moments of the unlimited state
    int* troubled = nullptr;         // [K/16 * 16]
    unsigned* count = nullptr;       // troubled cells (accumulated)
    int detect_only = 0;             // 1: flag troubled cells, leave q unchanged
    const double* invV = nullptr;    // Np x Np (modal coefficients, MLP features)
    int use_mlp = 0;                 // 1: the MLP decides which cells are troubled (4 variables)
    int all = 0;                     // 1: every cell is troubled (no detection step)
    MlpInfer mlp;
    int positivity = 0;              // 1: Zhang-Shu positivity scaling after limiting (4 variables)
    double gamma = 1.4, eps = 1e-10;
    const double* vpos = nullptr;    // (Np + Nq + 3 Nfq) x Np point evaluation: nodes, cubature, face Gauss
    int npos = 0;
    unsigned* pcount = nullptr;      // cells scaled for positivity (accumulated)
    const double* p4f = nullptr;     // moment operator as f64 MFMA A fragments, [lim_msteps(Np)][64]
    int moments_ready = 0;           // 1: mom already holds the moments of q (written by the stage kernel)
    const double* cff = nullptr;     // feature operator fragments (lim_cff_slot), or null above the MLP's orders
};

// K steps of 4 nodes covering Np (f64 16x16x4 MFMA)
__host__ __device__ constexpr int lim_ksteps(int Np) { return (Np + 3) / 4; }

// The moment product (rows 0..3: mean, edge-midpoint values) takes its K steps in the row order of the FP32 MFMA
// accumulators: at K step s row group kq supplies node lim_mnode(s, kq) = register s % 4 of tile s / 4 of row
// group kq.  A stage kernel holding the new state as FP32-layout accumulator rows therefore forms the same
// moments, in the same order, without moving data (stage_kernels.hip); the moments kernel reads the nodes in this
// order too, so both give the same bits.  lim_msteps(Np) K steps (4 per 16-node tile, zero beyond Np).
__host__ __device__ constexpr int lim_msteps(int Np) { return 4 * ((Np + 15) / 16); }
__host__ __device__ constexpr int lim_mnode(int s, int kq) { return 16 * (s / 4) + 4 * kq + s % 4; }

// Feature operator: the modal coefficients of rho (w = 0) and E (w = 1) of 16 cells as f64 MFMA products
// D_t = sum_w sum_s A[t][w][s] B_w[s], one 16-row tile t per 16-deep K step of the network's first layer.  A
// row i of tile t holds the invV row of feature 16 t + 4 (i % 4) + i / 4 (features 0..Np-1: rho modes,
// Np..2Np-1: E modes), so that accumulator register r of row group rg -- row rg + 4 r -- is feature
// 16 t + 4 rg + r: the bf16 B layout of layer 1.  Only (t, w) pairs whose tile holds modes of w are stored,
// in the order of lim_cff_slot.
__host__ __device__ constexpr bool lim_cff_active(int Np, int t, int w) {
    return w == 0 ? 16 * t < Np : (16 * t < 2 * Np && 16 * t + 16 > Np);
}
__host__ __device__ constexpr int lim_cff_slot(int Np, int t, int w) {
    int n = 0;
    for (int tt = 0; tt < t; ++tt) n += int(lim_cff_active(Np, tt, 0)) + int(lim_cff_active(Np, tt, 1));
    return n + (w == 1 && lim_cff_active(Np, t, 0) ? 1 : 0);
}
__host__ __device__ constexpr int lim_cff_pairs(int Np, int KS) { return lim_cff_slot(Np, KS, 0); }

// Static scratch of dgh_limit_noalloc (limiter.hip): the moments, troubled flags and the two counters of one call
// on a mesh of at most kNoallocMaxK cells and 4 variables, at these byte offsets.
constexpr int kNoallocMaxK = 131072;
constexpr size_t kNoallocMomBytes = size_t(kNoallocMaxK / 16) * 4 * 4 * 16 * sizeof(double);   // 16 MiB
constexpr size_t kNoallocFlagOff = kNoallocMomBytes;
constexpr size_t kNoallocCountOff = kNoallocFlagOff + size_t(kNoallocMaxK) * sizeof(int);
constexpr size_t kNoallocBytes = kNoallocCountOff + 256;
struct LimScratch {
    double* mom;
    int* troubled;
    unsigned* count;   // [2]: troubled cells, cells scaled for positivity
};
// Device address of this device's copy of the scratch array: looked up once per device with that device
// current (dgh_create calls it), then cached.
hipError_t noalloc_scratch(int device, void** ptr);
// zeroes the two counters by a kernel (no memset node when a caller captures the stream in a graph)
hipError_t launch_zero_counts(unsigned* count, hipStream_t s);

// q: field in the device layout (nvar variables).  Limits q in place.
hipError_t launch_tvb(int N, const LimArgs& a, double* q, hipStream_t s);
// Zhang-Shu positivity scaling alone (4 variables), as run after limiting when a.positivity is set
hipError_t launch_positivity(int N, const LimArgs& a, double* q, hipStream_t s);
// MLP input features of every cell, X[K][2 Np + 6] (4 variables; moments and ghosts as the limiter)
hipError_t launch_features(int N, const LimArgs& a, const double* q, float* X, hipStream_t s);

}  // namespace dg
