// Batched ensembles of independent LDG advection-diffusion cases (different meshes and orders) on one GPU:
// device tables and launch arguments of the two persistent kernels in ensemble.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dg {

// geometry slots per element, stored [kEnsGeo][Kp] per case: rx ry sx sy, then for face f:
// nx, ny, fscale = Js/J (own side), tau_f = tau max(fscale own, fscale other)
constexpr int kEnsGeo = 16;
// elements per wave: one element per lane
constexpr int kEnsLanes = 64;
// waves per workgroup; they share one LDS copy of the order's operators
constexpr int kEnsWaves = 8;
// operator rows padded to this count (>= P R for every supported split; zero rows)
constexpr int kEnsRows = 48;
// split code of the MFMA variant: a wave takes 16 elements and applies Dr, Ds and LIFT as 16x16x4 FP64 MFMA tiles
constexpr int kEnsMfma = 16;
// nodal arrays per case, each Kp x Np in the group-of-16 layout: u (two buffers), res, qx, qy
enum EnsArray { ENS_U0 = 0, ENS_U1 = 1, ENS_RES = 2, ENS_QX = 3, ENS_QY = 4, kEnsArrays = 5 };

struct EnsCase {
    int N = 0, K = 0, Kp = 0, pad = 0;
    size_t off_u = 0;      // first of the kEnsArrays nodal arrays (doubles)
    size_t off_geo = 0;    // geometry (doubles)
    size_t off_conn = 0;   // connectivity [3][Kp] (ints): en << 3 | fn << 1 | (alternating-flux switch > 0)
    double ax = 0, ay = 0, nu = 0, dt = 0;
};

// one workgroup's work: element blocks [b0, b0 + nb) of case c (64 elements each), each split into P row
// slices (rows [p R, p R + R), R = ceil(Np / P)); the nb P tasks (block, slice) go round-robin to the waves.
// With P = kEnsMfma the blocks are groups of 16 elements, one per wave.
struct EnsItem {
    int c = 0, b0 = 0, nb = 1, P = 1;
};
// largest operator table (order 8): Dr^T, Ds^T and LIFT^T, (2 Np + 3 Nfp) columns of kEnsRows rows
constexpr int kEnsOpMax = (2 * 45 + 3 * 9) * 48;

struct EnsArgs {
    const EnsCase* cases = nullptr;
    const EnsItem* items = nullptr;
    int chunk[9] = {0};          // XCD x walks items [chunk[x], chunk[x+1])
    double* data = nullptr;      // nodal arrays of all cases
    const double* geo = nullptr;
    const int* conn = nullptr;
    const double* ops[9] = {nullptr};   // per order, transposed and row-padded: Dr^T [Np][kEnsRows] | Ds^T |
                                        // LIFT^T [3 Nfp][kEnsRows]
    const double* opsM[9] = {nullptr};  // per order, MFMA A fragments: Dr [MT][KS][64] | Ds | LIFT [MT][KSL][64]
    int cur = 0;                 // the u buffer read by this stage (ENS_U0 or ENS_U1)
    int mode = 1;                // 0: rhs into res; 1: LSERK4 stage res = a res + dt rhs, u_next = u + b res
    double rk_a = 0.0, rk_b = 0.0;
};

// row splits the kernels are built for, per order: R = ceil(Np / P) <= 8 rows per lane, or kEnsMfma (N >= 2)
bool ens_split_supported(int N, int P);
// MFMA fragment tiles of order N: output-row tiles, K steps of Dr/Ds, K steps of LIFT
inline constexpr int ens_mfma_mt(int N) { return ((N + 1) * (N + 2) / 2 + 15) / 16; }
inline constexpr int ens_mfma_ks(int N) { return ((N + 1) * (N + 2) / 2 + 3) / 4; }
inline constexpr int ens_mfma_ksl(int N) { return (3 * (N + 1) + 3) / 4; }
// pass 1: q = grad u - LIFT(...) for the items of order N in a.chunk; pass 2: right-hand side (+ stage update)
hipError_t launch_ens_q(int N, int P, const EnsArgs& a, int grid, hipStream_t s);
hipError_t launch_ens_rhs(int N, int P, const EnsArgs& a, int grid, hipStream_t s);
// resident workgroups per CU of the order-N kernels with split P
int ens_blocks_per_cu(int N, int P);

}  // namespace dg
