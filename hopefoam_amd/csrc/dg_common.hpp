// Shared definitions between the host code and the HIP kernels.
#pragma once
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

namespace dg {

constexpr int kMaxN = 8;

// cubature points of the order-3(N+1) rule, N = 0..8 (index 0 unused)
constexpr int kNq[kMaxN + 1] = {0, 12, 19, 36, 54, 73, 93, 118, 145};

constexpr int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Per-order sizes of the MFMA tiling used by the stage kernels.
template <int N>
struct Dims {
    static constexpr int Np = (N + 1) * (N + 2) / 2;
    static constexpr int Nfp = N + 1;
    static constexpr int Nfq = N + 2;
    static constexpr int Nq = kNq[N];
    static constexpr int MT = round_up(Np, 16) / 16;     // output-node tiles
    static constexpr int QT = round_up(Nq, 16) / 16;     // cubature-point tiles
    static constexpr int KS1 = round_up(Np, 4) / 4;      // K steps of the interpolation GEMM
    static constexpr int SF = round_up(Nfq, 4) / 4;      // K steps per face (Gauss points padded to 4)
    static constexpr int KSF = 3 * SF;                   // K steps of the face lift GEMM
    static constexpr int FT = round_up(KSF, 4) / 4;      // interpolation tiles holding the own face traces
    // fragment array sizes (64 lanes each); VqF holds QT cubature tiles then FT face-trace tiles
    static constexpr int nVqF = (QT + FT) * KS1 * 64;
    static constexpr int nPF = QT * 4 * MT * 2 * 64;
    static constexpr int nLF = KSF * MT * 64;
    // bf16x3 split products (v_mfma_f32_16x16x16_bf16): one K16 step holds four of the K4 steps above, in the
    // four bf16 elements of a lane's operand; each lane's fragment is 4 words (hi pair, hi pair, lo pair, lo pair)
    static constexpr int KS16 = (KS1 + 3) / 4;           // K16 steps of the interpolation GEMM
    static constexpr int nVqS = (QT + FT) * KS16 * 64 * 4;   // 32-bit words
    static constexpr int nPS = QT * MT * 2 * 64 * 4;
    static constexpr int nLS = FT * MT * 64 * 4;
};

// Row i of the 16x16 MFMA accumulator held by register r (0..3) of a lane in row group rg = lane / 16:
// f32 (v_mfma_f32_16x16x4f32) packs four consecutive rows per lane; f64 (v_mfma_f64_16x16x4f64)
// interleaves them.  Checked on the device by dgh_selftest_mfma_layout.
template <typename T>
constexpr int mfma_row(int rg, int r) {
    return sizeof(T) == 8 ? rg + 4 * r : 4 * rg + r;
}
// Node supplied by row group k' at K step s of the interpolation GEMM.  With the f64 row map this is also
// the output row 16 m + mfma_row(k', r) for s = 4 m + r, so the FP64 stage update reuses the B fragments
// as q_in.
template <typename T>
constexpr int gemm1_node(int s, int kq) {
    return 4 * s + kq;
}
// inverse: the (row group, register) holding accumulator row i
template <typename T>
constexpr int mfma_rg(int i) {
    return sizeof(T) == 8 ? i % 4 : i / 4;
}
template <typename T>
constexpr int mfma_reg(int i) {
    return sizeof(T) == 8 ? i / 4 : i % 4;
}

// Node k of face f of the reference triangle of order N, in the order of RefElem::fmask (face 0 along s = -1,
// face 1 the hypotenuse, face 2 along r = -1), as a compile-time function so that the stage kernels look nothing
// up.  The host checks it against the reference element's table when it builds the element (dgh_api.cpp).
__host__ __device__ constexpr int face_node_n(int N, int f, int k) {
    const int Nfp = N + 1, Np = (N + 1) * (N + 2) / 2;
    if (f == 0) return k;
    int n = f == 1 ? N : Np - 1;
    for (int i = 1; i <= k; ++i) n += f == 1 ? Nfp - i : -(i + 1);
    return n;
}
template <int N>
__host__ __device__ constexpr int face_node(int f, int k) {
    return face_node_n(N, f, k);
}

// Device layout of nodal fields: elements in groups of 16 (one MFMA column each), node-major inside
// a group: value (e, v, n) of an nvar-variable field sits at ((e/16 * nvar + v) * Np + n) * 16 + e%16.
// K is padded to a multiple of 16 (padding elements repeat element K-1).
__host__ __device__ inline size_t dev_index(int e, int v, int n, int nvar, int Np) {
    return ((size_t(e >> 4) * nvar + v) * Np + n) * 16 + (e & 15);
}

// Carpenter-Kennedy five-stage fourth-order low-storage Runge-Kutta coefficients (the values of
// createFields.H:119-138): res = A[s] res + dt rhs, q += B[s] res
inline constexpr double kLserkA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                      -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
inline constexpr double kLserkB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                      1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                      2277821191437.0 / 14882151754819.0};
// stage times t^n + C[s] dt of the same scheme
inline constexpr double kLserkC[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                                      2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

// geometric-factor slots, stored [kGeoSlots][K]
enum GeoSlot { G_RX = 0, G_RY, G_SX, G_SY, G_J, G_NX0, G_NY0, G_FS0, G_NX1, G_NY1, G_FS1, G_NX2, G_NY2, G_FS2,
               kGeoSlots = 16 };

// boundary tags for faces without a neighbour (stored in conn as -1 - tag)
enum BcTag { BC_NONE = 0, BC_FIXED = 1, BC_ZEROGRAD = 2, BC_WALL = 3 };

enum Phys { PHYS_ADVECT = 0, PHYS_EULER = 1 };
enum Flux { FLUX_UPWIND = 0, FLUX_LF = 1, FLUX_ROE = 2, FLUX_AVERAGE = 3 };
enum Prec { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2 };

}  // namespace dg
