// The reference's Triangle limiter (weighted Green-Gauss P1 reconstruction in primitive variables) on the
// GPU: launch arguments of the three kernels in triangle.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

// Per-element geometry in the group layout [K/16][kTriGeo][16]:
//   0..3        xr, xs, yr, ys (x - centroid = xr (r + 1/3) + xs (s + 1/3), y likewise)
//   4 + 8 f     for face f: dxc, dyc (neighbour or ghost centroid minus own centroid), ex, ey (second minus
//               first face endpoint), A (signed diamond area (dxc ey - ex dyc) / 2), A_2 (A0_o + A0_n, or
//               2 A0_o on the boundary; A0 = area / 3), nx, ny (outward normal)
constexpr int kTriGeo = 28;
// per-cell data: 4 means, then the 4 x 3 vertex values (variable-major)
constexpr int kTriCell = 16;

struct TriArgs {
    int K = 0;
    double gamma = 1.4;
    double tol = 1e-2;               // density floor of the reconstruction
    const double* tgeo = nullptr;    // [K/16][kTriGeo][16]
    const double* pmean = nullptr;   // Np mean weights
    const double* rsn = nullptr;     // Np x 2: r + 1/3, s + 1/3 of the nodes
    const int* conn = nullptr;       // [3][K] en * 4 + fn, or -1 - tag
    const double* bcq = nullptr;     // 4 x K x 3 x Nfp fixedValue data, or null
    const int* pfirst = nullptr;     // [K][3]: first face e' 3 + f' of a fixedValue face's patch (its ghost mean is
                                     // that face's first value), -1 or null: the own face
    double* cell = nullptr;          // [K/16][kTriCell][16]
    double* cv = nullptr;            // [K/16][8][16] cell gradients (x, y) of rho, u, v, p
    double* vg = nullptr;            // [K/16][3][8][16] ghost gradients (boundary faces only)
    int* flags = nullptr;            // [K]: in: selection (select = 1); out: bit 0 rebuilt, bit 1 density
                                     // floor, bit 2 pressure below tol
    int select = 0;                  // 1: rebuild only the cells whose flags entry is non-zero on entry
};

// q: 4-variable field (rho, rhou, rhov, E) in the device layout; rebuilt in place
hipError_t launch_triangle(int N, const TriArgs& a, double* q, hipStream_t s);

}  // namespace dg
