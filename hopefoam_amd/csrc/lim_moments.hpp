// Moments of the TVB limiter (cell mean and the three edge-midpoint values of the P1 part) of one variable of a
// group of 16 elements, formed on the VALU in one fixed order that the bf16x3 stage kernel's epilogue and the
// moments kernel share, so that the two give the same bits.
//
// Lane l (element l % 16, row group rg = l / 16) holds, at K step s < lim_msteps(Np), the value at node
// lim_mnode(s, rg) (zero beyond Np).  pf is the moment operator in the f64 MFMA A-fragment layout [s][64 lanes]
// (lane i + 16 rg: row i, node lim_mnode(s, rg)), so the coefficient of moment k at the lane's node of step s is
// pf[s * 64 + 16 rg + k].  Each lane sums its nodes for the four moments in ascending s (fused multiply-adds);
// the row groups' partial sums are then combined as (rg 0 + rg 2) + (rg 1 + rg 3), leaving moment rg of element
// l % 16 in lane l -- the layout of the moments array ([group][variable][4][16]).
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

__device__ __forceinline__ void lim_moment_acc(const double* pf, int s, int lane, double x, double (&part)[4]) {
    const double* c = pf + s * 64 + 16 * (lane >> 4);
#pragma unroll
    for (int k = 0; k < 4; ++k) part[k] = fma(c[k], x, part[k]);
}

// Every lane of the wavefront must call it (cross-lane reads).
__device__ __forceinline__ double lim_moment_reduce(const double (&part)[4], int lane) {
    const int rg = lane >> 4;
    const bool lo = rg < 2, even = (rg & 1) == 0;
    // moments 0, 1 gather in row groups 0, 1 and moments 2, 3 in row groups 2, 3 (partner: lane ^ 32)
    const double r0 = __shfl_xor(lo ? part[2] : part[0], 32, 64);
    const double r1 = __shfl_xor(lo ? part[3] : part[1], 32, 64);
    const double q0 = (lo ? part[0] : part[2]) + r0;
    const double q1 = (lo ? part[1] : part[3]) + r1;
    // then moment rg in row group rg (partner: lane ^ 16)
    const double r = __shfl_xor(even ? q1 : q0, 16, 64);
    return (even ? q0 : q1) + r;
}

}  // namespace dg
