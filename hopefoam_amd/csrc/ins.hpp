// Operators of the incompressible projection solver (host launch interface of ins.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dg {

struct InsArgs {
    int K = 0;
    const double* geo = nullptr;    // [kGeoSlots][K] FP64 geometric factors
    const int* conn = nullptr;      // [3][K]
    const int* fmask = nullptr;     // [3][Nfp]
    const double* bcq = nullptr;    // fixedValue ghost states of two variables [2][K][3][Nfp] or null
    const double* Sr = nullptr;     // Np x Np: Pr Vq, Ps Vq (the weak derivative of nodal fluxes)
    const double* Ss = nullptr;
    const double* Lift = nullptr;   // Np x 3 Nfq
    const double* Vf = nullptr;     // Nfq x Nfp
    const double* Dr = nullptr;     // Np x Np nodal derivatives
    const double* Ds = nullptr;
};

constexpr int kLinCombMax = 8;
struct LinComb {
    int n = 0;
    double c[kLinCombMax];
    const double* x[kLinCombMax];
};

hipError_t launch_ins_convection(int N, const InsArgs& a, const double* U, double* out, hipStream_t st);
// mode 0: out (2 variables) = grad u (1 variable); mode 1: out (1 variable) = div u (2 variables)
hipError_t launch_nodal_deriv(int N, const InsArgs& a, int mode, const double* u, double* out, hipStream_t st);
hipError_t launch_lincomb(const LinComb& lc, size_t n, double* out, int ncu, hipStream_t st);
hipError_t launch_copy_var(int K, int Np, const double* src, int snv, int sv, double* dst, int dnv, int dv,
                           hipStream_t st);

}  // namespace dg
