// Device evaluation of time-dependent fixedValue ghost states (host launch interface).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dg {

enum { BCFN_NONE = 0, BCFN_ISENTROPIC_VORTEX = 1, BCFN_DOUBLE_MACH = 2 };

struct BcFunction {
    int kind = BCFN_NONE;
    double gamma = 1.4;
    double x0 = 0.0, y0 = 0.0, beta = 0.0, u_inf = 0.0, v_inf = 0.0;
};

// writes the 4 conserved variables of the function at time t into bcq[v * var_stride + slot[i]] for the n
// boundary face nodes at xy[2 i], xy[2 i + 1]
hipError_t launch_bc_eval(const BcFunction& f, double t, int n, const double* xy, const int* slot, size_t var_stride,
                          double* bcq, hipStream_t st);

}  // namespace dg
