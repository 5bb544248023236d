// fixedValue ghost states as functions of time, evaluated on the device.
//
// The reference's Euler tutorials refresh the fixedValue patch data at the start of every time step with t^n
// (setBoundaryValues(rho1, rhoU1, Ener1, gamma, runTime - runTime.deltaT()), dgEulerFoam.C:73), node by node at
// the owner's face nodes (ownerDofMapping).  One thread per boundary face node writes the four conserved
// variables of that node into the mesh's ghost array [4][K][3][Nfp] (the layout the stage kernel and the
// limiters read).
#include <hip/hip_runtime.h>

#include "boundary.hpp"

namespace dg {

// HopeFOAM-0.1/tutorials/DG/2D/isentropicVortex/dgEulerFoam/setBoundaryValues.H:20-46: r = squared distance to
// the advected centre; rho = (1 - (gamma - 1) beta^2 e^(2 (1 - r)) / (16 gamma pi^2))^(1 / (gamma - 1)),
// U = U_inf + beta e^(1 - r) (-(y - yc), x - xc) / (2 pi), E = rho^gamma / (gamma - 1) + |rho U|^2 / (2 rho)
__device__ inline void vortex_state(const BcFunction& f, double t, double x, double y, double* q) {
    const double pi = 3.14159265358979323846;
    const double g = f.gamma;
    const double xc = f.x0 + f.u_inf * t, yc = f.y0 + f.v_inf * t;
    const double dx = x - xc, dy = y - yc;
    const double r = dx * dx + dy * dy;
    const double rho = pow(1.0 - (g - 1.0) * (f.beta * f.beta) * exp(2.0 * (1.0 - r)) / (16.0 * g * pi * pi),
                           1.0 / (g - 1.0));
    const double e1 = f.beta * exp(1.0 - r) / (2.0 * pi);
    const double mx = (f.u_inf - e1 * dy) * rho, my = (f.v_inf + e1 * dx) * rho;
    q[0] = rho;
    q[1] = mx;
    q[2] = my;
    q[3] = pow(rho, g) / (g - 1.0) + 0.5 * (mx * mx + my * my) / rho;
}

// HopeFOAM-0.1/tutorials/DG/2D/doubleMach/dgEulerFoam/setBoundaryValues.H:18-60: the Mach 10 shock through
// (1/6, 0) at 60 degrees to the wall, moving with it; position = (x - 1/6) / gTime - y, gTime =
// (1 + 20 t) / sqrt(3); behind the shock (position < 0) rho = 8, rho U = 8 * 8.25 (cos 30deg, -sin 30deg),
// p = 116.5; ahead of it rho = 1.4, U = 0, p = 1
__device__ inline void double_mach_state(const BcFunction& f, double t, double x, double y, double* q) {
    const double pi = 3.14159265358979323846;
    const double g = f.gamma;
    const double gt = (1.0 + 20.0 * t) / sqrt(3.0);
    const double position = 1.0 / gt * (x - 1.0 / 6.0) - y;
    if (position < 0.0) {
        const double mx = 8.25 * cos(pi / 6.0) * 8.0, my = -8.25 * sin(pi / 6.0) * 8.0;
        q[0] = 8.0;
        q[1] = mx;
        q[2] = my;
        q[3] = 116.5 / (g - 1.0) + (mx * mx + my * my) / (2.0 * 8.0);
    } else {
        q[0] = 1.4;
        q[1] = 0.0;
        q[2] = 0.0;
        q[3] = 1.0 / (g - 1.0);
    }
}

__global__ __launch_bounds__(256) void bc_eval_kernel(BcFunction f, double t, int n, const double* __restrict__ xy,
                                                      const int* __restrict__ slot, size_t var_stride,
                                                      double* __restrict__ bcq) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double x = xy[2 * i], y = xy[2 * i + 1];
    double q[4];
    if (f.kind == BCFN_DOUBLE_MACH)
        double_mach_state(f, t, x, y, q);
    else
        vortex_state(f, t, x, y, q);
    const size_t s = size_t(slot[i]);
#pragma unroll
    for (int v = 0; v < 4; ++v) bcq[v * var_stride + s] = q[v];
}

hipError_t launch_bc_eval(const BcFunction& f, double t, int n, const double* xy, const int* slot, size_t var_stride,
                          double* bcq, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bc_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, st, f, t, n, xy, slot, var_stride, bcq);
    return hipGetLastError();
}

}  // namespace dg
