// Operators of the incompressible projection solver (HopeFOAM-0.1/tutorials/DG/2D/pearsonVortex/dgChorinFoam),
// on the device, FP64.
//
//  * convection  convecU = div(U U)  -- dg::solveEquation(dgm::Sp(convecU) == dgc::div(U, U)), dgChorinFoam.C:74,
//    with the scheme the tutorial selects ("div(U,U) default LF", system/dgSchemes).  dgc::div of two dof fields
//    (EquationDgcConvectionSchemeType3) forms the fluxes from nodal values: the volume flux is the interpolant of
//    the nodal products U_k U_c (defaultConvectionScheme.C:215-303, gaussCellVandermonde applied to U_k vf), the
//    face flux is the Lax-Friedrichs flux at the face nodes with one dissipation speed per face,
//    maxV = max over the face nodes of |n.U_O|, |n.U_N|, interpolated to the face Gauss points
//    (LFFlux.C:103-211).  After the mass solve of Sp: convecU = M^-1 [ -(grad phi, U U) + <phi, F*> ].
//    Neighbour states on boundary faces are the patch values: fixedValue data (the mesh's ghost states of two
//    variables), else the own trace (zeroGradient and fixedGradient patches evaluate to the internal values,
//    fixedGradientDgPatchField.C evaluate()).
//  * nodal gradient and divergence -- dgc::grad / dgc::div of a dof field: element-local strong derivatives
//    at the nodes, sum_j (Dr_ij rx + Ds_ij sx, Dr_ij ry + Ds_ij sy) . u_j (defaultGrad.C:45-84,
//    defaultDivScheme.C:45-83; their source term -grad u, -div u enters b as M (-grad u), so an Sp equation
//    returns the nodal derivative itself).
//
// One thread per element: the projection solver's meshes are small and these passes are not its cost (the
// SIPG solves are).
#include <hip/hip_runtime.h>

#include "dg_common.hpp"
#include "ins.hpp"

namespace dg {

template <int N>
__global__ __launch_bounds__(128) void ins_convection_kernel(InsArgs a, const double* __restrict__ U,
                                                             double* __restrict__ out) {
    constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1, Nfq = N + 2;
    const int eo = blockIdx.x * 128 + threadIdx.x;   // the output column; padding columns repeat element K - 1
    if (eo >= ((a.K + 15) & ~15)) return;
    const int K = a.K;
    const int e = eo < K ? eo : K - 1;
    double ux[Np], uy[Np], rx_[Np], ry_[Np];
#pragma unroll
    for (int n = 0; n < Np; ++n) {
        ux[n] = U[dev_index(e, 0, n, 2, Np)];
        uy[n] = U[dev_index(e, 1, n, 2, Np)];
        rx_[n] = 0.0;
        ry_[n] = 0.0;
    }
    const double rx = a.geo[G_RX * K + e], ry = a.geo[G_RY * K + e];
    const double sx = a.geo[G_SX * K + e], sy = a.geo[G_SY * K + e];
    // volume: Pr Vq (rx F + ry G) + Ps Vq (sx F + sy G) of the nodal fluxes F_c = ux u_c, G_c = uy u_c
    for (int j = 0; j < Np; ++j) {
        const double fxx = ux[j] * ux[j], fxy = ux[j] * uy[j], fyy = uy[j] * uy[j];
        const double ar_x = rx * fxx + ry * fxy, as_x = sx * fxx + sy * fxy;   // component x: (F, G) = (ux ux, uy ux)
        const double ar_y = rx * fxy + ry * fyy, as_y = sx * fxy + sy * fyy;   // component y: (ux uy, uy uy)
        for (int n = 0; n < Np; ++n) {
            const double sr = a.Sr[n * Np + j], ss = a.Ss[n * Np + j];
            rx_[n] += sr * ar_x + ss * as_x;
            ry_[n] += sr * ar_y + ss * as_y;
        }
    }
    // faces
    for (int f = 0; f < 3; ++f) {
        const int cn = a.conn[f * K + e];
        const double nx = a.geo[(G_NX0 + 3 * f) * K + e], ny = a.geo[(G_NY0 + 3 * f) * K + e];
        const double fs = a.geo[(G_FS0 + 3 * f) * K + e];
        double oxv[Nfp], oyv[Nfp], nxv[Nfp], nyv[Nfp];
#pragma unroll
        for (int k = 0; k < Nfp; ++k) {
            const int node = a.fmask[f * Nfp + k];
            oxv[k] = ux[node];
            oyv[k] = uy[node];
            if (cn >= 0) {   // the neighbour traverses the shared face reversed
                const int en = cn >> 2, fn = cn & 3;
                const int nn = a.fmask[fn * Nfp + (N - k)];
                nxv[k] = U[dev_index(en, 0, nn, 2, Np)];
                nyv[k] = U[dev_index(en, 1, nn, 2, Np)];
            } else if (-cn - 1 == BC_FIXED && a.bcq) {
                nxv[k] = a.bcq[((size_t(0) * K + e) * 3 + f) * Nfp + k];
                nyv[k] = a.bcq[((size_t(1) * K + e) * 3 + f) * Nfp + k];
            } else {
                nxv[k] = oxv[k];
                nyv[k] = oyv[k];
            }
        }
        double vo[Nfp], vn[Nfp], maxv = 0.0;
#pragma unroll
        for (int k = 0; k < Nfp; ++k) {
            vo[k] = nx * oxv[k] + ny * oyv[k];
            vn[k] = nx * nxv[k] + ny * nyv[k];
            maxv = fmax(maxv, fmax(fabs(vo[k]), fabs(vn[k])));
        }
        double fxn[Nfp], fyn[Nfp];
#pragma unroll
        for (int k = 0; k < Nfp; ++k) {
            fxn[k] = (vo[k] * oxv[k] + vn[k] * nxv[k]) * 0.5 + maxv * (oxv[k] - nxv[k]) * 0.5;
            fyn[k] = (vo[k] * oyv[k] + vn[k] * nyv[k]) * 0.5 + maxv * (oyv[k] - nyv[k]) * 0.5;
        }
        for (int q = 0; q < Nfq; ++q) {
            double gx = 0.0, gy = 0.0;
#pragma unroll
            for (int k = 0; k < Nfp; ++k) {
                gx += a.Vf[q * Nfp + k] * fxn[k];
                gy += a.Vf[q * Nfp + k] * fyn[k];
            }
            for (int n = 0; n < Np; ++n) {
                const double l = fs * a.Lift[n * 3 * Nfq + f * Nfq + q];
                rx_[n] -= l * gx;
                ry_[n] -= l * gy;
            }
        }
    }
    // Sp(convecU) == dgc::div(U, U): the weak form above is -div(U U); the equation returns its negative
#pragma unroll
    for (int n = 0; n < Np; ++n) {
        out[dev_index(eo, 0, n, 2, Np)] = -rx_[n];
        out[dev_index(eo, 1, n, 2, Np)] = -ry_[n];
    }
}

// mode 0: out[2] = grad u (u one variable); mode 1: out[1] = div u (u two variables)
template <int N>
__global__ __launch_bounds__(128) void ins_nodal_deriv_kernel(InsArgs a, int mode, const double* __restrict__ u,
                                                              double* __restrict__ out) {
    constexpr int Np = (N + 1) * (N + 2) / 2;
    const int eo = blockIdx.x * 128 + threadIdx.x;   // the output column; padding columns repeat element K - 1
    if (eo >= ((a.K + 15) & ~15)) return;
    const int K = a.K;
    const int e = eo < K ? eo : K - 1;
    const double rx = a.geo[G_RX * K + e], ry = a.geo[G_RY * K + e];
    const double sx = a.geo[G_SX * K + e], sy = a.geo[G_SY * K + e];
    const int nin = mode == 0 ? 1 : 2;
    for (int i = 0; i < Np; ++i) {
        double gx = 0.0, gy = 0.0, dv = 0.0;
        for (int j = 0; j < Np; ++j) {
            const double dx = a.Dr[i * Np + j] * rx + a.Ds[i * Np + j] * sx;
            const double dy = a.Dr[i * Np + j] * ry + a.Ds[i * Np + j] * sy;
            if (mode == 0) {
                const double uj = u[dev_index(e, 0, j, nin, Np)];
                gx += dx * uj;
                gy += dy * uj;
            } else {
                dv += dx * u[dev_index(e, 0, j, nin, Np)] + dy * u[dev_index(e, 1, j, nin, Np)];
            }
        }
        if (mode == 0) {
            out[dev_index(eo, 0, i, 2, Np)] = gx;
            out[dev_index(eo, 1, i, 2, Np)] = gy;
        } else {
            out[dev_index(eo, 0, i, 1, Np)] = dv;
        }
    }
}

// out = sum_t c[t] x[t] over n values (the time-integration combinations of the projection solver)
__global__ __launch_bounds__(256) void lincomb_kernel(LinComb lc, size_t n, double* __restrict__ out) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        double s = 0.0;
        for (int t = 0; t < lc.n; ++t) s += lc.c[t] * lc.x[t][i];
        out[i] = s;
    }
}

// dst variable dv <- src variable sv (same mesh size); the padding columns of the last element group receive
// element K - 1, as on upload
__global__ __launch_bounds__(256) void copy_var_kernel(int K, int Np, const double* __restrict__ src, int snv, int sv,
                                                       double* __restrict__ dst, int dnv, int dv) {
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    const int Kp = (K + 15) & ~15;
    if (i >= size_t(Kp) * Np) return;
    const int e = int(i / Np), n = int(i - size_t(e) * Np);
    const int es = e < K ? e : K - 1;
    dst[dev_index(e, dv, n, dnv, Np)] = src[dev_index(es, sv, n, snv, Np)];
}

template <int N>
static hipError_t conv_n(const InsArgs& a, const double* U, double* out, hipStream_t st) {
    hipLaunchKernelGGL((ins_convection_kernel<N>), dim3((a.K + 15 + 127) / 128), dim3(128), 0, st, a, U, out);
    return hipGetLastError();
}
template <int N>
static hipError_t deriv_n(const InsArgs& a, int mode, const double* u, double* out, hipStream_t st) {
    hipLaunchKernelGGL((ins_nodal_deriv_kernel<N>), dim3((a.K + 15 + 127) / 128), dim3(128), 0, st, a, mode, u, out);
    return hipGetLastError();
}

hipError_t launch_ins_convection(int N, const InsArgs& a, const double* U, double* out, hipStream_t st) {
    switch (N) {
        case 1: return conv_n<1>(a, U, out, st);
        case 2: return conv_n<2>(a, U, out, st);
        case 3: return conv_n<3>(a, U, out, st);
        case 4: return conv_n<4>(a, U, out, st);
        case 5: return conv_n<5>(a, U, out, st);
        case 6: return conv_n<6>(a, U, out, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_nodal_deriv(int N, const InsArgs& a, int mode, const double* u, double* out, hipStream_t st) {
    switch (N) {
        case 1: return deriv_n<1>(a, mode, u, out, st);
        case 2: return deriv_n<2>(a, mode, u, out, st);
        case 3: return deriv_n<3>(a, mode, u, out, st);
        case 4: return deriv_n<4>(a, mode, u, out, st);
        case 5: return deriv_n<5>(a, mode, u, out, st);
        case 6: return deriv_n<6>(a, mode, u, out, st);
        case 7: return deriv_n<7>(a, mode, u, out, st);
        case 8: return deriv_n<8>(a, mode, u, out, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lincomb(const LinComb& lc, size_t n, double* out, int ncu, hipStream_t st) {
    size_t blocks = (n + 255) / 256;
    if (blocks > size_t(8) * ncu) blocks = size_t(8) * ncu;
    hipLaunchKernelGGL(lincomb_kernel, dim3(unsigned(blocks > 0 ? blocks : 1)), dim3(256), 0, st, lc, n, out);
    return hipGetLastError();
}

hipError_t launch_copy_var(int K, int Np, const double* src, int snv, int sv, double* dst, int dnv, int dv,
                           hipStream_t st) {
    const size_t n = size_t(K) * Np;
    hipLaunchKernelGGL(copy_var_kernel, dim3(unsigned((n + 16 * Np + 255) / 256)), dim3(256), 0, st, K, Np, src, snv,
                       sv, dst, dnv, dv);
    return hipGetLastError();
}

}  // namespace dg
