// TVB minmod limiter for triangles (Cockburn & Shu 1998, the triangle limiter), component-wise on the
// conserved variables, for gfx950.  Not part of the reference (whose Trianglelimite.C:61-852 is a
// different scheme); it sits where the reference's limiteSchemes::limite hook is called
// (limiteScheme.H:139-159) and replaces that whole-field call.
//
// Two kernels, one block of 256 threads per group of 16 elements (thread t: element t % 16, slot t / 16):
//   tvb_moments_kernel  cell mean and the P1 part at the three edge midpoints of every variable
//                       (4 x Np operator on the group's nodal values, staged in LDS);
//   tvb_limit_kernel    per (variable, edge): d = mbar(ut(m) - ubar, nu D(m)) from the neighbour (or ghost)
//                       means; a cell is troubled when any d differs from its argument; troubled cells are
//                       rewritten in place as ubar + sum dh_i phi_i with the positivity rescaling of dh.
// The second kernel reads only moments of the unlimited state, so limiting in place is race-free.
#include <hip/hip_runtime.h>

#include "dg_common.hpp"
#include "limiter.hpp"

namespace dg {

template <int N>
struct LShape {
    static constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
};

template <int N, int NV>
__global__ __launch_bounds__(256) void tvb_moments_kernel(LimArgs a, const double* __restrict__ q) {
    constexpr int Np = LShape<N>::Np;
    __shared__ double sq[NV * Np][16];
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const size_t base = size_t(g) * NV * Np * 16;
    for (int i = threadIdx.x; i < NV * Np * 16; i += 256) sq[i >> 4][i & 15] = q[base + i];
    __syncthreads();
    if (slot < 4 * NV) {
        const int v = slot >> 2, k = slot & 3;
        double acc = 0.0;
        for (int n = 0; n < Np; ++n) acc += a.p4[k * Np + n] * sq[v * Np + n][el];
        a.mom[((size_t(g) * NV + v) * 4 + k) * 16 + el] = acc;
    }
}

__device__ __forceinline__ double minmod2(double x, double y) {
    if (x > 0.0 && y > 0.0) return fmin(x, y);
    if (x < 0.0 && y < 0.0) return fmax(x, y);
    return 0.0;
}

// mean of variable v behind face f of element e (neighbour, or the ghost cell of a boundary face)
template <int N, int NV>
__device__ __forceinline__ double outer_mean(const LimArgs& a, int g, int el, int e, int v, int f) {
    constexpr int Nfp = LShape<N>::Nfp;
    const int cn = a.conn[f * a.K + e];
    if (cn >= 0) {
        const int en = cn >> 2;
        return a.mom[((size_t(en >> 4) * NV + v) * 4) * 16 + (en & 15)];
    }
    const int tag = -cn - 1;
    if (tag == BC_FIXED && a.bcq) {
        const double* b = a.bcq + ((size_t(v) * a.K + e) * 3 + f) * Nfp;
        double s = 0.0;
        for (int k = 0; k < Nfp; ++k) s += a.wbc[k] * b[k];
        return s;
    }
    const double own = a.mom[((size_t(g) * NV + v) * 4) * 16 + el];
    if (tag == BC_WALL && NV == 4 && (v == 1 || v == 2)) {
        // mirror the mean momentum through the wall (reflectiveDgPatchField.C:111-154)
        const double mu = a.mom[((size_t(g) * NV + 1) * 4) * 16 + el];
        const double mv = a.mom[((size_t(g) * NV + 2) * 4) * 16 + el];
        const double nx = a.lgeo[(size_t(g) * kLimGeo + 7 + 2 * f) * 16 + el];
        const double ny = a.lgeo[(size_t(g) * kLimGeo + 8 + 2 * f) * 16 + el];
        const double un = mu * nx + mv * ny;
        return v == 1 ? mu - 2.0 * un * nx : mv - 2.0 * un * ny;
    }
    return own;
}

template <int N, int NV>
__global__ __launch_bounds__(256) void tvb_limit_kernel(LimArgs a, double* __restrict__ q) {
    constexpr int Np = LShape<N>::Np;
    __shared__ double sd[NV][3][16];
    __shared__ int schg[3 * NV][16];
    __shared__ unsigned nbad;
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int e0 = g * 16 + el;
    const bool live = e0 < a.K;
    const int e = live ? e0 : a.K - 1;
    if (threadIdx.x == 0) nbad = 0u;
    if (slot < 3 * NV) {
        const int v = slot / 3, f = slot - 3 * (slot / 3);
        const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + el];
        const double dev = a.mom[((size_t(g) * NV + v) * 4 + 1 + f) * 16 + el] - u0;
        const int j = a.jsel[f * a.K + e];
        const double a1 = a.lgeo[(size_t(g) * kLimGeo + 2 * f) * 16 + el];
        const double a2 = a.lgeo[(size_t(g) * kLimGeo + 2 * f + 1) * 16 + el];
        const double h = a.lgeo[(size_t(g) * kLimGeo + 6) * 16 + el];
        const double D = a1 * (outer_mean<N, NV>(a, g, el, e, v, f) - u0) +
                         a2 * (outer_mean<N, NV>(a, g, el, e, v, j) - u0);
        const double r = fabs(dev) <= a.M * h * h ? dev : minmod2(dev, a.nu * D);
        sd[v][f][el] = r;
        schg[slot][el] = r != dev;
    }
    __syncthreads();
    int bad = 0;
    for (int i = 0; i < 3 * NV; ++i) bad |= schg[i][el];
    if (bad && slot < NV) {
        const int v = slot;
        const double d0 = sd[v][0][el], d1 = sd[v][1][el], d2 = sd[v][2][el];
        if (d0 + d1 + d2 != 0.0) {
            const double pos = fmax(0.0, d0) + fmax(0.0, d1) + fmax(0.0, d2);
            const double neg = fmax(0.0, -d0) + fmax(0.0, -d1) + fmax(0.0, -d2);
            const double tp = pos > 0.0 ? fmin(1.0, neg / pos) : 0.0;
            const double tn = neg > 0.0 ? fmin(1.0, pos / neg) : 0.0;
            sd[v][0][el] = tp * fmax(0.0, d0) - tn * fmax(0.0, -d0);
            sd[v][1][el] = tp * fmax(0.0, d1) - tn * fmax(0.0, -d1);
            sd[v][2][el] = tp * fmax(0.0, d2) - tn * fmax(0.0, -d2);
        }
    }
    __syncthreads();
    if (bad && live) {
        for (int idx = slot; idx < NV * Np; idx += 16) {
            const int v = idx / Np, n = idx - v * Np;
            const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + el];
            q[dev_index(e0, v, n, NV, Np)] =
                u0 + sd[v][0][el] * a.phi[n * 3] + sd[v][1][el] * a.phi[n * 3 + 1] + sd[v][2][el] * a.phi[n * 3 + 2];
        }
    }
    if (slot == 0 && live) {
        a.troubled[e0] = bad;
        if (bad) atomicAdd(&nbad, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && nbad) atomicAdd(a.count, nbad);
}

template <int N, int NV>
static hipError_t launch_nv(const LimArgs& a, double* q, hipStream_t s) {
    const int groups = (a.K + 15) / 16;
    hipLaunchKernelGGL((tvb_moments_kernel<N, NV>), dim3(groups), dim3(256), 0, s, a, q);
    hipLaunchKernelGGL((tvb_limit_kernel<N, NV>), dim3(groups), dim3(256), 0, s, a, q);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_n(const LimArgs& a, double* q, hipStream_t s) {
    if (a.nvar == 1) return launch_nv<N, 1>(a, q, s);
    if (a.nvar == 4) return launch_nv<N, 4>(a, q, s);
    return hipErrorInvalidValue;
}

hipError_t launch_tvb(int N, const LimArgs& a, double* q, hipStream_t s) {
    switch (N) {
        case 1: return launch_n<1>(a, q, s);
        case 2: return launch_n<2>(a, q, s);
        case 3: return launch_n<3>(a, q, s);
        case 4: return launch_n<4>(a, q, s);
        case 5: return launch_n<5>(a, q, s);
        case 6: return launch_n<6>(a, q, s);
        case 7: return launch_n<7>(a, q, s);
        case 8: return launch_n<8>(a, q, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dg
