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
//
// Positivity (4 variables): a third kernel scales every cell towards its mean (Zhang & Shu 2010; on
// triangles Zhang, Xia & Shu 2012) so that density and pressure stay >= eps_c = min(eps, rhobar, p(Ubar))
// at the nodes, cubature points and face Gauss points -- every point where the right-hand side evaluates
// the state:  rho <- rhobar + t1 (rho - rhobar), then U <- Ubar + t2 (U - Ubar), t2 the smallest
// admissible fraction over the points (60 bisection steps on p(Ubar + t (U_x - Ubar)) >= eps_c).
//
// Learned detection (use_mlp): the troubled-cell test is replaced by the MLP on the cell's features
//   c_k / (|c_0| + 1e-12), k < Np, for rho and E (c = V^-1 u, the modal coefficients),
//   (ubar_f - ubar_0) / (|ubar_0| + 1e-12) for the three neighbour (or ghost) means of rho and E, and
//   log10(max_i |ut(m_i) - ubar_0| / (M h^2) + 1e-30) clipped to [-10, 10] for rho and E (the TVB scale),
// evaluated for the 16 cells of a group by one wavefront on BF16 MFMA (mlp_infer.hpp); the limiting
// itself is unchanged.
#include <hip/hip_runtime.h>

#include "dg_common.hpp"
#include "limiter.hpp"
#include "mlp_infer.hpp"

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

// MLP features of the 16 cells of group g into sx[mlp_pad(2 Np + 8)][16] (rows >= 2 Np + 8 zero); 4 variables.
template <int N>
__device__ void group_features(const LimArgs& a, const double* __restrict__ q, int g, float (*sx)[16],
                               double (*sq)[16], double (*sc)[16]) {
    constexpr int Np = LShape<N>::Np, NV = 4;
    const int el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int e0 = g * 16 + el;
    const int e = e0 < a.K ? e0 : a.K - 1;
    for (int i = threadIdx.x; i < 2 * Np * 16; i += 256) {
        const int w = i / (Np * 16), r = i - w * Np * 16;   // w = 0: rho (variable 0), 1: E (variable 3)
        sq[w * Np + (r >> 4)][r & 15] = q[(size_t(g) * NV + 3 * w) * Np * 16 + r];
    }
    for (int i = threadIdx.x; i < mlp_pad(mlp_inputs(Np)) * 16; i += 256) sx[i >> 4][i & 15] = 0.f;
    __syncthreads();
    for (int it = slot; it < 2 * Np; it += 16) {
        const int w = it / Np, k = it - w * Np;
        double c = 0.0;
        for (int j = 0; j < Np; ++j) c += a.invV[k * Np + j] * sq[w * Np + j][el];
        sc[it][el] = c;
    }
    if (slot < 6) {
        const int w = slot / 3, f = slot - 3 * w, v = 3 * w;
        const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + el];
        const double nb = outer_mean<N, NV>(a, g, el, e, v, f);
        sx[2 * Np + slot][el] = float((nb - u0) / (fabs(u0) + 1e-12));
    } else if (slot < 8) {
        const int v = 3 * (slot - 6);
        const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + el];
        double dmax = 0.0;
        for (int i = 0; i < 3; ++i) dmax = fmax(dmax, fabs(a.mom[((size_t(g) * NV + v) * 4 + 1 + i) * 16 + el] - u0));
        const double h = a.lgeo[(size_t(g) * kLimGeo + 6) * 16 + el];
        const double t = log10(dmax / (a.M * h * h) + 1e-30);
        sx[2 * Np + slot][el] = float(fmin(10.0, fmax(-10.0, t)));
    }
    __syncthreads();
    for (int it = slot; it < 2 * Np; it += 16) {
        const int w = it / Np;
        sx[it][el] = float(sc[it][el] / (fabs(sc[w * Np][el]) + 1e-12));
    }
    __syncthreads();
}

template <int N>
__global__ __launch_bounds__(256) void tvb_features_kernel(LimArgs a, const double* __restrict__ q,
                                                           float* __restrict__ X) {
    constexpr int Np = LShape<N>::Np, nin = mlp_inputs(Np);
    __shared__ float sx[mlp_pad(nin)][16];
    __shared__ double sq[2 * Np][16], sc[2 * Np][16];
    const int g = blockIdx.x;
    group_features<N>(a, q, g, sx, sq, sc);
    for (int i = threadIdx.x; i < nin * 16; i += 256) {
        const int c = i / nin, k = i - c * nin;
        if (g * 16 + c < a.K) X[size_t(g * 16 + c) * nin + k] = sx[k][c];
    }
}

template <int N, int NV, bool MLP>
__global__ __launch_bounds__(256) void tvb_limit_kernel(LimArgs a, double* __restrict__ q) {
    constexpr int Np = LShape<N>::Np;
    __shared__ double sd[NV][3][16];
    __shared__ int schg[3 * NV][16];
    __shared__ unsigned nbad;
    __shared__ int smlp[16];
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int e0 = g * 16 + el;
    const bool live = e0 < a.K;
    const int e = live ? e0 : a.K - 1;
    if constexpr (MLP) {
        constexpr int PAD = mlp_pad(mlp_inputs(Np));
        __shared__ float sx[PAD][16];
        __shared__ double sq[2 * Np][16], sc[2 * Np][16];
        group_features<N>(a, q, g, sx, sq, sc);
        if (threadIdx.x < 64) {
            const float z = mlp_infer_wave<PAD / 16>(a.mlp, sx, threadIdx.x);
            if (threadIdx.x < 16) smlp[threadIdx.x] = z >= a.mlp.zthr;
        }
    }
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
    if constexpr (MLP) {
        bad = smlp[el];
    } else {
        for (int i = 0; i < 3 * NV; ++i) bad |= schg[i][el];
        bad |= a.all;
    }
    if (bad && slot < NV && !a.detect_only) {
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
    if (bad && live && !a.detect_only) {
        for (int idx = slot; idx < NV * Np; idx += 16) {
            const int v = idx / Np, n = idx - v * Np;
            const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + el];
            q[dev_index(e0, v, n, NV, Np)] =
                u0 + sd[v][0][el] * a.phi[n * 3] + sd[v][1][el] * a.phi[n * 3 + 1] + sd[v][2][el] * a.phi[n * 3 + 2];
        }
    }
    if (slot == 0 && live) {
        a.troubled[e0] = bad;
        if (bad && !a.all) atomicAdd(&nbad, 1u);
    }
    __syncthreads();
    // every cell limited: the host knows the count; otherwise one atomic per block that has troubled cells
    if (threadIdx.x == 0 && nbad) atomicAdd(a.count, nbad);
}

__device__ __forceinline__ double pressure(double r, double mx, double my, double E, double gamma) {
    return (gamma - 1.0) * (E - 0.5 * (mx * mx + my * my) / r);
}

template <int N>
__global__ __launch_bounds__(256) void positivity_kernel(LimArgs a, double* __restrict__ q) {
    constexpr int Np = LShape<N>::Np, NV = 4;
    __shared__ double sq[NV * Np][16];
    __shared__ double sub[NV][16];
    __shared__ double sred[16][16];
    __shared__ unsigned nsc;
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const bool live = g * 16 + el < a.K;
    const size_t base = size_t(g) * NV * Np * 16;
    for (int i = threadIdx.x; i < NV * Np * 16; i += 256) sq[i >> 4][i & 15] = q[base + i];
    if (threadIdx.x == 0) nsc = 0u;
    __syncthreads();
    if (slot < NV) {
        double acc = 0.0;
        for (int j = 0; j < Np; ++j) acc += a.p4[j] * sq[slot * Np + j][el];
        sub[slot][el] = acc;
    }
    __syncthreads();
    const double rb = sub[0][el], mxb = sub[1][el], myb = sub[2][el], Eb = sub[3][el];
    const double pb = pressure(rb, mxb, myb, Eb, a.gamma);
    const bool adm = rb > 0.0 && pb > 0.0;
    const double ec = fmin(a.eps, fmin(rb, pb));
    // density
    double rmin = 1e300;
    if (adm)
        for (int x = slot; x < a.npos; x += 16) {
            double r = 0.0;
            for (int j = 0; j < Np; ++j) r += a.vpos[x * Np + j] * sq[j][el];
            rmin = fmin(rmin, r);
        }
    sred[slot][el] = rmin;
    __syncthreads();
    for (int k = 0; k < 16; ++k) rmin = fmin(rmin, sred[k][el]);
    const bool s1 = adm && rmin < ec;
    const double t1 = s1 ? (rb - ec) / (rb - rmin) : 1.0;
    __syncthreads();
    if (s1)
        for (int j = slot; j < Np; j += 16) sq[j][el] = rb + t1 * (sq[j][el] - rb);
    __syncthreads();
    // pressure
    double tmin = 1.0;
    if (adm)
        for (int x = slot; x < a.npos; x += 16) {
            double ux[4];
            for (int v = 0; v < 4; ++v) {
                double acc = 0.0;
                for (int j = 0; j < Np; ++j) acc += a.vpos[x * Np + j] * sq[v * Np + j][el];
                ux[v] = acc;
            }
            if (pressure(ux[0], ux[1], ux[2], ux[3], a.gamma) >= ec) continue;
            double lo = 0.0, hi = 1.0;
            for (int it = 0; it < 60; ++it) {
                const double t = 0.5 * (lo + hi);
                const double r = rb + t * (ux[0] - rb), mx = mxb + t * (ux[1] - mxb);
                const double my = myb + t * (ux[2] - myb), E = Eb + t * (ux[3] - Eb);
                if (pressure(r, mx, my, E, a.gamma) >= ec) lo = t;
                else hi = t;
            }
            tmin = fmin(tmin, lo);
        }
    sred[slot][el] = tmin;
    __syncthreads();
    for (int k = 0; k < 16; ++k) tmin = fmin(tmin, sred[k][el]);
    const bool s2 = adm && tmin < 1.0;
    if (s2)
        for (int idx = slot; idx < NV * Np; idx += 16) {
            const int v = idx / Np;
            sq[idx][el] = sub[v][el] + tmin * (sq[idx][el] - sub[v][el]);
        }
    __syncthreads();
    if ((s1 || s2) && live) {
        for (int idx = slot; idx < NV * Np; idx += 16) q[base + size_t(idx) * 16 + el] = sq[idx][el];
        if (slot == 0) atomicAdd(&nsc, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && nsc) atomicAdd(a.pcount, nsc);
}

template <int N, int NV>
static hipError_t launch_positivity(const LimArgs& a, double* q, hipStream_t s) {
    if constexpr (NV == 4) {
        if (a.positivity && !a.detect_only)
            hipLaunchKernelGGL((positivity_kernel<N>), dim3((a.K + 15) / 16), dim3(256), 0, s, a, q);
    }
    return hipGetLastError();
}

template <int N, int NV>
static hipError_t launch_nv(const LimArgs& a, double* q, hipStream_t s) {
    const int groups = (a.K + 15) / 16;
    hipLaunchKernelGGL((tvb_moments_kernel<N, NV>), dim3(groups), dim3(256), 0, s, a, q);
    if constexpr (NV == 4 && mlp_inputs(LShape<N>::Np) <= kMlpInMax) {
        if (a.use_mlp) {
            hipLaunchKernelGGL((tvb_limit_kernel<N, NV, true>), dim3(groups), dim3(256), 0, s, a, q);
            return launch_positivity<N, NV>(a, q, s);
        }
    }
    if (a.use_mlp) return hipErrorInvalidValue;
    hipLaunchKernelGGL((tvb_limit_kernel<N, NV, false>), dim3(groups), dim3(256), 0, s, a, q);
    return launch_positivity<N, NV>(a, q, s);
}

template <int N>
static hipError_t launch_n(const LimArgs& a, double* q, hipStream_t s) {
    if (a.nvar == 1) return launch_nv<N, 1>(a, q, s);
    if (a.nvar == 4) return launch_nv<N, 4>(a, q, s);
    return hipErrorInvalidValue;
}

template <int N>
static hipError_t features_n(const LimArgs& a, const double* q, float* X, hipStream_t s) {
    if constexpr (mlp_inputs(LShape<N>::Np) <= kMlpInMax) {
        const int groups = (a.K + 15) / 16;
        hipLaunchKernelGGL((tvb_moments_kernel<N, 4>), dim3(groups), dim3(256), 0, s, a, q);
        hipLaunchKernelGGL((tvb_features_kernel<N>), dim3(groups), dim3(256), 0, s, a, q, X);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

hipError_t launch_features(int N, const LimArgs& a, const double* q, float* X, hipStream_t s) {
    if (a.nvar != 4) return hipErrorInvalidValue;
    switch (N) {
        case 1: return features_n<1>(a, q, X, s);
        case 2: return features_n<2>(a, q, X, s);
        case 3: return features_n<3>(a, q, X, s);
        case 4: return features_n<4>(a, q, X, s);
        case 5: return features_n<5>(a, q, X, s);
        case 6: return features_n<6>(a, q, X, s);
        default: return hipErrorInvalidValue;   // 2 Np + 8 inputs must fit kMlpInMax
    }
}

hipError_t launch_positivity(int N, const LimArgs& a, double* q, hipStream_t s) {
    if (a.nvar != 4) return hipErrorInvalidValue;
    switch (N) {
        case 1: return launch_positivity<1, 4>(a, q, s);
        case 2: return launch_positivity<2, 4>(a, q, s);
        case 3: return launch_positivity<3, 4>(a, q, s);
        case 4: return launch_positivity<4, 4>(a, q, s);
        case 5: return launch_positivity<5, 4>(a, q, s);
        case 6: return launch_positivity<6, 4>(a, q, s);
        case 7: return launch_positivity<7, 4>(a, q, s);
        case 8: return launch_positivity<8, 4>(a, q, s);
        default: return hipErrorInvalidValue;
    }
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
