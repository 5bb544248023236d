// TVB minmod limiter for triangles (Cockburn & Shu 1998, the triangle limiter), component-wise on the
// conserved variables, for gfx950.  Not part of the reference (whose Trianglelimite.C:61-852 is a
// different scheme); it sits where the reference's limiteSchemes::limite hook is called
// (limiteScheme.H:139-159) and replaces that whole-field call.
//
// Two kernels, one wavefront per group of 16 elements (lane l: element l % 16, row group l / 16), no block
// barriers inside the group loop:
//   tvb_moments_kernel  cell mean and the P1 part at the three edge midpoints of every variable: the 4 x Np
//                       moment operator times the group's nodal values, per-lane partial sums and a fixed
//                       cross-lane combination (lim_moments.hpp), the state read one contiguous 512-byte row
//                       set per K step;
//   tvb_limit_kernel    lanes of row group v take variable v: per edge f, d = mbar(ut(m_f) - ubar,
//                       nu D(m_f)) from the neighbour (or ghost) means; a cell is troubled when any d differs
//                       from its argument; troubled cells are rewritten in place as ubar + sum dh_f phi_f with
//                       the positivity rescaling of dh (all three edges of a variable in one lane).
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
//   log10(max_i |ut(m_i) - ubar_0| / (M h^2) + 1e-30) clipped to [-10, 10] for rho and E (the TVB scale).
// The modal coefficients come from f64 MFMA products whose accumulators are already in the bf16 B layout of
// the network's first layer (lim_cff_slot), so the features never leave registers, and each wavefront
// evaluates its 16 cells on BF16 MFMA (mlp_infer.hpp).  Only cells the network flags compute limited slopes.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dg_common.hpp"
#include "lim_moments.hpp"
#include "limiter.hpp"
#include "mlp_infer.hpp"

namespace dg {

template <int N>
struct LShape {
    static constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
};

typedef double lim_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ lim_d4 mfma_f64(double a, double b, lim_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// B fragments of one variable of a group (base: its first value): node 4 s + lane / 16 of cell lane % 16 at
// K step s, zero beyond Np
template <int Np, int NS>
__device__ __forceinline__ void lim_load_var(const double* __restrict__ q, size_t base, int lane, double (&b)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int n = 4 * s + (lane >> 4);
        b[s] = n < Np ? q[base + 64 * s + lane] : 0.0;
    }
}

// Moments of every variable of a group (lim_moments.hpp): lane l holds the nodes lim_mnode(s, l / 16) of element
// l % 16, four 128-byte rows of the group per K step s, and stores moment l / 16.
template <int N, int NV>
__global__ __launch_bounds__(256) void tvb_moments_kernel(LimArgs a, const double* __restrict__ q) {
    constexpr int Np = LShape<N>::Np, NS = lim_msteps(Np);
    const int lane = threadIdx.x & 63, groups = (a.K + 15) >> 4;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g < groups) {
        double b[NV][NS];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int n = lim_mnode(s, lane >> 4);
                b[v][s] = n < Np ? q[(size_t(g) * NV + v) * Np * 16 + n * 16 + (lane & 15)] : 0.0;
            }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < NS; ++s) lim_moment_acc(a.p4f, s, lane, b[v][s], part);
            a.mom[(size_t(g) * NV + v) * 64 + lane] = lim_moment_reduce(part, lane);
        }
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

// Limited slopes r[0..2] of variable v of cell col (mean u0, longest edge h): per edge the edge-midpoint
// deviation, left alone when |dev| <= M h^2, else minmod(dev, nu D) with D the neighbour-pair reconstruction;
// returns whether some r differs from its deviation.
template <int N, int NV>
__device__ __forceinline__ int tvb_slopes(const LimArgs& a, int g, int col, int e, int v, double u0, double h,
                                          double (&r)[3]) {
    int chg = 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const double dev = a.mom[((size_t(g) * NV + v) * 4 + 1 + f) * 16 + col] - u0;
        const int j = a.jsel[f * a.K + e];
        const double a1 = a.lgeo[(size_t(g) * kLimGeo + 2 * f) * 16 + col];
        const double a2 = a.lgeo[(size_t(g) * kLimGeo + 2 * f + 1) * 16 + col];
        const double D = a1 * (outer_mean<N, NV>(a, g, col, e, v, f) - u0) +
                         a2 * (outer_mean<N, NV>(a, g, col, e, v, j) - u0);
        r[f] = fabs(dev) <= a.M * h * h ? dev : minmod2(dev, a.nu * D);
        chg |= r[f] != dev;
    }
    return chg;
}

// MLP input 2 Np + x (x = 0..7) of cell col: x < 6 the jump to neighbour mean f = x % 3 of rho (x < 3) or E,
// x = 6, 7 the TVB scale of rho, E
template <int N>
__device__ float lim_extra_feature(const LimArgs& a, int g, int col, int e, int x) {
    constexpr int NV = 4;
    if (x < 6) {
        const int w = x / 3, f = x - 3 * w, v = 3 * w;
        const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + col];
        const double nb = outer_mean<N, NV>(a, g, col, e, v, f);
        return float((nb - u0) / (fabs(u0) + 1e-12));
    }
    const int v = 3 * (x - 6);
    const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + col];
    double dmax = 0.0;
    for (int i = 0; i < 3; ++i) dmax = fmax(dmax, fabs(a.mom[((size_t(g) * NV + v) * 4 + 1 + i) * 16 + col] - u0));
    const double h = a.lgeo[(size_t(g) * kLimGeo + 6) * 16 + col];
    const double t = log10(dmax / (a.M * h * h) + 1e-30);
    return float(fmin(10.0, fmax(-10.0, t)));
}

template <int N>
struct LimMlp {
    static constexpr int Np = LShape<N>::Np, NS = lim_ksteps(Np), nin = mlp_inputs(Np), KS = mlp_pad(nin) / 16;
    static constexpr int pairs = lim_cff_pairs(Np, KS);
};

// Global-memory inputs of the features of group g: the B fragments of rho and E and the 8 inputs after the
// modes, of which row group kg computes x = kg and kg + 4.
template <int N>
struct LimFeatIn {
    double br[LimMlp<N>::NS], bE[LimMlp<N>::NS];
    float ex0, ex1;
};

template <int N>
__device__ __forceinline__ void lim_features_load(const LimArgs& a, const double* __restrict__ q, int g, int lane,
                                                  LimFeatIn<N>& in) {
    using S = LimMlp<N>;
    const int col = lane & 15, kg = lane >> 4;
    const int e0 = g * 16 + col, e = e0 < a.K ? e0 : a.K - 1;
    lim_load_var<S::Np, S::NS>(q, size_t(g) * 4 * S::Np * 16, lane, in.br);
    lim_load_var<S::Np, S::NS>(q, (size_t(g) * 4 + 3) * S::Np * 16, lane, in.bE);
    in.ex0 = lim_extra_feature<N>(a, g, col, e, kg);
    in.ex1 = lim_extra_feature<N>(a, g, col, e, kg + 4);
}

// MLP inputs of the 16 cells of a group (4 variables) in the bf16 B layout of layer 1: x[t][r] = input
// 16 t + 4 (lane / 16) + r of cell lane % 16, zero beyond 2 Np + 8.  cff: the feature-operator fragments.
// Every lane of the wavefront must call it (cross-lane reads).
template <int N>
__device__ __forceinline__ void lim_features_compute(const double* cff, const LimFeatIn<N>& in, int lane,
                                                     float (&x)[LimMlp<N>::KS][4]) {
    using S = LimMlp<N>;
    constexpr int Np = S::Np, NS = S::NS, KS = S::KS;
    const int col = lane & 15, kg = lane >> 4;
    lim_d4 c[KS];
#pragma unroll
    for (int t = 0; t < KS; ++t) c[t] = lim_d4{0.0, 0.0, 0.0, 0.0};
    // K steps outer, tiles inner: consecutive MFMAs feed different accumulators (each still sums its rho
    // steps, then its E steps, in order)
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int t = 0; t < KS; ++t)
                if (lim_cff_active(Np, t, w))
                    c[t] = mfma_f64(cff[(lim_cff_slot(Np, t, w) * NS + s) * 64 + lane], w ? in.bE[s] : in.br[s], c[t]);
    // mean modes: input 0 (rho) in row group 0, input Np (E) in tile Np / 16
    constexpr int tE = Np / 16, rgE = (Np % 16) / 4, rE = Np % 4;
    const double ir = 1.0 / (fabs(__shfl(c[0][0], col, 64)) + 1e-12);
    const double iE = 1.0 / (fabs(__shfl(c[tE][rE], col + 16 * rgE, 64)) + 1e-12);
#pragma unroll
    for (int t = 0; t < KS; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int f = 16 * t + 4 * kg + r;
            float val = 0.f;
            if (f < Np) val = float(c[t][r] * ir);
            else if (f < 2 * Np) val = float(c[t][r] * iE);
            if (16 * t + r + 12 >= 2 * Np && 16 * t + r < S::nin) {   // some row group of this slot holds an extra
                const int xi = f - 2 * Np, src = col + 16 * (xi & 3);
                const float v0 = __shfl(in.ex0, src, 64), v1 = __shfl(in.ex1, src, 64);
                if (xi >= 0 && f < S::nin) val = xi < 4 ? v0 : v1;
            }
            x[t][r] = val;
        }
}

template <int N>
__global__ __launch_bounds__(256) void tvb_features_kernel(LimArgs a, const double* __restrict__ q,
                                                           float* __restrict__ X) {
    using S = LimMlp<N>;
    const int lane = threadIdx.x & 63, col = lane & 15, kg = lane >> 4, groups = (a.K + 15) >> 4;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g < groups) {
        LimFeatIn<N> in;
        lim_features_load<N>(a, q, g, lane, in);
        float x[S::KS][4];
        lim_features_compute<N>(a.cff, in, lane, x);
        if (g * 16 + col < a.K) {
#pragma unroll
            for (int t = 0; t < S::KS; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int f = 16 * t + 4 * kg + r;
                    if (f < S::nin) X[size_t(g * 16 + col) * S::nin + f] = x[t][r];
                }
        }
    }
}

// Troubled cell (bad) of group g: lanes of row group v rewrite variable v from the mean u0 and the limited
// slopes r, rescaled so that they sum to zero (the dh of Cockburn-Shu); flags and the troubled count.
template <int N, int NV>
__device__ __forceinline__ void tvb_finish(const LimArgs& a, double* __restrict__ q, int g, int lane, int bad,
                                           double u0, const double (&r)[3]) {
    constexpr int Np = LShape<N>::Np;
    const int col = lane & 15, kg = lane >> 4, e0 = g * 16 + col;
    const bool live = e0 < a.K;
    if (bad && kg < NV && live && !a.detect_only) {
        double d0 = r[0], d1 = r[1], d2 = r[2];
        if (d0 + d1 + d2 != 0.0) {
            const double pos = fmax(0.0, d0) + fmax(0.0, d1) + fmax(0.0, d2);
            const double neg = fmax(0.0, -d0) + fmax(0.0, -d1) + fmax(0.0, -d2);
            const double tp = pos > 0.0 ? fmin(1.0, neg / pos) : 0.0;
            const double tn = neg > 0.0 ? fmin(1.0, pos / neg) : 0.0;
            d0 = tp * fmax(0.0, d0) - tn * fmax(0.0, -d0);
            d1 = tp * fmax(0.0, d1) - tn * fmax(0.0, -d1);
            d2 = tp * fmax(0.0, d2) - tn * fmax(0.0, -d2);
        }
        double* qv = q + (size_t(g) * NV + kg) * Np * 16 + col;
        for (int n = 0; n < Np; ++n)
            qv[n * 16] = u0 + d0 * a.phi[n * 3] + d1 * a.phi[n * 3 + 1] + d2 * a.phi[n * 3 + 2];
    }
    if (kg == 0 && live) a.troubled[e0] = bad;
    // every cell limited: the host knows the count; otherwise one atomic per wavefront with troubled cells
    const unsigned long long m = __ballot(kg == 0 && live && bad);
    if (lane == 0 && m && !a.all) atomicAdd(a.count, unsigned(__popcll(m)));
}

// TVB detection (or every cell): one group per wavefront, row group kg = variable kg.
template <int N, int NV>
__global__ __launch_bounds__(256) void tvb_limit_kernel(LimArgs a, double* __restrict__ q) {
    const int lane = threadIdx.x & 63, col = lane & 15, kg = lane >> 4;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= (a.K + 15) >> 4) return;
    const int e0 = g * 16 + col, e = e0 < a.K ? e0 : a.K - 1;
    const bool vl = kg < NV;
    const int v = vl ? kg : 0;
    const double u0 = a.mom[((size_t(g) * NV + v) * 4) * 16 + col];
    const double h = a.lgeo[(size_t(g) * kLimGeo + 6) * 16 + col];
    double r[3] = {0.0, 0.0, 0.0};
    int chg = vl ? tvb_slopes<N, NV>(a, g, col, e, v, u0, h, r) : 0;
    chg |= __shfl_xor(chg, 16, 64);
    chg |= __shfl_xor(chg, 32, 64);
    tvb_finish<N, NV>(a, q, g, lane, chg | a.all, u0, r);
}

// Learned detection (4 variables): one group of 16 cells per wavefront.  The feature operator and the
// network's fragments (about 35 KB at N = 6) are read by every wavefront straight from memory: they stay
// resident in the caches, and no block barrier holds the wavefronts back (staging them in LDS measured slower).
// Only flagged cells compute limited slopes.
template <int N>
__global__ __launch_bounds__(256) void tvb_mlp_limit_kernel(LimArgs a, double* __restrict__ q) {
    constexpr int NV = 4;
    using S = LimMlp<N>;
    const int lane = threadIdx.x & 63, col = lane & 15, kg = lane >> 4;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= (a.K + 15) >> 4) return;
    LimFeatIn<N> in;
    lim_features_load<N>(a, q, g, lane, in);
    float x[S::KS][4];
    lim_features_compute<N>(a.cff, in, lane, x);
    mlp_s4 xb[S::KS];
#pragma unroll
    for (int t = 0; t < S::KS; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) xb[t][j] = static_cast<short>(f2bf_dev(x[t][j]));
    const float z = mlp_infer_regs<S::KS>(reinterpret_cast<const mlp_s4*>(a.mlp.W1f),
                                          reinterpret_cast<const mlp_s4*>(a.mlp.W2f), a.mlp.b1, a.mlp.b2, a.mlp.w3,
                                          a.mlp.b3, xb, lane);
    const int bad = z >= a.mlp.zthr;
    const int e0 = g * 16 + col, e = e0 < a.K ? e0 : a.K - 1;
    const double u0 = a.mom[((size_t(g) * NV + kg) * 4) * 16 + col];
    double r[3] = {0.0, 0.0, 0.0};
    if (bad && !a.detect_only) {
        const double h = a.lgeo[(size_t(g) * kLimGeo + 6) * 16 + col];
        tvb_slopes<N, NV>(a, g, col, e, kg, u0, h, r);
    }
    tvb_finish<N, NV>(a, q, g, lane, bad, u0, r);
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

// blocks of the wave-per-group kernels: one group per wavefront
static int lim_grid(const LimArgs& a, int waves) {
    const int groups = (a.K + 15) / 16;
    return (groups + waves - 1) / waves;
}

template <int N, int NV>
static hipError_t launch_nv(const LimArgs& a, double* q, hipStream_t s) {
    if (!a.moments_ready) hipLaunchKernelGGL((tvb_moments_kernel<N, NV>), dim3(lim_grid(a, 4)), dim3(256), 0, s, a, q);
    if constexpr (NV == 4 && mlp_inputs(LShape<N>::Np) <= kMlpInMax) {
        if (a.use_mlp) {
            if (!a.cff) return hipErrorInvalidValue;
            hipLaunchKernelGGL((tvb_mlp_limit_kernel<N>), dim3(lim_grid(a, 4)), dim3(256), 0, s, a, q);
            return launch_positivity<N, NV>(a, q, s);
        }
    }
    if (a.use_mlp) return hipErrorInvalidValue;
    hipLaunchKernelGGL((tvb_limit_kernel<N, NV>), dim3(lim_grid(a, 4)), dim3(256), 0, s, a, q);
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
        if (!a.cff) return hipErrorInvalidValue;
        hipLaunchKernelGGL((tvb_moments_kernel<N, 4>), dim3(lim_grid(a, 4)), dim3(256), 0, s, a, q);
        hipLaunchKernelGGL((tvb_features_kernel<N>), dim3(lim_grid(a, 4)), dim3(256), 0, s, a, q, X);
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

// ---------------------------------------------------------------- scratch of dgh_limit_noalloc
// Moments, troubled flags and counters of one limiter call without device allocation (layout in limiter.hpp).
// Each GPU holds its own copy until the process ends; calls on one GPU share it, so calls that overlap on one
// GPU race.
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

hipError_t noalloc_scratch(int device, void** ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices || !ptr) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));   // the caller made `device` current
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

__global__ void zero_counts_kernel(unsigned* count) {
    if (threadIdx.x < 2) count[threadIdx.x] = 0u;
}

hipError_t launch_zero_counts(unsigned* count, hipStream_t s) {
    hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, s, count);
    return hipGetLastError();
}

}  // namespace dg
