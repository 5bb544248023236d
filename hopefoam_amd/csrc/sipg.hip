// Matrix-free SIPG Laplacian (dgm::laplacian) and the kernels of its preconditioned CG, gfx950.
//
// Operator: the reference's NEGATIVE SIPG form (HopeFOAM-0.1/src/DG/DG/laplacianSchemes/
// defaultLaplacianScheme/defaultLaplacianScheme.C:62-288), applied without assembling it:
//   (A u)_e = -K_e u_e + sum over faces of
//        interior:  Vf^T W ( -tau/2 [u] + {n.grad u} ) + Dn^T W [u] / 2
//        Dirichlet: Vf^T W ( -tau u + n.grad u ) + Dn^T W u
// with [u] = u_e - u_other and {n.grad u} = (n_e.grad u_e + n_e.grad u_other) / 2 at the face Gauss
// points, W = diag(Js w_g) of the face's owner, Dn = n_e.grad of e's basis at the Gauss points and
// K_e = J (c1 Srr + c2 (Srs + Srs^T) + c3 Sss) the cubature stiffness (:96-102).
//
// Every kernel works on one group of 16 elements (the field layout's unit): thread t handles element
// t % 16 of the group and node / face slot t / 16 + 16 i, so each global access of a node row is 16
// consecutive values.  One application = trace kernel (u and n.grad u at all face Gauss points) +
// apply kernel (reads its own and its neighbours' traces).
//
// The CG kernels run PCG in the Chronopoulos & Gear form (KSPCG's iterates and convergence test; dgMatrix.C:163-283
// selects KSPCG) in two launches per iteration: the update kernel (direction and iterate updates, u = B r, the face
// traces of u, u.u and r.u partials) and the apply kernel (w = S u, u.w partials).  Dot products are reduced per
// block; every consumer re-sums the partials in one fixed order (deterministic, the same bits in every block) and
// evaluates the scalars and the convergence test there.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dg_common.hpp"
#include "lds_stage.hpp"
#include "sipg.hpp"

namespace dg {

template <int N>
struct SG {
    static constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1, Nfq = N + 2;
    static constexpr int NFQ3 = 3 * Nfq, NFP3 = 3 * Nfp;
};

__device__ __forceinline__ size_t trace_index(int g, int item, int c, int el, int nfq3) {
    return ((size_t(g) * nfq3 + item) * 2 + c) * 16 + el;
}

// ---------------------------------------------------------------- reductions
// Block sum of NV values (256 threads); returns the sums in thread 0.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
    __shared__ double red[NV][4];
    for (int k = 0; k < NV; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < NV; ++k) red[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) v[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    __syncthreads();
}

// LDS written by some lanes of a wavefront and read by others: no block barrier, the wave's own ordering
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// s0 = u.u (KSPCG's preconditioned norm), s1 = r.u, s2 = u.S u of the iterate the partials belong to.
// Chronopoulos & Gear: gamma = r.u, beta = gamma / gamma_old, p.S p = u.S u - beta gamma / alpha_old,
// alpha = gamma / p.S p; the convergence and breakdown tests are KSPCG's (dgMatrix.C:163-283 selects KSPCG).
__device__ void cg_finish(CgState* st, int mode, double s0, double s1, double s2) {
    if (mode == CG_FINISH_NORM) {
        st->dp = sqrt(s0);
        return;
    }
    const double dp = sqrt(s0);
    st->dp = dp;
    if (mode == CG_FINISH_INIT) {
        st->rnorm0 = dp;
        st->ttol = fmax(st->rtol * dp, st->abstol);
        st->beta = s1;
        st->its = 0;
        if (!std::isfinite(dp)) {
            st->done = 1;
            st->reason = -9;
            return;
        } else if (dp <= st->ttol) {
            st->done = 1;
            st->reason = 1;
            return;
        } else if (st->maxit <= 0) {
            st->done = 1;
            st->reason = -3;
            return;
        }
        st->bcoef = 0.0;
    } else {   // CG_FINISH_ITER
        st->betaold = st->beta;
        st->beta = s1;
        st->its += 1;
        if (!std::isfinite(dp)) {
            st->done = 1;
            st->reason = -9;
            return;
        } else if (dp <= st->ttol) {
            st->done = 1;
            st->reason = 1;
            return;
        } else if (dp >= 1e4 * st->rnorm0) {   // PETSc default divergence tolerance
            st->done = 1;
            st->reason = -4;
            return;
        } else if (st->its >= st->maxit) {
            st->done = 1;
            st->reason = -3;
            return;
        } else if (s1 == 0.0) {                 // the next iteration would stop on beta = 0
            st->done = 1;
            st->its += 1;
            st->reason = 2;
            return;
        }
        st->bcoef = st->beta / st->betaold;
    }
    // KSP_DIVERGED_INDEFINITE_MAT: p.Ap must stay positive for the SPD operator
    const double pw = mode == CG_FINISH_INIT ? s2 : s2 - st->bcoef * st->beta / st->alpha;
    st->pw = pw;
    if (!(pw > 0.0)) {
        st->done = 1;
        st->reason = std::isfinite(pw) ? -8 : -9;
        return;
    }
    st->alpha = st->beta / pw;
}

// Writes this block's partial sums (NV <= 2) to part[blockIdx.x][0..1]; the finishes sum them.
template <int NV>
__device__ void write_partials(double (&v)[NV], double* part) {
    block_reduce<NV>(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) part[size_t(blockIdx.x) * 2 + k] = v[k];
}

// Sums the nb partials in a fixed order (deterministic) into st->dp; one block.
__global__ __launch_bounds__(256) void norm_finish_kernel(CgState* __restrict__ st, const double* __restrict__ part,
                                                          int nb) {
    double s[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += 256) s[0] += part[size_t(b) * 2];
    block_reduce<1>(s);
    if (threadIdx.x == 0) cg_finish(st, CG_FINISH_NORM, s[0], 0.0, 0.0);
}

// The CG state one finish further, evaluated inside the kernel that consumes it: every block (256 threads)
// reads st_in, sums the update kernel's (u.u, r.u) partials pb and the apply's u.w partials pa (nb blocks each)
// in a fixed order (so every block gets the same bits) and applies the finish (INIT while its < 0, else ITER;
// none for CG_FINISH_NONE); block 0 publishes the result to st_out.  st_in is never written by the same launch,
// so the blocks cannot see each other's update.  A state that is already done passes through unchanged.
__device__ void cg_advance(const CgState* __restrict__ st_in, CgState* __restrict__ st_out,
                           const double* __restrict__ pb, const double* __restrict__ pa, int nb, int mode,
                           CgState& ls) {
    // the partials are read together with the state (a done state ignores them), so the two latencies overlap
    double s[3] = {0.0, 0.0, 0.0};
    if (mode != CG_FINISH_NONE)
        for (int b = threadIdx.x; b < nb; b += 256) {
            s[0] += pb[size_t(b) * 2];
            s[1] += pb[size_t(b) * 2 + 1];
            s[2] += pa[size_t(b) * 2];
        }
    if (threadIdx.x == 0) ls = *st_in;
    __syncthreads();
    if (!ls.done && mode != CG_FINISH_NONE) {
        block_reduce<3>(s);
        if (threadIdx.x == 0) cg_finish(&ls, ls.its < 0 ? CG_FINISH_INIT : CG_FINISH_ITER, s[0], s[1], s[2]);
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) *st_out = ls;
}

// st_out = st_in one finish further (the state the next update would see), for the host's look at the device
// state between queued iterations
__global__ __launch_bounds__(256) void cg_publish_kernel(const CgState* __restrict__ st_in, CgState* __restrict__ st_out,
                                                         const double* __restrict__ pb, const double* __restrict__ pa,
                                                         int nb) {
    __shared__ CgState ls;
    cg_advance(st_in, st_out, pb, pa, nb, CG_FINISH_AUTO, ls);
}

// ---------------------------------------------------------------- operator kernels
typedef double sipg_d4 __attribute__((ext_vector_type(4)));
typedef float sipg_f4 __attribute__((ext_vector_type(4)));
template <typename T>
struct SV4;
template <>
struct SV4<double> { typedef sipg_d4 type; };
template <>
struct SV4<float> { typedef sipg_f4 type; };

__device__ __forceinline__ sipg_d4 smfma(double a, double b, sipg_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ sipg_f4 smfma(float a, float b, sipg_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ T select3(int f, T a0, T a1, T a2) {
    return f == 0 ? a0 : (f == 1 ? a1 : a2);
}

// Face terms of one Gauss point: av multiplies Vface^T (values), cv multiplies Dn^T (normal derivatives).
template <typename T>
__device__ __forceinline__ void face_terms(int cn, T w, T tau, T ue, T dne, const T* __restrict__ trace, int item_n,
                                           int nfq3, T& av, T& cv) {
    av = T(0);
    cv = T(0);
    if (cn >= 0) {
        const int en = cn >> 2;
        const size_t tn = trace_index(en >> 4, item_n, 0, en & 15, nfq3);
        const T un = trace[tn];
        const T dnn = -trace[tn + 16];   // n_e = -n_other
        av = w * (T(-0.5) * tau * (ue - un) + T(0.5) * (dne + dnn));
        cv = w * T(0.5) * (ue - un);
    } else if (-cn - 1 == BC_FIXED) {
        av = w * (-tau * ue + dne);
        cv = w * ue;
    }
}

// The face geometry the traces of element lane % 16 of group g need: (ar_f, as_f) of its three faces
template <typename T>
__device__ __forceinline__ void load_face_geo(const SipgArgs<T>& a, int g, int lane, T (&fg)[6]) {
    const T* gg = a.geo + size_t(g) * kSipgGeo * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 6; ++i) fg[i] = gg[(4 + i) * 16];
}

// Face traces of one group of 16 elements from its nodal values in B-fragment layout (ub[s] = node 4 s + l / 16
// of element l % 16): the trace GEMM on the matrix cores with the A fragments sA staged in LDS, then the value and
// n_e . grad u at every face Gauss point item = f * Nfq + q into trace[g][item][c][16].  fg: load_face_geo.
template <int N, typename T>
__device__ __forceinline__ void trace_group(const T* __restrict__ sA, const T (&ub)[SipgTiles(N).KS],
                                            const T (&fg)[6], int g, int lane, T* __restrict__ trace) {
    constexpr SipgTiles S(N);
    constexpr int Nfq = S.Nfq, NFQ3 = S.NFQ3, TI = S.TI, KS = S.KS;
    typedef typename SV4<T>::type V4;
    const int j = lane & 15, kq = lane >> 4;
    V4 acc[3 * TI];
#pragma unroll
    for (int t = 0; t < 3 * TI; ++t) {
        acc[t] = V4{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < KS; ++s) acc[t] = smfma(sA[(t * KS + s) * 64 + lane], ub[s], acc[t]);
    }
    const T ar0 = fg[0], as0 = fg[1], ar1 = fg[2], as1 = fg[3], ar2 = fg[4], as2 = fg[5];
#pragma unroll
    for (int t = 0; t < TI; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int item = 16 * t + mfma_row<T>(kq, r);
            if (item < NFQ3) {
                const int f = item / Nfq;
                const T dn = select3(f, ar0, ar1, ar2) * acc[TI + t][r] + select3(f, as0, as1, as2) * acc[2 * TI + t][r];
                trace[trace_index(g, item, 0, j, NFQ3)] = acc[t][r];
                trace[trace_index(g, item, 1, j, NFQ3)] = dn;
            }
        }
}

// Persistent: every block stages the A fragments in LDS once; each of its 4 wavefronts then takes groups of
// 16 elements (group = MFMA N dimension; lane l handles element l % 16 and K lane l / 16).
template <int N, typename T>
__global__ __launch_bounds__(256) void sipg_trace_kernel(SipgArgs<T> a, const T* __restrict__ u, T* __restrict__ trace) {
    constexpr SipgTiles S(N);
    constexpr int Np = S.Np, TI = S.TI, KS = S.KS;
    __shared__ T sA[3 * TI * KS * 64];
    for (int i = threadIdx.x; i < 3 * TI * KS * 64; i += 256) sA[i] = a.TrF[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, kq = lane >> 4;
    const int groups = (a.K + 15) >> 4;
    for (int g = blockIdx.x * 4 + wave; g < groups; g += gridDim.x * 4) {
        const size_t base = size_t(g) * Np * 16;
        T ub[KS], fg[6];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + kq;
            ub[s] = k < Np ? u[base + size_t(k) * 16 + j] : T(0);
        }
        load_face_geo<T>(a, g, lane, fg);
        trace_group<N, T>(sA, ub, fg, g, lane, trace);
    }
}

// What the update kernel reads of one group ahead of the CG scalars (lane: element lane % 16, nodes 4 s + lane / 16
// < Np; other entries unset): the inputs of the chain r -> u = B r -> traces (r, w, S p; init: rhs in r, and w), the
// point-Jacobi diagonal or the element's inverse-block index, and the face geometry of the traces.  The rest (u, p,
// x) is read when needed: nothing downstream waits for the p and x updates.
template <typename T, int KS>
struct UpdateIn {
    T w[KS], ap[KS], r[KS], d[KS];
    T fg[6];
    int bt;
};

template <int N, typename T, bool BLOCK>
__device__ __forceinline__ void update_load(const SipgArgs<T>& a, const CgLoop<T>& c, int init, int g, int lane,
                                            UpdateIn<T, SipgTiles(N).KS>& in) {
    constexpr SipgTiles S(N);
    constexpr int Np = S.Np, KS = S.KS;
    const int j = lane & 15, kq = lane >> 4;
    const size_t base = size_t(g) * Np * 16;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + kq;
        if (k < Np) {
            const size_t idx = base + size_t(k) * 16 + j;
            in.w[s] = c.w[idx];
            if (init) {
                in.r[s] = c.rhs[idx];
            } else {
                in.ap[s] = c.ap[idx];
                in.r[s] = c.r[idx];
            }
            if constexpr (!BLOCK) in.d[s] = c.dinv[idx];
        }
    }
    if constexpr (BLOCK) in.bt = c.bidx[g * 16 + j];   // padded to whole groups
    load_face_geo<T>(a, g, lane, in.fg);
}

// The update kernel of PCG iteration i (Chronopoulos & Gear; one wavefront per group as in the trace kernel):
// finishes the previous reduction (cg_advance), then p = u + beta p, ap = w + beta ap (p = u, ap = w in the first
// iteration), x += alpha p, r -= alpha ap, u = B r with its face traces, and the u.u, r.u partials.  init: only
// r = rhs - w, u = B r, traces and partials (w = S x on entry).  B = diag(dinv) (point Jacobi) or the element's
// block inverse binv[bidx[e]] applied from a per-wavefront LDS copy of the group's r.
template <int N, typename T, bool BLOCK>
__global__ __launch_bounds__(256, N <= 6 ? 2 : 1) void cg_update_kernel(SipgArgs<T> a, CgLoop<T> c, int init,
                                                        const CgState* __restrict__ st_in, CgState* __restrict__ st_out,
                                                        const double* __restrict__ pb_in, double* __restrict__ pb_out) {
    constexpr SipgTiles S(N);
    constexpr int Np = S.Np, TI = S.TI, KS = S.KS, MT = S.MT;
    typedef typename SV4<T>::type V4;
    static_assert(KS <= 4 * MT, "node steps beyond the row tiles");
    __shared__ CgState ls;
    constexpr size_t kTrBytes = size_t(3 * TI * KS * 64) * sizeof(T);
    __shared__ __attribute__((aligned(16))) T sA[dma_span(kTrBytes) / sizeof(T)];
    const int nb = gridDim.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, kq = lane >> 4;
    const int groups = (a.K + 15) >> 4;
    // The trace fragments (LDS DMA) and the first group's inputs do not depend on the CG scalars: they are
    // requested before the reduction that yields them, so the latencies overlap.
    dma_stage_to_lds<kTrBytes>(sA, a.TrF);
    const int g0 = blockIdx.x * 4 + wave;
    UpdateIn<T, KS> in;
    if (g0 < groups) update_load<N, T, BLOCK>(a, c, init, g0, lane, in);
    cg_advance(st_in, st_out, pb_in, c.part_a, nb, init ? CG_FINISH_NONE : CG_FINISH_AUTO, ls);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the fragments have landed (and no DMA outlives the block)
    if (ls.done) return;
    __syncthreads();
    const bool first = !init && ls.its == 0;
    const T alpha = init ? T(0) : T(ls.alpha), bc = T(ls.bcoef);
    double acc[2] = {0.0, 0.0};
    for (int g = g0; g < groups; g += gridDim.x * 4) {
        if (g != g0) update_load<N, T, BLOCK>(a, c, init, g, lane, in);
        const size_t base = size_t(g) * Np * 16;
        const int e = g * 16 + j;
        const bool live = e < a.K;
        T rv[KS], ub[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + kq;
            rv[s] = T(0);
            if (k < Np) {
                const size_t idx = base + size_t(k) * 16 + j;
                T r;
                if (init) {
                    r = in.r[s] - in.w[s];
                } else {
                    const T av = first ? in.w[s] : in.w[s] + bc * in.ap[s];
                    // one rounding (fused), whatever the compiler makes of the init branch beside it
                    r = fma(-alpha, av, in.r[s]);
                    const T pv = first ? c.u[idx] : c.u[idx] + bc * c.p[idx];
                    c.p[idx] = pv;
                    c.ap[idx] = av;
                    c.x[idx] += alpha * pv;
                }
                c.r[idx] = r;
                rv[s] = r;
            }
        }
        if constexpr (BLOCK) {
            // u = B_t r for each distinct inverse block t of the group, on the matrix cores: r is already in
            // B-fragment layout and the fragments' row order puts output node 4 s' + kq in register s' % 4 of
            // row tile s' / 4 (block_fragments), so the result lands in the same layout
            const int bt = in.bt;
            bool mine = false;
#pragma unroll
            for (int s = 0; s < KS; ++s) ub[s] = T(0);
            for (;;) {
                const unsigned long long todo = __ballot(!mine);
                if (todo == 0ull) break;
                const int t = __shfl(bt, __ffsll(static_cast<long long>(todo)) - 1, 64);
                const T* F = c.binv + size_t(t) * MT * KS * 64 + lane;
                V4 bacc[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    bacc[m] = V4{T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int s = 0; s < KS; ++s) bacc[m] = smfma(F[(m * KS + s) * 64], rv[s], bacc[m]);
                }
                if (bt == t) {
                    mine = true;
#pragma unroll
                    for (int s = 0; s < KS; ++s) ub[s] = live ? bacc[s / 4][s % 4] : T(0);
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + kq;
                ub[s] = k < Np ? in.d[s] * rv[s] : T(0);
            }
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + kq;
            if (k < Np) {
                c.u[base + size_t(k) * 16 + j] = ub[s];
                if (live) {
                    acc[0] += double(ub[s]) * double(ub[s]);
                    acc[1] += double(ub[s]) * double(rv[s]);
                }
            }
        }
        trace_group<N, T>(sA, ub, in.fg, g, lane, c.trace);
    }
    write_partials<2>(acc, pb_out);
}

// What the apply kernel needs of one group (lane: element lane % 16, K lane lane / 16): u in B-fragment layout,
// J and the stiffness coefficients -J c1..c3, and the face terms av (times Vface^T) and (ar, as) cv (times Dr^T,
// Ds^T) at the face Gauss points 4 s + lane / 16 from its own and its neighbours' traces.
template <typename T, int KU, int KI>
struct ApplyIn {
    T uv[KU], av[KI], cr[KI], cs[KI];
    T J, m1, m2, m3;
};

template <int N, typename T>
__device__ __forceinline__ void apply_load(const SipgArgs<T>& a, const T* __restrict__ u, const T* __restrict__ trace,
                                           int g, int lane, ApplyIn<T, SipgTiles(N).NPP / 4, SipgTiles(N).NIP / 4>& in) {
    constexpr SipgTiles S(N);
    constexpr int Np = S.Np, Nfq = S.Nfq, NFQ3 = S.NFQ3, KU = S.NPP / 4, KI = S.NIP / 4;
    const int j = lane & 15, kq = lane >> 4;
    const size_t base = size_t(g) * Np * 16;
    const int e = g * 16 + j;
    const int ec = e < a.K ? e : a.K - 1;
#pragma unroll
    for (int s = 0; s < KU; ++s) {
        const int k = 4 * s + kq;
        in.uv[s] = k < Np ? u[base + size_t(k) * 16 + j] : T(0);
    }
    const T* gg = a.geo + size_t(g) * kSipgGeo * 16 + j;
    const T J = gg[0];
    in.J = J;
    in.m1 = -J * gg[16];
    in.m2 = -J * gg[2 * 16];
    in.m3 = -J * gg[3 * 16];
#pragma unroll
    for (int s = 0; s < KI; ++s) {
        const int it = 4 * s + kq;
        in.av[s] = T(0);
        in.cr[s] = T(0);
        in.cs[s] = T(0);
        if (it < NFQ3) {
            const int f = it / Nfq, q = it - f * Nfq;
            const int cn = a.conn[f * a.K + ec];
            const T w = gg[(10 + f) * 16] * a.gw[q];
            const T ue = trace[trace_index(g, it, 0, j, NFQ3)];
            const T dne = trace[trace_index(g, it, 1, j, NFQ3)];
            const int item_n = cn >= 0 ? (cn & 3) * Nfq + (Nfq - 1 - q) : 0;   // the other side runs reversed
            T avv, cvv;
            face_terms<T>(cn, w, gg[(13 + f) * 16], ue, dne, trace, item_n, NFQ3, avv, cvv);
            in.av[s] = avv;
            in.cr[s] = gg[(4 + 2 * f) * 16] * cvv;
            in.cs[s] = gg[(5 + 2 * f) * 16] * cvv;
        }
    }
}

// out = A u (negate = 0) or -A u (negate = 1).  CG mode (st != null): out = w = S u and the u.w partials, nothing
// once st->done is set.
template <int N, typename T>
__global__ __launch_bounds__(256, N <= 6 ? 2 : 1) void sipg_apply_kernel(SipgArgs<T> a, const T* __restrict__ u,
                                                         const T* __restrict__ trace, T* __restrict__ out, int negate,
                                                         CgState* __restrict__ st, double* __restrict__ part) {
    constexpr SipgTiles S(N);
    constexpr int Np = S.Np, MT = S.MT, KA = S.KA, KU = S.NPP / 4, KI = S.NIP / 4;
    typedef typename SV4<T>::type V4;
    constexpr size_t kApBytes = size_t(MT * KA * 64) * sizeof(T);
    __shared__ __attribute__((aligned(16))) T sA[dma_span(kApBytes) / sizeof(T)];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, kq = lane >> 4;
    const int groups = (a.K + 15) >> 4;
    // The fragments (LDS DMA) and the first group's inputs are requested before anything waits, so their latencies
    // overlap (a finished CG state reads them for nothing, once).
    const int done = st ? st->done : 0;
    dma_stage_to_lds<kApBytes>(sA, a.ApF);
    const int g0 = blockIdx.x * 4 + wave;
    ApplyIn<T, KU, KI> in;
    if (g0 < groups) apply_load<N, T>(a, u, trace, g0, lane, in);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the fragments have landed (and no DMA outlives the block)
    if (done) return;
    __syncthreads();
    double acc_dot[1] = {0.0};
    for (int g = g0; g < groups; g += gridDim.x * 4) {
        if (g != g0) apply_load<N, T>(a, u, trace, g, lane, in);
        const size_t base = size_t(g) * Np * 16;
        const int e = g * 16 + j;
        const bool live = e < a.K;
        const T J = in.J, m1 = in.m1, m2 = in.m2, m3 = in.m3;
        T uv[KU], av[KI], cr[KI], cs[KI];
#pragma unroll
        for (int s = 0; s < KU; ++s) uv[s] = in.uv[s];
#pragma unroll
        for (int s = 0; s < KI; ++s) {
            av[s] = in.av[s];
            cr[s] = in.cr[s];
            cs[s] = in.cs[s];
        }
#pragma unroll 1
        for (int m = 0; m < MT; ++m) {
            V4 acc = V4{T(0), T(0), T(0), T(0)};
            const T* A = sA + size_t(m) * KA * 64 + lane;
#pragma unroll
            for (int s = 0; s < KU; ++s) acc = smfma(A[s * 64], m1 * uv[s], acc);
#pragma unroll
            for (int s = 0; s < KU; ++s) acc = smfma(A[(KU + s) * 64], m2 * uv[s], acc);
#pragma unroll
            for (int s = 0; s < KU; ++s) acc = smfma(A[(2 * KU + s) * 64], m3 * uv[s], acc);
#pragma unroll
            for (int s = 0; s < KI; ++s) acc = smfma(A[(3 * KU + s) * 64], av[s], acc);
#pragma unroll
            for (int s = 0; s < KI; ++s) acc = smfma(A[(3 * KU + KI + s) * 64], cr[s], acc);
#pragma unroll
            for (int s = 0; s < KI; ++s) acc = smfma(A[(3 * KU + 2 * KI + s) * 64], cs[s], acc);
            if (a.sigma != T(0)) {   // mass shift: the reference-sign operator gains -sigma J M (fragments from cache)
                const T ms = -a.sigma * J;
                const T* F = a.MF + size_t(m) * KU * 64 + lane;
#pragma unroll
                for (int s = 0; s < KU; ++s) acc = smfma(F[s * 64], ms * uv[s], acc);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = 16 * m + mfma_row<T>(kq, r);
                if (n < Np && live) {
                    const T o = negate ? -acc[r] : acc[r];
                    out[base + size_t(n) * 16 + j] = o;
                    if (st) {
                        // f64 rows n = 16 m + kq + 4 r keep this lane's K lane (n % 4 = kq): u from registers
                        const T un = sizeof(T) == 8 ? uv[4 * m + r] : u[base + size_t(n) * 16 + j];
                        acc_dot[0] += double(un) * double(o);
                    }
                }
            }
        }
    }
    if (st) write_partials<1>(acc_dot, part);
}

// Lifts the face terms for the right-hand side: sz = Vf^T cv and sl = Vf^T av at the face nodes; then for
// node n r += sum_f (ar_f Dr[fmask_f,:]^T + as_f Ds[fmask_f,:]^T) sz_f + sl at n's face slots.
template <int N, typename T>
__device__ __forceinline__ T face_lift(const SipgArgs<T>& a, int n, int el, const T (*sg)[16], const T (*sz)[16],
                                       const T (*sl)[16]) {
    using S = SG<N>;
    constexpr int Np = S::Np, Nfp = S::Nfp;
    T r = T(0);
    for (int f = 0; f < 3; ++f) {
        T tr = T(0), ts = T(0);
        for (int k = 0; k < Nfp; ++k) {
            const int node = a.fmask[f * Nfp + k];
            tr += a.Dr[node * Np + n] * sz[f * Nfp + k][el];
            ts += a.Ds[node * Np + n] * sz[f * Nfp + k][el];
        }
        r += sg[4 + 2 * f][el] * tr + sg[5 + 2 * f][el] * ts;
    }
    for (int h = 0; h < 2; ++h) {
        const int s = a.nface[2 * n + h];
        if (s >= 0) r += sl[s][el];
    }
    return r;
}

// b = J M f + Dirichlet terms of g (reference sign, defaultLaplacianScheme.C:153-209 and
// EquationGeometricDofField.H:84-102); gbc is [K][3][Nfp] nodal boundary values or null.  gnb ([K][3][Nfp] or
// null): the normal-derivative data of the other boundary faces (fixedGradient patches,
// defaultLaplacianScheme.C:210-220): b -= Vf^T W Vf g on those faces.
template <int N>
__global__ __launch_bounds__(256) void sipg_rhs_kernel(SipgArgs<double> a, const double* __restrict__ f,
                                                       const double* __restrict__ gbc, const double* __restrict__ gnb,
                                                       double* __restrict__ b) {
    using S = SG<N>;
    constexpr int Np = S::Np, Nfp = S::Nfp, Nfq = S::Nfq, NFQ3 = S::NFQ3, NFP3 = S::NFP3;
    __shared__ double sf[Np][16];
    __shared__ double sg[kSipgGeo][16];
    __shared__ double sa[NFQ3][16], sc[NFQ3][16];
    __shared__ double sz[NFP3][16], sl[NFP3][16];
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int e = g * 16 + el;
    const bool live = e < a.K;
    const int ec = live ? e : a.K - 1;
    const size_t base = size_t(g) * Np * 16;
    for (int i = threadIdx.x; i < Np * 16; i += 256) sf[i >> 4][i & 15] = f[base + i];
    for (int i = threadIdx.x; i < kSipgGeo * 16; i += 256) sg[i >> 4][i & 15] = a.geo[size_t(g) * kSipgGeo * 16 + i];
    __syncthreads();
    for (int it = slot; it < NFQ3; it += 16) {
        const int fc = it / Nfq, q = it - fc * Nfq;
        const int cn = a.conn[fc * a.K + ec];
        double av = 0.0, cv = 0.0;
        if (cn < 0 && -cn - 1 == BC_FIXED && gbc) {
            double gv = 0.0;
            for (int k = 0; k < Nfp; ++k) gv += a.Vf[q * Nfp + k] * gbc[(size_t(ec) * 3 + fc) * Nfp + k];
            const double w = sg[10 + fc][el] * a.gw[q];
            av = -w * sg[13 + fc][el] * gv;
            cv = w * gv;
        } else if (cn < 0 && -cn - 1 != BC_FIXED && gnb) {
            double gv = 0.0;
            for (int k = 0; k < Nfp; ++k) gv += a.Vf[q * Nfp + k] * gnb[(size_t(ec) * 3 + fc) * Nfp + k];
            av = -sg[10 + fc][el] * a.gw[q] * gv;
        }
        sa[it][el] = av;
        sc[it][el] = cv;
    }
    __syncthreads();
    for (int it = slot; it < NFP3; it += 16) {
        const int fc = it / Nfp, k = it - fc * Nfp;
        double zs = 0.0, ls = 0.0;
        for (int q = 0; q < Nfq; ++q) {
            zs += a.Vf[q * Nfp + k] * sc[fc * Nfq + q][el];
            ls += a.Vf[q * Nfp + k] * sa[fc * Nfq + q][el];
        }
        sz[it][el] = zs;
        sl[it][el] = ls;
    }
    __syncthreads();
    const double J = sg[0][el];
    for (int n = slot; n < Np; n += 16) {
        double r = 0.0;
        for (int j = 0; j < Np; ++j) r += J * a.M[n * Np + j] * sf[j][el];
        r += face_lift<N, double>(a, n, el, sg, sz, sl);
        if (live) b[base + n * 16 + el] = r;
    }
}

// ---------------------------------------------------------------- small vector kernels (group layout)
__global__ __launch_bounds__(256) void neg_live_kernel(double* __restrict__ out, const double* __restrict__ in,
                                                       size_t nfull, size_t n, int krem) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        out[i] = (i < nfull || int(i & 15) < krem) ? -in[i] : 0.0;
}
__global__ __launch_bounds__(256) void sub_live_kernel(double* __restrict__ r, const double* __restrict__ a,
                                                       const double* __restrict__ b, size_t nfull, size_t n,
                                                       int krem) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        r[i] = (i < nfull || int(i & 15) < krem) ? a[i] - b[i] : 0.0;
}
__global__ __launch_bounds__(256) void d2f_kernel(float* __restrict__ y, const double* __restrict__ x, size_t nfull,
                                                  size_t n, int krem) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        y[i] = (i < nfull || int(i & 15) < krem) ? float(x[i]) : 0.0f;
}
__global__ __launch_bounds__(256) void add_f32_kernel(double* __restrict__ y, const float* __restrict__ x, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        y[i] += double(x[i]);
}
__global__ __launch_bounds__(256) void zero_kernel(unsigned* __restrict__ x, size_t nwords) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < nwords; i += size_t(gridDim.x) * blockDim.x)
        x[i] = 0u;
}
// x.x over live entries, finished by the last block into st->dp as a 2-norm
template <typename T>
__global__ __launch_bounds__(256) void norm_kernel(const T* __restrict__ x, size_t nfull, size_t n, int krem,
                                                   double* __restrict__ part) {
    double acc[2] = {0.0, 0.0};
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        if (i < nfull || int(i & 15) < krem) acc[0] += double(x[i]) * double(x[i]);
    write_partials<2>(acc, part);
}

// ---------------------------------------------------------------- launchers
#define DG_SIPG_DISPATCH(N, CALL) \
    switch (N) { \
        case 1: CALL(1); break; case 2: CALL(2); break; case 3: CALL(3); break; case 4: CALL(4); break; \
        case 5: CALL(5); break; case 6: CALL(6); break; case 7: CALL(7); break; case 8: CALL(8); break; \
        default: return hipErrorInvalidValue; }

static int groups_of(int K) { return (K + 15) / 16; }

VecShape vec_shape(int K, int Np) {
    VecShape s;
    s.n = size_t(groups_of(K)) * 16 * Np;
    s.nfull = size_t(K / 16) * 16 * Np;
    s.krem = K % 16;
    return s;
}

int vec_blocks(size_t n) {
    const size_t b = (n + 255) / 256;
    return int(b > 2048 ? 2048 : (b < 1 ? 1 : b));
}

// persistent grid: 4 groups per block per pass, at most 4 blocks per compute unit
template <typename T>
static int sipg_blocks(const SipgArgs<T>& a) {
    const int want = (groups_of(a.K) + 3) / 4;
    return want < 4 * a.ncu ? want : 4 * a.ncu;
}

template <typename T>
static hipError_t sipg_trace(int N, const SipgArgs<T>& a, const T* u, T* trace, hipStream_t s) {
#define CALL(n) hipLaunchKernelGGL((sipg_trace_kernel<n, T>), dim3(sipg_blocks(a)), dim3(256), 0, s, a, u, trace)
    DG_SIPG_DISPATCH(N, CALL)
#undef CALL
    return hipGetLastError();
}

template <typename T>
static hipError_t sipg_apply_only(int N, const SipgArgs<T>& a, const T* u, const T* trace, T* out, int negate,
                                  CgState* st, double* part, hipStream_t s) {
#define CALL(n) \
    hipLaunchKernelGGL((sipg_apply_kernel<n, T>), dim3(sipg_blocks(a)), dim3(256), 0, s, a, u, trace, out, negate, st, \
                       part)
    DG_SIPG_DISPATCH(N, CALL)
#undef CALL
    return hipGetLastError();
}

template <typename T>
hipError_t sipg_apply(int N, const SipgArgs<T>& a, const T* u, T* trace, T* out, int negate, hipStream_t s) {
    hipError_t err = sipg_trace<T>(N, a, u, trace, s);
    if (err != hipSuccess) return err;
    return sipg_apply_only<T>(N, a, u, trace, out, negate, nullptr, nullptr, s);
}
template hipError_t sipg_apply<double>(int, const SipgArgs<double>&, const double*, double*, double*, int, hipStream_t);
template hipError_t sipg_apply<float>(int, const SipgArgs<float>&, const float*, float*, float*, int, hipStream_t);

hipError_t sipg_rhs(int N, const SipgArgs<double>& a, const double* f, const double* gbc, const double* gnb, double* b,
                    hipStream_t s) {
#define CALL(n) hipLaunchKernelGGL((sipg_rhs_kernel<n>), dim3(groups_of(a.K)), dim3(256), 0, s, a, f, gbc, gnb, b)
    DG_SIPG_DISPATCH(N, CALL)
#undef CALL
    return hipGetLastError();
}

// the update kernel: reads state st_in and partials pb_in (+ the apply's), writes st_out and pb_out
template <typename T>
static hipError_t cg_update_launch(int N, const SipgArgs<T>& a, const CgLoop<T>& c, int init, const CgState* st_in,
                                   CgState* st_out, const double* pb_in, double* pb_out, hipStream_t s) {
    if (c.binv) {
#define CALL(n)                                                                                                      \
    hipLaunchKernelGGL((cg_update_kernel<n, T, true>), dim3(sipg_blocks(a)), dim3(256), 0, s, a, c, init, st_in, \
                       st_out, pb_in, pb_out)
        DG_SIPG_DISPATCH(N, CALL)
#undef CALL
    } else {
#define CALL(n)                                                                                                       \
    hipLaunchKernelGGL((cg_update_kernel<n, T, false>), dim3(sipg_blocks(a)), dim3(256), 0, s, a, c, init, st_in, \
                       st_out, pb_in, pb_out)
        DG_SIPG_DISPATCH(N, CALL)
#undef CALL
    }
    return hipGetLastError();
}

template <typename T>
hipError_t cg_start(int N, const SipgArgs<T>& a, const CgLoop<T>& c, hipStream_t s) {
    hipError_t err = cg_update_launch<T>(N, a, c, 1, c.st + 1, c.st, c.part_b[0], c.part_b[1], s);
    if (err != hipSuccess) return err;
    return sipg_apply_only<T>(N, a, c.u, c.trace, c.w, 1, c.st, c.part_a, s);
}

template <typename T>
hipError_t cg_iteration(int N, const SipgArgs<T>& a, const CgLoop<T>& c, int i, hipStream_t s) {
    const int q = i & 1;
    hipError_t err = cg_update_launch<T>(N, a, c, 0, c.st + q, c.st + (q ^ 1), c.part_b[q ^ 1], c.part_b[q], s);
    if (err != hipSuccess) return err;
    return sipg_apply_only<T>(N, a, c.u, c.trace, c.w, 1, c.st + (q ^ 1), c.part_a, s);
}

template <typename T>
hipError_t cg_publish(const SipgArgs<T>& a, const CgLoop<T>& c, hipStream_t s) {
    hipLaunchKernelGGL(cg_publish_kernel, dim3(1), dim3(256), 0, s, c.st, c.st + 2, c.part_b[1], c.part_a,
                       sipg_blocks(a));
    return hipGetLastError();
}

template hipError_t cg_start<double>(int, const SipgArgs<double>&, const CgLoop<double>&, hipStream_t);
template hipError_t cg_start<float>(int, const SipgArgs<float>&, const CgLoop<float>&, hipStream_t);
template hipError_t cg_iteration<double>(int, const SipgArgs<double>&, const CgLoop<double>&, int, hipStream_t);
template hipError_t cg_iteration<float>(int, const SipgArgs<float>&, const CgLoop<float>&, int, hipStream_t);
template hipError_t cg_publish<double>(const SipgArgs<double>&, const CgLoop<double>&, hipStream_t);
template hipError_t cg_publish<float>(const SipgArgs<float>&, const CgLoop<float>&, hipStream_t);

hipError_t vec_neg_live(const VecShape& s, double* out, const double* in, hipStream_t st) {
    hipLaunchKernelGGL(neg_live_kernel, dim3(vec_blocks(s.n)), dim3(256), 0, st, out, in, s.nfull, s.n, s.krem);
    return hipGetLastError();
}
hipError_t vec_sub_live(const VecShape& s, double* r, const double* a, const double* b, hipStream_t st) {
    hipLaunchKernelGGL(sub_live_kernel, dim3(vec_blocks(s.n)), dim3(256), 0, st, r, a, b, s.nfull, s.n, s.krem);
    return hipGetLastError();
}
hipError_t vec_d2f(const VecShape& s, float* y, const double* x, hipStream_t st) {
    hipLaunchKernelGGL(d2f_kernel, dim3(vec_blocks(s.n)), dim3(256), 0, st, y, x, s.nfull, s.n, s.krem);
    return hipGetLastError();
}
hipError_t vec_add_f32(const VecShape& s, double* y, const float* x, hipStream_t st) {
    hipLaunchKernelGGL(add_f32_kernel, dim3(vec_blocks(s.n)), dim3(256), 0, st, y, x, s.n);
    return hipGetLastError();
}
hipError_t vec_zero(void* x, size_t bytes, hipStream_t st) {
    const size_t nw = bytes / 4;
    hipLaunchKernelGGL(zero_kernel, dim3(vec_blocks(nw)), dim3(256), 0, st, static_cast<unsigned*>(x), nw);
    return hipGetLastError();
}
template <typename T>
hipError_t vec_norm(const VecShape& s, const T* x, CgState* st, double* part, hipStream_t stream) {
    const int nb = vec_blocks(s.n) < 1024 ? vec_blocks(s.n) : 1024;
    hipLaunchKernelGGL((norm_kernel<T>), dim3(nb), dim3(256), 0, stream, x, s.nfull, s.n, s.krem, part);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(norm_finish_kernel, dim3(1), dim3(256), 0, stream, st, part, nb);
    return hipGetLastError();
}
template hipError_t vec_norm<double>(const VecShape&, const double*, CgState*, double*, hipStream_t);
template hipError_t vec_norm<float>(const VecShape&, const float*, CgState*, double*, hipStream_t);

}  // namespace dg
