// Matrix-free SIPG operator and PCG building blocks (kernels in sipg.hip, host driver in sipg.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dg {

// Per-element SIPG geometry, kSipgGeo values per element, stored per group of 16 elements as
// [group][slot][16]:
//   0 J, 1 c1 = rx^2+ry^2, 2 c2 = rx sx+ry sy, 3 c3 = sx^2+sy^2,
//   4+2f ar_f = nx rx + ny ry, 5+2f as_f = nx sx + ny sy     (n . grad = ar d/dr + as d/ds on face f)
//   10+f Js of the face's owner (the smaller element index; its quadrature weights serve both sides)
//   13+f penalty tau_f
constexpr int kSipgGeo = 16;

// MFMA tiling of the operator for order N (16x16x4 MFMA; one wavefront per group of 16 elements):
//   trace GEMM   T[3 TI 16 x 16] = [Vface; DrF; DsF] (items = f Nfq + q, padded to TI tiles of 16 rows)
//                                   x u[NPP x 16]                              (KS = NPP / 4 K steps)
//   apply GEMM   out[MT 16 x 16] = [Srr Srs Sss | Vface^T DrF^T DsF^T]
//                                   x [-J c1 u; -J c2 u; -J c3 u; a_v; ar c_v; as c_v]   (KA K steps)
// with Vface = face interpolation of the face nodes, DrF = Vface Dr, DsF = Vface Ds.  Nodes are padded to
// NPP and items to NIP (multiples of 4), so node k and item i keep their K-step lane (k % 4, i % 4) in
// every block of the K dimension.
struct SipgTiles {
    int Np, Nfq, NFQ3, NPP, NIP, TI, KS, MT, KA;
    __host__ __device__ constexpr explicit SipgTiles(int N)
        : Np((N + 1) * (N + 2) / 2), Nfq(N + 2), NFQ3(3 * (N + 2)), NPP(((N + 1) * (N + 2) / 2 + 3) / 4 * 4),
          NIP((3 * (N + 2) + 3) / 4 * 4), TI((3 * (N + 2) + 15) / 16), KS(((N + 1) * (N + 2) / 2 + 3) / 4),
          MT(((N + 1) * (N + 2) / 2 + 15) / 16),
          KA((3 * (((N + 1) * (N + 2) / 2 + 3) / 4 * 4) + 3 * ((3 * (N + 2) + 3) / 4 * 4)) / 4) {}
};

template <typename T>
struct SipgArgs {
    int K = 0;
    const T* geo = nullptr;        // [K/16][kSipgGeo][16]
    const int* conn = nullptr;     // [3][K]: en * 4 + fn, or -1 - boundary tag
    const T* Dr = nullptr;         // Np x Np nodal derivative matrices
    const T* Ds = nullptr;
    const T* Vf = nullptr;         // Nfq x Nfp face interpolation
    const T* gw = nullptr;         // Nfq face Gauss weights
    const T* Srr = nullptr;        // Np x Np cubature stiffness pieces Drq^T W Drq, Drq^T W Dsq + Dsq^T W Drq,
    const T* Srs = nullptr;        //   Dsq^T W Dsq
    const T* Sss = nullptr;
    const T* M = nullptr;          // Np x Np reference mass matrix
    const int* fmask = nullptr;    // 3 x Nfp
    const int* nface = nullptr;    // Np x 2: face slots f * Nfp + k of node n (-1 when fewer)
    const T* TrF = nullptr;        // trace GEMM A fragments [3 TI][KS][64 lanes]
    const T* ApF = nullptr;        // apply GEMM A fragments [MT][KA][64 lanes]
    const T* MF = nullptr;         // reference mass matrix as apply A fragments [MT][NPP / 4][64 lanes]
    T sigma = T(0);                // mass shift: the operator is A + sigma M
    int ncu = 256;                 // compute units (persistent grid size)
};

// Device-side state of one PCG solve.  The recurrence is PCG in the Chronopoulos & Gear form (one reduction per
// iteration; the same iterates as PETSc KSPCG in exact arithmetic, KSPCG's convergence test on |z|).  The kernels
// of an iteration read `done` first and do nothing once it is set, so iterations can be queued in batches and
// the host checks the state once per batch.
struct CgState {
    double beta, betaold;   // gamma = r.z of the current and the previous iterate
    double pw;              // p.A p of the current direction
    double dp, ttol, rnorm0, rtol, abstol;
    double alpha, bcoef;    // the step and the direction coefficient the next update applies
    int its, done, maxit;
    int reason;       // 1 converged, 2 beta = 0, -3 iteration limit, -4 diverged, -8 indefinite, -9 not finite
    int pad[4];
};

enum CgFinish { CG_FINISH_INIT = 0, CG_FINISH_ITER = 2, CG_FINISH_NORM = 3, CG_FINISH_AUTO = 4, CG_FINISH_NONE = 5 };

// Vector kernels see only live entries: the field layout pads K to a multiple of 16 and the
// padding of the last element group (e % 16 >= K % 16) is skipped in dot products.
struct VecShape {
    size_t n;       // entries in the padded layout
    size_t nfull;   // entries of the complete groups of 16 elements
    int krem;       // K % 16
};
VecShape vec_shape(int K, int Np);
int vec_blocks(size_t n);

// out = A u (negate 0) or -A u (negate 1); trace holds K/16 * 3 (N+2) * 2 * 16 values
template <typename T>
hipError_t sipg_apply(int N, const SipgArgs<T>& a, const T* u, T* trace, T* out, int negate, hipStream_t st);
hipError_t sipg_rhs(int N, const SipgArgs<double>& a, const double* f, const double* gbc, const double* gnb, double* b,
                    hipStream_t st);

// The vectors and state of one PCG solve (u = B r is the preconditioned residual, w = S u, ap = S p).
// Iteration i runs two kernels: the update kernel reads st[i & 1] and the partials part_b[(i & 1) ^ 1] and
// part_a, finishes the previous reduction into st[(i & 1) ^ 1], updates p, ap, x, r, forms u = B r and its face
// traces, and writes its u.u, r.u partials to part_b[i & 1]; the apply kernel forms w = S u and its u.w partials
// in part_a.  The start is the update kernel in its initial form (r = rhs - w for w = S x, state st[1] passed to
// st[0], partials to part_b[1]) followed by the apply.  st[2] is the published state the host reads.  Each
// partial buffer holds 2 * max(K/16, 2048) doubles.  binv != null selects block Jacobi (binv[bidx[e]] is element
// e's Np x Np inverse block).
template <typename T>
struct CgLoop {
    T *x, *r, *u, *w, *p, *ap, *trace;
    const T *rhs, *dinv, *binv;
    const int* bidx;
    CgState* st;
    double *part_a, *part_b[2];
    int ncu;
};
// r = rhs - w (w = S x on entry), u = B r and its traces, then w = S u; the state st[1] (its = -1) passes to st[0]
template <typename T>
hipError_t cg_start(int N, const SipgArgs<T>& a, const CgLoop<T>& c, hipStream_t s);
// Iteration i (parity i & 1) of PCG in two launches
template <typename T>
hipError_t cg_iteration(int N, const SipgArgs<T>& a, const CgLoop<T>& c, int i, hipStream_t s);
// st[2] = the state after the pending finish (what the next iteration would see), after an even number of
// iterations
template <typename T>
hipError_t cg_publish(const SipgArgs<T>& a, const CgLoop<T>& c, hipStream_t s);

hipError_t vec_neg_live(const VecShape& s, double* out, const double* in, hipStream_t st);   // -in, 0 on padding
hipError_t vec_sub_live(const VecShape& s, double* r, const double* a, const double* b, hipStream_t st);   // a - b
hipError_t vec_d2f(const VecShape& s, float* y, const double* x, hipStream_t st);           // float(x), 0 on padding
hipError_t vec_add_f32(const VecShape& s, double* y, const float* x, hipStream_t st);       // y += x
hipError_t vec_zero(void* x, size_t bytes, hipStream_t st);
template <typename T>
hipError_t vec_norm(const VecShape& s, const T* x, CgState* st, double* part, hipStream_t stream);   // st->dp = |x|

}  // namespace dg
