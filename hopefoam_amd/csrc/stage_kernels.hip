// DG right-hand side + fused LSERK4 stage kernels for gfx950 (CDNA4).
//
// One wavefront owns 16 elements (one per MFMA column).  Per element e the kernel evaluates
// the reference's over-integrated weak form
//     rhs_e = Pr (rx F + ry G) + Ps (sx F + sy G) - sum_f (Js_f / J) Lift_f F*_f
// (Pr = M^-1 Drq^T W, Ps = M^-1 Dsq^T W, Lift = M^-1 E Vf^T W_f folded on the host), i.e.
// HopeFOAM-0.1/src/DG/DG/convectionSchemes/defaultConvectionScheme/defaultConvectionScheme.C:47-129
// followed by the block mass solve of src/DG/dgMesh/dgMesh.C:129-172, in three MFMA stages:
//   1. cubature interpolation  qc[Nq x 16] = Vq[Nq x Np] . q[Np x 16]     (one GEMM per variable)
//   2. pointwise flux at the cubature points, in registers: the 16x16 accumulator layout puts
//      all variables of one (point, element) pair in one lane,
//   3. projection   rhs[Np x 16] += P[Np x Nq] . F[Nq x 16]  with the point index of each
//      K step permuted so that step (tile t, register r) consumes exactly the accumulator
//      register r of stage 1 -- the B operand never leaves the VGPRs,
// then the face fluxes (one (face, Gauss point) item per lane and K step, again matching the
// B-operand layout) and the lift GEMM, and finally either the RHS store or the fused
// low-storage RK update  res = a res + dt rhs;  q_out = q_in + b res  in FP64.
//
// MFMA 16x16x4 operand layout: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane
// j + 16 rg, register r with i = mfma_row<T>(rg, r) (dg_common.hpp; f32 and f64 differ).  Operator
// fragments are pre-arranged on the host in exactly the per-lane order, so every A load is a
// contiguous 64-lane vector load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "dg_common.hpp"
#include "lds_stage.hpp"
#include "lim_moments.hpp"
#include "stage_kernels.hpp"

namespace dg {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Vec4;
template <>
struct Vec4<double> { typedef d4_t type; };
template <>
struct Vec4<float> { typedef f4_t type; };

// Row of the 16x16 accumulator held in register r of a lane of row group rg (= lane / 16).
template <typename T>
struct RowMap {
    __host__ __device__ static constexpr int row(int rg, int r) { return dg::mfma_row<T>(rg, r); }
};

__device__ __forceinline__ d4_t mfma4(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma4(float a, float b, f4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Reciprocal and square root of the flux arithmetic.  FP64: the square root correctly rounded; the reciprocal
// v_rcp_f64 refined by two Newton steps r += r (1 - x r) in fused multiply-adds (each step squares the relative
// error, so the result is within an ulp of 1 / x), five instructions instead of the eleven of an IEEE division
// (the densities it is taken of are positive and normal).  FP32: the hardware's v_rcp_f32 / v_sqrt_f32 (about
// 1 ulp), far below the FP32 and bf16x3 paths' contraction round-off, in one instruction instead of the ten of an
// IEEE division or square root.
__device__ __forceinline__ double inv(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ float inv(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// ---------------------------------------------------------------- bf16x3 split products
// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi), both rounded to nearest even, so the pair carries about
// 16 significant bits.  A product a b is formed on the bf16 matrix cores as a_lo b_hi + a_hi b_lo + a_hi b_hi
// with FP32 accumulation; the dropped a_lo b_lo term is about 2^-16 of the product.
// v_mfma_f32_16x16x16_bf16: A[i][k] in lane i + 16 (k / 4), element k % 4; B[k][j] in lane j + 16 (k / 4),
// element k % 4; D as for v_mfma_f32_16x16x4f32.  One K16 step carries four K4 steps of the FP32 kernel, one per
// element, so the B fragments and accumulator rows of the FP32 kernel are reused without moving data.
typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
typedef unsigned u2_t __attribute__((ext_vector_type(2)));

struct Split4 {
    s4_t h, l;
};

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f2_t{a, b}, bf2_t));
}
__device__ __forceinline__ Split4 split4(float x0, float x1, float x2, float x3) {
    const unsigned h01 = pack_bf16(x0, x1), h23 = pack_bf16(x2, x3);
    const unsigned l01 = pack_bf16(x0 - __uint_as_float(h01 << 16), x1 - __uint_as_float(h01 & 0xffff0000u));
    const unsigned l23 = pack_bf16(x2 - __uint_as_float(h23 << 16), x3 - __uint_as_float(h23 & 0xffff0000u));
    return Split4{__builtin_bit_cast(s4_t, u2_t{h01, h23}), __builtin_bit_cast(s4_t, u2_t{l01, l23})};
}
// c += A B; A: a lane's operator fragment {hi of elements 0,1 | 2,3, lo of elements 0,1 | 2,3}
__device__ __forceinline__ f4_t mfma3(const uint4 A, const Split4& B, f4_t c) {
    const s4_t ah = __builtin_bit_cast(s4_t, u2_t{A.x, A.y}), al = __builtin_bit_cast(s4_t, u2_t{A.z, A.w});
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, B.h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, B.l, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, B.h, c, 0, 0, 0);
}

// ---------------------------------------------------------------- physics
template <int PHYS, typename T>
struct VolumeFlux;

template <typename T>
struct VolumeFlux<PHYS_ADVECT, T> {
    static constexpr int NV = 1;
    __device__ static void eval(const T* q, T ax, T ay, T, T rx, T ry, T sx, T sy, T* fr, T* fs) {
        const T F = ax * q[0], G = ay * q[0];
        fr[0] = rx * F + ry * G;
        fs[0] = sx * F + sy * G;
    }
};

// dgEulerFoam.C:81-90: gther_U = rhoU / rho, gther_p = (gamma-1)(E - rho |U|^2 / 2); volume
// terms div(U rho), div(U rhoU) + grad(p), div(U E) + div(U p).  Formed directly in the reference directions:
// with the contravariant momenta m_r = rx rhou + ry rhov, m_s = sx rhou + sy rhov and U_r = m_r / rho,
//   rx F + ry G = (m_r, rhou U_r + p rx, rhov U_r + p ry, (E + p) U_r),  likewise for s,
// the same values as rx F + ry G of the Cartesian fluxes in about half the operations.
template <typename T>
struct VolumeFlux<PHYS_EULER, T> {
    static constexpr int NV = 4;
    __device__ static void eval(const T* q, T, T, T gamma, T rx, T ry, T sx, T sy, T* fr, T* fs) {
        const T ru = q[1], rv = q[2], E = q[3];
        const T irho = inv(q[0]);
        const T p = (gamma - T(1)) * (E - T(0.5) * ((ru * ru + rv * rv) * irho));
        const T mr = rx * ru + ry * rv, ms = sx * ru + sy * rv;
        const T Ur = mr * irho, Us = ms * irho, Ep = E + p;
        fr[0] = mr;
        fr[1] = ru * Ur + p * rx;
        fr[2] = rv * Ur + p * ry;
        fr[3] = Ep * Ur;
        fs[0] = ms;
        fs[1] = ru * Us + p * sx;
        fs[2] = rv * Us + p * sy;
        fs[3] = Ep * Us;
    }
};

// Roe flux in the face-normal frame without entropy fix (RoeFlux.C:109-181); M = own side, P = other.
template <typename T>
__device__ __forceinline__ void roe_flux(T g, T nx, T ny, const T* M, const T* P, T* F) {
    const T QP2 = nx * P[1] + ny * P[2], QM2 = nx * M[1] + ny * M[2];
    const T QP3 = nx * P[2] - ny * P[1], QM3 = nx * M[2] - ny * M[1];
    const T rhoP = P[0], rhoM = M[0];
    const T uM = QM2 / rhoM, uP = QP2 / rhoP, vM = QM3 / rhoM, vP = QP3 / rhoP;
    const T pM = (g - T(1)) * (M[3] - T(0.5) * (QM2 * uM + QM3 * vM));
    const T pP = (g - T(1)) * (P[3] - T(0.5) * (QP2 * uP + QP3 * vP));
    const T HM = (M[3] + pM) / rhoM, HP = (P[3] + pP) / rhoP;
    T fr = (QM2 + QP2) / T(2);
    T fu = (QM2 * uM + pM + QP2 * uP + pP) / T(2);
    T fv = (QM3 * uM + QP3 * uP) / T(2);
    T fe = (uM * (M[3] + pM) + uP * (P[3] + pP)) / T(2);
    const T rMs = sqrt(rhoM), rPs = sqrt(rhoP);
    const T rhob = rMs * rPs;
    const T u = (rMs * uM + rPs * uP) / (rMs + rPs);
    const T v = (rMs * vM + rPs * vP) / (rMs + rPs);
    const T H = (rMs * HM + rPs * HP) / (rMs + rPs);
    const T c2 = (g - T(1)) * (H - T(0.5) * (u * u + v * v));
    const T c = sqrt(fabs(c2));
    const T dw1 = (T(-0.5) * rhob * (uP - uM) / c + T(0.5) * (pP - pM) / c2) * fabs(u - c);
    const T dw2 = ((rhoP - rhoM) - (pP - pM) / c2) * fabs(u);
    const T dw3 = (rhob * (vP - vM)) * fabs(u);
    const T dw4 = (T(0.5) * rhob * (uP - uM) / c + T(0.5) * (pP - pM) / c2) * fabs(u + c);
    fr -= (dw1 + dw2 + dw4) / T(2);
    fu -= (dw1 * (u - c) + dw2 * u + dw4 * (u + c)) / T(2);
    fv -= (dw1 * v + dw2 * v + dw3 + dw4 * v) / T(2);
    fe -= (dw1 * (H - u * c) + dw2 * (u * u + v * v) / T(2) + dw3 * v + dw4 * (H + u * c)) / T(2);
    F[0] = fr;
    F[1] = nx * fu - ny * fv;
    F[2] = ny * fu + nx * fv;
    F[3] = fe;
}

// local Lax-Friedrichs (Rusanov) flux with pointwise |u.n| + c
template <typename T>
__device__ __forceinline__ void llf_flux(T g, T nx, T ny, const T* M, const T* P, T* F) {
    const T iM = inv(M[0]), iP = inv(P[0]);
    const T uM = M[1] * iM, vM = M[2] * iM, uP = P[1] * iP, vP = P[2] * iP;
    const T pM = (g - T(1)) * (M[3] - T(0.5) * M[0] * (uM * uM + vM * vM));
    const T pP = (g - T(1)) * (P[3] - T(0.5) * P[0] * (uP * uP + vP * vP));
    const T unM = uM * nx + vM * ny, unP = uP * nx + vP * ny;
    const T lM = fabs(unM) + fsqrt(fabs(g * pM * iM)), lP = fabs(unP) + fsqrt(fabs(g * pP * iP));
    const T lam = lM > lP ? lM : lP;
    const T FM[4] = {M[0] * unM, M[1] * unM + pM * nx, M[2] * unM + pM * ny, (M[3] + pM) * unM};
    const T FP[4] = {P[0] * unP, P[1] * unP + pP * nx, P[2] * unP + pP * ny, (P[3] + pP) * unP};
#pragma unroll
    for (int v = 0; v < 4; ++v) F[v] = T(0.5) * (FM[v] + FP[v]) + T(0.5) * lam * (M[v] - P[v]);
}

// ---------------------------------------------------------------- stage kernel
// Orders a wave's LDS writes before its later LDS reads by other lanes of the same wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Persistent blocks, one per CU: 8 waves (2 per SIMD) for FP32 up to N = 6, else 4 waves with the
// full 512-register budget each (the bf16x3 kernel holds its split operands and shifted fluxes in registers:
// at two waves per SIMD it spills, measured 14 % slower at N = 6).  When they fit, the operator fragments are
// staged into LDS once per block and shared by its waves.
template <int N, typename T, bool SPLIT = false>
constexpr int waves_per_block() { return (N <= 6 && sizeof(T) == 4 && !SPLIT) ? 8 : 4; }

template <int N, int NV, typename T, bool SPLIT>
struct StageLds {
    using D = Dims<N>;
    static constexpr int NFI = 4 * D::SF;
    static constexpr int TR_G = NV * 16 + 16;              // padded stride of one Gauss point (bank spread)
    static constexpr size_t trace_wave = size_t(NFI) * TR_G;   // one face: [g][v][col]
    static constexpr int WPB = waves_per_block<N, T, SPLIT>();
    static constexpr size_t trace_elems = size_t(WPB) * trace_wave;
    static constexpr size_t budget = 150 * 1024 / sizeof(T);
    // fragment sizes in T-sized units (the bf16x3 fragments count 32-bit words)
    static constexpr size_t nP = SPLIT ? D::nPS : D::nPF;
    static constexpr size_t nV = SPLIT ? D::nVqS : D::nVqF;
    static constexpr size_t nL = SPLIT ? D::nLS : D::nLF;
    // a per-wave buffer that receives the next element group's q_in (FP64, field layout) by LDS DMA while the
    // wave works on the current group; used where it fits in the 160 KB of LDS beside every operator fragment
    static constexpr size_t qpre_bytes = size_t(NV) * D::Np * 16 * sizeof(double);
    static constexpr size_t qpre_wave = qpre_bytes / sizeof(T);
    // the face interpolation matrix Vf [Nfq][Nfp], read by every lane in the face phase (uniform addresses)
    static constexpr size_t nvf = size_t(D::Nfq) * D::Nfp;
    static constexpr bool prefetch = (trace_elems + WPB * qpre_wave + nP + nV + nL + nvf) * sizeof(T) <= 160 * 1024;
    static constexpr size_t qpre_elems = prefetch ? WPB * qpre_wave : 0;
    // operator fragments staged into LDS in order of reuse, as far as the budget allows (all of them when the
    // prefetch buffer is in use)
    static constexpr bool pf_lds = prefetch || trace_elems + nP <= budget;
    static constexpr bool vq_lds = prefetch || (pf_lds && trace_elems + nP + nV <= budget);
    static constexpr bool lf_lds = prefetch || (vq_lds && trace_elems + nP + nV + nL <= budget);
    static constexpr size_t elems = trace_elems + (pf_lds ? nP : 0) + (vq_lds ? nV : 0) + (lf_lds ? nL : 0);
};


// Other side's face nodes of variable rg for face f of element e (connectivity c): neighbour nodes in reversed
// order, or fixedValue ghost nodes.  The Nfp loads are issued with no branch around them: a face without such
// data (its flux copies the own trace) reads a valid address of its own element, and the caller drops the values
// (the return value says whether they are used), so the loads of several faces can be in flight together.
template <int N, int NV, typename OT>
__device__ __forceinline__ bool gather_face(const StageArgs& a, int c, int f, int e, int rg, OT (&o)[N + 1]) {
    constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
    constexpr size_t GS = size_t(NV) * Np * 16;
    const int v = rg < NV ? rg : 0;
    const bool nb = c >= 0;
    const bool fx = !nb && -c - 1 == BC_FIXED && a.bcq != nullptr;
    const int en = nb ? c >> 2 : e, fn = nb ? c & 3 : 0;
    const double* qn = a.q_in + size_t(en >> 4) * GS + size_t(v) * Np * 16 + (en & 15);
    const double* bq = fx ? a.bcq + ((size_t(v) * a.K + e) * 3 + f) * Nfp : qn;
#pragma unroll
    for (int k = 0; k < Nfp; ++k) {
        const int node = fn == 0 ? face_node<N>(0, N - k) : (fn == 1 ? face_node<N>(1, N - k) : face_node<N>(2, N - k));
        const double* src = fx ? bq + k : qn + node * 16;   // one load from the selected address
        o[k] = OT(*src);
    }
    return nb || fx;
}

// SPLIT: bf16x3 split products for the three contractions (T = float for everything else)
template <int N, int PHYS, typename T, bool SPLIT, bool MOM>
__global__ __launch_bounds__((64 * waves_per_block<N, T, SPLIT>())) void stage_kernel(StageArgs a) {
    using D = Dims<N>;
    using T4 = typename Vec4<T>::type;
    using VF = VolumeFlux<PHYS, T>;
    static_assert(!SPLIT || sizeof(T) == 4, "bf16x3 products accumulate in FP32");
    constexpr int NV = VF::NV;
    constexpr int Np = D::Np, Nq = D::Nq, Nfp = D::Nfp, Nfq = D::Nfq, SF = D::SF;
    constexpr size_t GS = size_t(NV) * Np * 16;   // doubles per element group in the field layout
    using L = StageLds<N, NV, T, SPLIT>;
    constexpr int TR_G = L::TR_G;

    __shared__ __attribute__((aligned(16))) T lds[L::elems];
    // the prefetch buffers are an LDS variable of their own, so the compiler can tell them from the fragments
    __shared__ __attribute__((aligned(16))) double qpre[L::prefetch ? L::qpre_elems * sizeof(T) / 8 : 2];
    __shared__ T vfl[L::nvf];
    for (int i = threadIdx.x; i < int(L::nvf); i += blockDim.x) vfl[i] = static_cast<const T*>(a.Vf)[i];
    // bf16x3 with a moments output: the moment operator (lim_moments.hpp), read by every lane in the update
    // MOM (bf16x3 only): the launch also forms the limiter's moments of the new state (its own instantiation, so
    // the launches without them keep their own register allocation and schedule)
    constexpr bool kMom = SPLIT && MOM;
    constexpr int kMomFrag = kMom ? 4 * D::MT * 64 : 1;
    __shared__ double p4l[kMomFrag];
    if (kMom)
        for (int i = threadIdx.x; i < kMomFrag; i += blockDim.x) p4l[i] = a.p4f[i];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15;   // element within the group
    const int rg = lane >> 4;    // row group of the 16x16 tiles
    const int K = a.K;
    const int groups = (K + 15) >> 4;
    const T* __restrict__ VqF = static_cast<const T*>(a.VqF);
    const T* __restrict__ PF = static_cast<const T*>(a.PF);
    const T* __restrict__ LF = static_cast<const T*>(a.LF);
    {
        T* fr = lds + L::trace_elems;
        static_assert((L::trace_elems * sizeof(T)) % 16 == 0 && L::qpre_bytes % 16 == 0,
                      "the prefetch buffers and fragments start on 16-byte boundaries");
        if (L::pf_lds) {
            stage_to_lds<L::nP * sizeof(T)>(fr, PF);
            PF = fr;
            fr += L::nP;
        }
        if (L::vq_lds) {
            stage_to_lds<L::nV * sizeof(T)>(fr, VqF);
            VqF = fr;
            fr += L::nV;
        }
        if (L::lf_lds) {
            stage_to_lds<L::nL * sizeof(T)>(fr, LF);
            LF = fr;
        }
        __syncthreads();   // the fragments and Vf are in LDS
    }
    // the same fragments seen as per-lane bf16x3 operands (SPLIT only)
    const uint4* __restrict__ VqU = reinterpret_cast<const uint4*>(VqF);
    const uint4* __restrict__ PU = reinterpret_cast<const uint4*>(PF);
    const uint4* __restrict__ LU = reinterpret_cast<const uint4*>(LF);
    const T* __restrict__ Vf = vfl;
    const T* __restrict__ geo = static_cast<const T*>(a.geo);
    const int* __restrict__ conn = a.conn;
    const T ax = T(a.ax), ay = T(a.ay), gamma = T(a.gamma);
    T* tr = lds + size_t(wave) * L::trace_wave;

    // XCD-aware schedule: block b runs on XCD b % 8; each XCD walks a contiguous range of groups
    const int xcd = blockIdx.x & 7;
    const int nbx = gridDim.x >> 3;
    const int bx = blockIdx.x >> 3;
    const int g_lo = int((long long)groups * xcd / 8);
    const int g_hi = int((long long)groups * (xcd + 1) / 8);
    constexpr int WPB = waves_per_block<N, T, SPLIT>();


    // The connectivity of the next group is fetched one iteration ahead.  With the prefetch buffer, so are its
    // q_in (LDS DMA) and volume geometric factors: the loads a group issues after its DMA are then needed only
    // by the next group, and since vector memory operations complete in order, waiting for any of them early
    // would also wait for the DMA.
    const int gstride = nbx * WPB;
    int grp = g_lo + bx * WPB + wave;
    int cn_next[3] = {0, 0, 0};
    T geo_next[4] = {T(0), T(0), T(0), T(0)};
    if (grp < g_hi) {
        const int e0 = grp * 16 + col < K ? grp * 16 + col : K - 1;
#pragma unroll
        for (int f = 0; f < 3; ++f) cn_next[f] = conn[f * K + e0];
        if constexpr (L::prefetch) {
#pragma unroll
            for (int i = 0; i < 4; ++i) geo_next[i] = geo[(G_RX + i) * K + e0];
            dma_to_lds<L::qpre_bytes>(a.q_in + size_t(grp) * GS, qpre + wave * (L::qpre_bytes / 8), lane);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the first group's q_in has landed
        }
    }
    for (; grp < g_hi; grp += gstride) {
        const int e_raw = grp * 16 + col;
        const int e = e_raw < K ? e_raw : K - 1;
        // q_in of this group: the prefetch buffer (its DMA was waited for before the previous group's update)
        // or HBM
        const double* __restrict__ qg = L::prefetch ? qpre + wave * (L::qpre_bytes / 8) : a.q_in + size_t(grp) * GS;
        int cn[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) cn[f] = cn_next[f];
        const bool more = grp + gstride < g_hi;
        const int en = more ? ((grp + gstride) * 16 + col < K ? (grp + gstride) * 16 + col : K - 1) : 0;

        T rx, ry, sx, sy;
        if constexpr (L::prefetch) {
            rx = geo_next[0];
            ry = geo_next[1];
            sx = geo_next[2];
            sy = geo_next[3];
        } else {
            if (more) {
#pragma unroll
                for (int f = 0; f < 3; ++f) cn_next[f] = conn[f * K + en];
            }
            rx = geo[G_RX * K + e];
            ry = geo[G_RY * K + e];
            sx = geo[G_SX * K + e];
            sy = geo[G_SY * K + e];
        }

        // B fragments of the interpolation GEMM (coalesced loads); they are also this lane's rows of q_in
        T qb[NV][D::KS1];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int s = 0; s < D::KS1; ++s) {
                const int node = gemm1_node<T>(s, rg);
                qb[v][s] = node < Np ? T(qg[(v * Np + node) * 16 + col]) : T(0);
            }

        // bf16x3: the contractions see the state and the fluxes relative to the element's constant state q0 (its
        // node 0).  This is exact: interpolation reproduces constants (qc = q0 + Vq (q - q0)), and the constant
        // flux F(q0) contributes nothing to the right-hand side (its volume term Pr fr0 + Ps fs0 equals its face
        // term sum_f (Js_f / J) Lift_f F(q0).n_f: the divergence theorem, which the quadratures integrate
        // exactly).  The split products then round only the variation inside the element.
        T q0[NV], fr0[NV], fs0[NV];
        if constexpr (SPLIT) {
#pragma unroll
            for (int v = 0; v < NV; ++v) q0[v] = T(qg[(v * Np) * 16 + col]);
            VF::eval(q0, ax, ay, gamma, rx, ry, sx, sy, fr0, fs0);
        }
        // the other side's face nodes and the face geometry (the face phase below says when each kernel asks)
        constexpr bool kWide = WPB == 4;
        using OT = typename std::conditional<kWide, double, T>::type;
        OT oth[3][Nfp];
        bool keep[3];
        T fnx[3], fny[3], ffs[3];
        if (SPLIT && (NV == 4 || rg < NV)) {
#pragma unroll
            for (int f = 0; f < 3; ++f) keep[f] = gather_face<N, NV>(a, cn[f], f, e, rg, oth[f]);
        }
        if constexpr (SPLIT) {
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                fnx[f] = geo[(G_NX0 + 3 * f) * K + e];
                fny[f] = geo[(G_NY0 + 3 * f) * K + e];
                ffs[f] = geo[(G_FS0 + 3 * f) * K + e];
            }
        }
        if constexpr (L::prefetch) {
            // every lane's reads of the buffer have returned before the DMA of the next group overwrites it
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (more) {
                dma_to_lds<L::qpre_bytes>(a.q_in + size_t(grp + gstride) * GS, qpre + wave * (L::qpre_bytes / 8),
                                          lane);
#pragma unroll
                for (int f = 0; f < 3; ++f) cn_next[f] = conn[f * K + en];
#pragma unroll
                for (int i = 0; i < 4; ++i) geo_next[i] = geo[(G_RX + i) * K + en];
            }
        }
        // bf16x3: the B fragments (minus q0) split into hi and lo, four K4 steps per K16 step
        Split4 qs[NV][SPLIT ? D::KS16 : 1];
        if constexpr (SPLIT) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int S = 0; S < D::KS16; ++S) {
                    T x[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int s1 = 4 * S + e < D::KS1 ? 4 * S + e : 0;
                        x[e] = (4 * S + e < D::KS1 && gemm1_node<T>(s1, rg) < Np) ? qb[v][s1] - q0[v] : T(0);
                    }
                    qs[v][S] = split4(x[0], x[1], x[2], x[3]);
                }
        }

        T4 acc[NV][D::MT];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int m = 0; m < D::MT; ++m) acc[v][m] = T4{0, 0, 0, 0};

        // ---- volume: interpolate to the cubature points, pointwise flux, project.  Software-pipelined:
        // the interpolation MFMAs of tile t+1 are issued before the fluxes of tile t, so the matrix core
        // and the vector ALU work at the same time.
        T4 qc[NV], qn[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) qc[v] = T4{0, 0, 0, 0};
        if constexpr (SPLIT) {
#pragma unroll
            for (int S = 0; S < D::KS16; ++S) {
                const uint4 A = VqU[S * 64 + lane];
#pragma unroll
                for (int v = 0; v < NV; ++v) qc[v] = mfma3(A, qs[v][S], qc[v]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < D::KS1; ++s) {
                const T A = VqF[s * 64 + lane];
#pragma unroll
                for (int v = 0; v < NV; ++v) qc[v] = mfma4(A, qb[v][s], qc[v]);
            }
        }
        // unrolled in full for FP64 up to N = 6 (at N = 6 its stage measured 7 % shorter; beyond, it spills); the
        // bf16x3 kernel spills unrolled in full and gained nothing unrolled by two, the two-wave FP32 kernel spills
        constexpr int kTileUnroll = (sizeof(T) == 8 && N <= 6) ? D::QT : 1;
#pragma unroll kTileUnroll
        for (int t = 0; t < D::QT; ++t) {
            const int tn = t + 1 < D::QT ? t + 1 : D::QT + D::FT - 1;   // last step: a face tile (harmless)
#pragma unroll
            for (int v = 0; v < NV; ++v) qn[v] = T4{0, 0, 0, 0};
            if constexpr (SPLIT) {
#pragma unroll
                for (int S = 0; S < D::KS16; ++S) {
                    const uint4 A = VqU[(tn * D::KS16 + S) * 64 + lane];
#pragma unroll
                    for (int v = 0; v < NV; ++v) qn[v] = mfma3(A, qs[v][S], qn[v]);
                }
            } else {
#pragma unroll
                for (int s = 0; s < D::KS1; ++s) {
                    const T A = VqF[(tn * D::KS1 + s) * 64 + lane];
#pragma unroll
                    for (int v = 0; v < NV; ++v) qn[v] = mfma4(A, qb[v][s], qn[v]);
                }
            }
            T frr[NV][4], fsr[NV][4];   // SPLIT: the tile's fluxes, one K16 step of the projection
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int point = 16 * t + RowMap<T>::row(rg, r);
                T qp[NV], fr[NV], fs[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) qp[v] = SPLIT ? qc[v][r] + q0[v] : qc[v][r];
                if (point < Nq) {
                    VF::eval(qp, ax, ay, gamma, rx, ry, sx, sy, fr, fs);
                    if constexpr (SPLIT) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            fr[v] -= fr0[v];
                            fs[v] -= fs0[v];
                        }
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) fr[v] = fs[v] = T(0);
                }
                if constexpr (SPLIT) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        frr[v][r] = fr[v];
                        fsr[v][r] = fs[v];
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < D::MT; ++m) {
                        const int base = (((t * 4 + r) * D::MT + m) * 2) * 64 + lane;
                        const T Ar = PF[base], As = PF[base + 64];
#pragma unroll
                        for (int v = 0; v < NV; ++v) acc[v][m] = mfma4(Ar, fr[v], acc[v][m]);
#pragma unroll
                        for (int v = 0; v < NV; ++v) acc[v][m] = mfma4(As, fs[v], acc[v][m]);
                    }
                }
            }
            if constexpr (SPLIT) {
                Split4 br[NV], bs[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    br[v] = split4(frr[v][0], frr[v][1], frr[v][2], frr[v][3]);
                    bs[v] = split4(fsr[v][0], fsr[v][1], fsr[v][2], fsr[v][3]);
                }
#pragma unroll
                for (int m = 0; m < D::MT; ++m) {
                    const uint4 Ar = PU[((t * D::MT + m) * 2) * 64 + lane], As = PU[((t * D::MT + m) * 2 + 1) * 64 + lane];
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v][m] = mfma3(Ar, br[v], acc[v][m]);
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v][m] = mfma3(As, bs[v], acc[v][m]);
                }
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) qc[v] = qn[v];
        }

        const size_t gbase = size_t(grp) * GS;
        const double* __restrict__ qin_g = a.q_in + gbase;
        double* __restrict__ out_g = a.q_out + gbase;
        double* __restrict__ res_g = a.res + gbase;
        // bf16x3 update (the full register file of one wave per SIMD; up to 8 tile-variables, 128 VGPRs): the
        // residual and q_in of every output tile are requested here, before the faces, and arrive while the face
        // work runs.  (The FP32 kernel would spill them and reads them in the update; FP64 holds its own.)
        constexpr bool kHoist = sizeof(T) == 4 && WPB == 4 && D::MT * NV <= 8;
        constexpr int MH = kHoist ? D::MT : 1;
        double rmh[MH][4][NV], qmh[MH][4][NV];
        if (kHoist && a.mode != 0) {
#pragma unroll
            for (int m = 0; m < MH; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = 16 * m + RowMap<T>::row(rg, r);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        rmh[m][r][v] = j < Np ? res_g[(v * Np + j) * 16 + col] : 0.0;
                        qmh[m][r][v] = j < Np ? qin_g[(v * Np + j) * 16 + col] : 0.0;
                    }
                }
        }

        // ---- faces.  The other side's face nodes of all three faces (lane (col, rg): variable rg) are requested
        // together, so their latency is paid once and behind other work rather than once per face: bf16x3 requests
        // them at the group start (above) and holds them through the volume phase; FP64, short of registers for
        // that (it would spill), before the own-trace MFMAs below.  With one wave per SIMD (FP64, bf16x3) the face
        // geometry comes with them and the nodes stay FP64 until used; the two-wave FP32 kernel, short of
        // registers, converts the nodes as they arrive and loads each face's geometry when it gets there.
        if (kWide && !SPLIT && (NV == 4 || rg < NV)) {
#pragma unroll
            for (int f = 0; f < 3; ++f) keep[f] = gather_face<N, NV>(a, cn[f], f, e, rg, oth[f]);
        }
        if constexpr (kWide && !SPLIT) {
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                fnx[f] = geo[(G_NX0 + 3 * f) * K + e];
                fny[f] = geo[(G_NY0 + 3 * f) * K + e];
                ffs[f] = geo[(G_FS0 + 3 * f) * K + e];
            }
        }
        // Own traces at the face Gauss points come out of the interpolation GEMM (face tiles:
        // register r of tile tf holds lift K step 4 tf + r, Gauss point 4 (step % SF) + rg).
        // (FP64 forms them later, in the face loop, from the lanes' own nodal values shuffled within each element's
        // four lanes and Vf: a quarter of the flops of the face tiles, on the VALU, which takes as long per f64 flop
        // as the matrix pipe, the kernel's busiest unit)
        constexpr bool kOwnValu = sizeof(T) == 8;
        T4 own[kOwnValu ? 1 : D::FT][NV];
#pragma unroll
        for (int tf = 0; tf < (kOwnValu ? 0 : D::FT); ++tf) {
#pragma unroll
            for (int v = 0; v < NV; ++v) own[tf][v] = T4{0, 0, 0, 0};
            if constexpr (SPLIT) {
#pragma unroll
                for (int S = 0; S < D::KS16; ++S) {
                    const uint4 A = VqU[((D::QT + tf) * D::KS16 + S) * 64 + lane];
#pragma unroll
                    for (int v = 0; v < NV; ++v) own[tf][v] = mfma3(A, qs[v][S], own[tf][v]);
                }
            } else {
#pragma unroll
                for (int s = 0; s < D::KS1; ++s) {
                    const T A = VqF[((D::QT + tf) * D::KS1 + s) * 64 + lane];
#pragma unroll
                    for (int v = 0; v < NV; ++v) own[tf][v] = mfma4(A, qb[v][s], own[tf][v]);
                }
            }
        }
        if (!kWide && (NV == 4 || rg < NV)) {   // FP32: after the own-trace MFMAs, whose accumulators are live
#pragma unroll
            for (int f = 0; f < 3; ++f) keep[f] = gather_face<N, NV>(a, cn[f], f, e, rg, oth[f]);
        }
        // FP64: residual registers for the stage update, requested before the face work that hides them
        // (the FP32 kernel has no register room for them and reads them in the update)
        constexpr bool kPrefetchRes = sizeof(T) == 8;
        double rold[NV][kPrefetchRes ? D::KS1 : 1];
        if (kPrefetchRes && a.mode != 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int s = 0; s < D::KS1; ++s) {
                    const int node = gemm1_node<T>(s, rg);
                    rold[v][s] = node < Np ? a.res[size_t(grp) * GS + (v * Np + node) * 16 + col] : 0.0;
                }
        }
        T F0[NV], G0[NV];   // SPLIT: physical flux of the constant state, for its normal flux on the faces
        if constexpr (SPLIT) VF::eval(q0, ax, ay, gamma, T(1), T(0), T(0), T(1), F0, G0);
        T bq[NV][4];   // SPLIT: face fluxes of four consecutive lift K steps (one K16 step)
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int e = 0; e < 4; ++e) bq[v][e] = T(0);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const int fg = f;
            if constexpr (!kWide) {
                fnx[f] = geo[(G_NX0 + 3 * f) * K + e];
                fny[f] = geo[(G_NY0 + 3 * f) * K + e];
                ffs[f] = geo[(G_FS0 + 3 * f) * K + e];
            }
            if (rg < NV) {
#pragma unroll
                for (int g = 0; g < Nfq; ++g) {
                    T sm = T(0);
#pragma unroll
                    for (int k = 0; k < Nfp; ++k) sm += Vf[g * Nfp + k] * (keep[fg] ? T(oth[fg][k]) : T(0));
                    tr[g * TR_G + rg * 16 + col] = sm;
                }
            }
            wave_sync();
            const T nx = fnx[fg], ny = fny[fg], fsc = ffs[fg];
            const bool copy_own = cn[f] < 0 && -cn[f] - 1 != BC_FIXED;
            // FP64: the element's own nodes on face f, from the lanes of its row groups (node n sits in row group
            // n % 4 at K step n / 4); every lane takes part in the shuffles
            T fnode[kOwnValu ? NV : 1][kOwnValu ? Nfp : 1];
            if constexpr (kOwnValu) {
#pragma unroll
                for (int k = 0; k < Nfp; ++k) {
                    const int node = face_node<N>(f, k);
#pragma unroll
                    for (int v = 0; v < NV; ++v) fnode[v][k] = __shfl(qb[v][node >> 2], col + 16 * (node & 3), 64);
                }
            }
#pragma unroll
            for (int j = 0; j < SF; ++j) {
                const int st = f * SF + j;
                const int g = 4 * j + rg;
                T bf[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) bf[v] = T(0);
                if (g < Nfq) {
                    T qM[NV], qP[NV], Fs[NV];
                    if constexpr (kOwnValu) {   // the own trace at Gauss point g
#pragma unroll
                        for (int v = 0; v < NV; ++v) qM[v] = T(0);
#pragma unroll
                        for (int k = 0; k < Nfp; ++k) {
                            const T w = Vf[g * Nfp + k];
#pragma unroll
                            for (int v = 0; v < NV; ++v) qM[v] += w * fnode[v][k];
                        }
                    } else {
#pragma unroll
                        for (int v = 0; v < NV; ++v)
                            qM[v] = SPLIT ? own[st / 4][v][st % 4] + q0[v] : own[st / 4][v][st % 4];
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        qP[v] = copy_own ? qM[v] : tr[g * TR_G + v * 16 + col];
                    }
                    if (PHYS == PHYS_EULER && cn[f] < 0 && -cn[f] - 1 == BC_WALL) {   // mirror the momentum
                        const T mn = qP[1] * nx + qP[2] * ny;
                        qP[1] -= T(2) * mn * nx;
                        qP[2] -= T(2) * mn * ny;
                    }
                    if (PHYS == PHYS_ADVECT) {   // upwind, or the central "average" flux without the |vn| term
                        const T vn = nx * ax + ny * ay;
                        const T diss = a.flux == FLUX_AVERAGE ? T(0) : fabs(vn);
                        Fs[0] = (vn * qM[0] + vn * qP[0]) * T(0.5) + diss * (qM[0] - qP[0]) * T(0.5);
                    } else if (a.flux == FLUX_ROE) {
                        roe_flux<T>(gamma, nx, ny, qM, qP, Fs);
                    } else {
                        llf_flux<T>(gamma, nx, ny, qM, qP, Fs);
                    }
                    if constexpr (SPLIT) {   // minus the constant state's normal flux
#pragma unroll
                        for (int v = 0; v < NV; ++v) Fs[v] -= F0[v] * nx + G0[v] * ny;
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) bf[v] = -fsc * Fs[v];
                }
                if constexpr (SPLIT) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) bq[v][st & 3] = bf[v];
                    if ((st & 3) == 3 || st == D::KSF - 1) {
                        Split4 b[NV];
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                if (e > (st & 3)) bq[v][e] = T(0);
                            b[v] = split4(bq[v][0], bq[v][1], bq[v][2], bq[v][3]);
                        }
#pragma unroll
                        for (int m = 0; m < D::MT; ++m) {
                            const uint4 A = LU[((st >> 2) * D::MT + m) * 64 + lane];
#pragma unroll
                            for (int v = 0; v < NV; ++v) acc[v][m] = mfma3(A, b[v], acc[v][m]);
                        }
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < D::MT; ++m) {
                        const T A = LF[(st * D::MT + m) * 64 + lane];
#pragma unroll
                        for (int v = 0; v < NV; ++v) acc[v][m] = mfma4(A, bf[v], acc[v][m]);
                    }
                }
            }
            wave_sync();
        }

        if constexpr (L::prefetch) {
            // the DMA of the next group's q_in (issued at the start of this group) has landed before the stores
            // below; everything else outstanding here is already consumed, so this wait costs nothing, and the
            // buffer is complete when the next iteration reads it
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }

        // ---- output in the field layout: rows j = 16m + row(rg, r) of column col.  For FP64, node j is
        // gemm1_node(4m + r, rg), whose q_in this lane holds in qb and whose residual it prefetched.
        // bf16x3 with a moments output: K step s = 4 m + r of the f64 moment product takes row group rg's row
        // 16 m + 4 rg + r = node lim_mnode(s, rg), the new state this lane has just formed, in the moments kernel's
        // order and with its fragments (so the two give the same bits); rows beyond Np enter as zeros.  (Only the
        // one-wave-per-SIMD bf16x3 kernel has the registers for the accumulators; FP32 and FP64 leave the moments to
        // the limiter's own kernel.)
        const bool want_mom = kMom;   // the launcher takes this instantiation for a moments output and an update
        double mpart[kMom ? NV : 1][4];   // the moments' per-lane partial sums (lim_moments.hpp)
#pragma unroll
        for (int v = 0; v < (kMom ? NV : 1); ++v)
#pragma unroll
            for (int k = 0; k < 4; ++k) mpart[v][k] = 0.0;
#pragma unroll
        for (int m = 0; m < D::MT; ++m) {
            double rm[4][NV], qm[4][NV];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s = 4 * m + r;
                const int j = 16 * m + RowMap<T>::row(rg, r);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    if (kPrefetchRes) {
                        rm[r][v] = s < D::KS1 ? rold[v][s < D::KS1 ? s : 0] : 0.0;
                        qm[r][v] = s < D::KS1 ? double(qb[v][s < D::KS1 ? s : 0]) : 0.0;
                    } else if (kHoist) {
                        rm[r][v] = rmh[m < MH ? m : 0][r][v];
                        qm[r][v] = qmh[m < MH ? m : 0][r][v];
                    } else if (a.mode != 0) {
                        rm[r][v] = j < Np ? res_g[(v * Np + j) * 16 + col] : 0.0;
                        qm[r][v] = j < Np ? qin_g[(v * Np + j) * 16 + col] : 0.0;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = 16 * m + RowMap<T>::row(rg, r);
                double qnew[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) qnew[v] = 0.0;
                if (j < Np) {   // padding columns (e_raw >= K) keep evolving a copy of element K-1
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const int idx = (v * Np + j) * 16 + col;
                        const double rhs = double(acc[v][m][r]);
                        if (a.mode == 0) {
                            out_g[idx] = rhs;
                        } else {
                            const double rn = a.rk_a * rm[r][v] + a.dt * rhs;
                            res_g[idx] = rn;
                            qnew[v] = qm[r][v] + a.rk_b * rn;
                            out_g[idx] = qnew[v];
                        }
                    }
                }
                if (kMom && want_mom) {
#pragma unroll
                    for (int v = 0; v < (kMom ? NV : 1); ++v) lim_moment_acc(p4l, 4 * m + r, lane, qnew[v], mpart[v]);
                }
            }
        }
        if (kMom && want_mom) {   // wave-uniform: every lane takes part in the reduction
#pragma unroll
            for (int v = 0; v < (kMom ? NV : 1); ++v)
                a.mom_out[(size_t(grp) * NV + v) * 64 + lane] = lim_moment_reduce(mpart[v], lane);
        }
    }
}

// Persistent grid: as many blocks as the device keeps resident at once (a multiple of 8 so that every
// XCD gets the same share), capped by the number of element groups.
template <int N, int PHYS, typename T, bool SPLIT, bool MOM>
static hipError_t launch_inst(const StageArgs& a, int ncu, hipStream_t st) {
    constexpr int threads = 64 * waves_per_block<N, T, SPLIT>();
    static int per_cu = -1;
    if (per_cu < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, stage_kernel<N, PHYS, T, SPLIT, MOM>, threads, 0) !=
            hipSuccess)
            n = 1;
        per_cu = n < 1 ? 1 : n;
    }
    const int groups = (a.K + 15) / 16;
    const int wpb = waves_per_block<N, T, SPLIT>();
    int nb = (groups + wpb - 1) / wpb;
    if (nb > ncu * per_cu) nb = ncu * per_cu;
    nb = nb < 8 ? 8 : (nb + 7) / 8 * 8;
    if (a.block_cap > 0 && nb > a.block_cap) nb = a.block_cap < 8 ? 8 : a.block_cap / 8 * 8;
    hipLaunchKernelGGL((stage_kernel<N, PHYS, T, SPLIT, MOM>), dim3(nb), dim3(threads), 0, st, a);
    return hipGetLastError();
}

template <int N, int PHYS, typename T, bool SPLIT>
static hipError_t launch_one(const StageArgs& a, int ncu, hipStream_t st) {
    if constexpr (SPLIT) {
        if (a.mom_out != nullptr && a.mode != 0) return launch_inst<N, PHYS, T, SPLIT, true>(a, ncu, st);
    }
    return launch_inst<N, PHYS, T, SPLIT, false>(a, ncu, st);
}

template <int PHYS, typename T, bool SPLIT = false>
static hipError_t launch_n(int N, const StageArgs& a, int ncu, hipStream_t st) {
    switch (N) {
        case 1: return launch_one<1, PHYS, T, SPLIT>(a, ncu, st);
        case 2: return launch_one<2, PHYS, T, SPLIT>(a, ncu, st);
        case 3: return launch_one<3, PHYS, T, SPLIT>(a, ncu, st);
        case 4: return launch_one<4, PHYS, T, SPLIT>(a, ncu, st);
        case 5: return launch_one<5, PHYS, T, SPLIT>(a, ncu, st);
        case 6: return launch_one<6, PHYS, T, SPLIT>(a, ncu, st);
        case 7: return launch_one<7, PHYS, T, SPLIT>(a, ncu, st);
        case 8: return launch_one<8, PHYS, T, SPLIT>(a, ncu, st);
        default: return hipErrorInvalidValue;
    }
}

template <int PHYS>
static hipError_t launch_prec(int N, int prec, const StageArgs& a, int ncu, hipStream_t st) {
    switch (prec) {
        case PREC_FP64: return launch_n<PHYS, double>(N, a, ncu, st);
        case PREC_FP32: return launch_n<PHYS, float>(N, a, ncu, st);
        case PREC_BF16X3: return launch_n<PHYS, float, true>(N, a, ncu, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stage(int N, int phys, int prec, const StageArgs& a, int ncu, hipStream_t st) {
    return phys == PHYS_ADVECT ? launch_prec<PHYS_ADVECT>(N, prec, a, ncu, st)
                               : launch_prec<PHYS_EULER>(N, prec, a, ncu, st);
}

// ---------------------------------------------------------------- MFMA layout probe
// D = A(16x4) B(4x16) with A, B given row-major; checks the lane layout the kernels assume.
__global__ void mfma_probe_kernel(const double* A, const double* B, double* D64, const float* Af, const float* Bf,
                                  float* D32) {
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, k = lane >> 4;
    d4_t c = {0, 0, 0, 0};
    c = mfma4(A[i * 4 + k], B[k * 16 + i], c);
    f4_t cf = {0, 0, 0, 0};
    cf = mfma4(Af[i * 4 + k], Bf[k * 16 + i], cf);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        D64[mfma_row<double>(k, r) * 16 + i] = c[r];
        D32[mfma_row<float>(k, r) * 16 + i] = cf[r];
    }
}

// Decodes the accumulator layout: with B = 1 and A = lane id, D[i][j] = sum_k lane(A[i][k]); with A = 1 and
// B = lane id, D[i][j] = sum_k lane(B[k][j]).  out: [f64: A-probe 256 | B-probe 256 | f32: A 256 | B 256],
// each 256 block indexed lane * 4 + register.
__global__ void mfma_layout_probe_kernel(double* out) {
    const int lane = threadIdx.x & 63;
    d4_t a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
    a = mfma4(double(lane), 1.0, a);
    b = mfma4(1.0, double(lane), b);
    f4_t af = {0, 0, 0, 0}, bf = {0, 0, 0, 0};
    af = mfma4(float(lane), 1.0f, af);
    bf = mfma4(1.0f, float(lane), bf);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        out[lane * 4 + r] = a[r];
        out[256 + lane * 4 + r] = b[r];
        out[512 + lane * 4 + r] = af[r];
        out[768 + lane * 4 + r] = bf[r];
    }
}

// D = A(16x16) B(16x16) through one K16 step of bf16x3 products (D3) and of the hi parts alone (D1).  A arrives
// as the host-split per-lane fragments of the operators; B (row-major) is split here, as the kernels split
// their data.  D is written row-major.
__global__ void mfma_bf16x3_probe_kernel(const uint4* A, const float* B, float* D3, float* D1) {
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, kq = lane >> 4;
    const Split4 b = split4(B[(4 * kq) * 16 + j], B[(4 * kq + 1) * 16 + j], B[(4 * kq + 2) * 16 + j],
                            B[(4 * kq + 3) * 16 + j]);
    f4_t c3 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
    c3 = mfma3(A[lane], b, c3);
    const uint4 Ah = A[lane];
    c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4_t, u2_t{Ah.x, Ah.y}), b.h, c1, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        D3[mfma_row<float>(kq, r) * 16 + j] = c3[r];
        D1[mfma_row<float>(kq, r) * 16 + j] = c1[r];
    }
}

hipError_t launch_mfma_bf16x3_probe(const unsigned* A, const float* B, float* D3, float* D1, hipStream_t st) {
    hipLaunchKernelGGL(mfma_bf16x3_probe_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<const uint4*>(A), B, D3,
                       D1);
    return hipGetLastError();
}

// out = a - b over n values (the SSP-RK2 second stage's q^n - q^(1) term)
__global__ __launch_bounds__(256) void field_diff_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                         double* __restrict__ out, size_t n) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) out[i] = a[i] - b[i];
}

hipError_t launch_field_diff(const double* a, const double* b, double* out, size_t n, int ncu, hipStream_t st) {
    const size_t blocks = std::min<size_t>((n + 255) / 256, size_t(8) * ncu);
    hipLaunchKernelGGL(field_diff_kernel, dim3(unsigned(blocks > 0 ? blocks : 1)), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_mfma_layout_probe(double* out, hipStream_t st) {
    hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, st, out);
    return hipGetLastError();
}

hipError_t launch_mfma_probe(const double* A, const double* B, double* D64, const float* Af, const float* Bf,
                             float* D32, hipStream_t st) {
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, st, A, B, D64, Af, Bf, D32);
    return hipGetLastError();
}

}  // namespace dg
