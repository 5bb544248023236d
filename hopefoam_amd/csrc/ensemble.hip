// Ensembles of independent LDG advection-diffusion cases (u_t + div(a u) = nu div(q), q = grad u; periodic
// triangle meshes of different sizes and orders) for gfx950.  The method is the local discontinuous
// Galerkin scheme of Cockburn & Shu (1998) with alternating fluxes (u^ from the side the switch vector
// points away from, q^ from the other, penalty tau on u jumps) in the nodal strong form; it has no
// counterpart in the reference, whose dgc::laplacian is empty.
//
// Each LSERK4 stage is two persistent launches over the whole ensemble (the auxiliary q of an element's
// neighbours must be complete before the right-hand side reads it -- a grid-wide dependency that a kernel
// boundary settles without fences or spin waits):
//   ens_kernel<0>  q = grad u - LIFT(fscale n (u- - u^))                         -> qx, qy
//   ens_kernel<1>  rhs = -div(a u - nu q) + LIFT(fscale (n.F- - n.F^)), then res = a res + dt rhs,
//                  u_next = u + b res
// Work items are (case, 64 elements, row slice): one element per lane, so the operator entries a wave
// needs are the same for all its lanes and come through the scalar data cache (uniform loads), with no LDS.
// The operators are stored transposed with rows padded to kEnsRows (zeros), so the R entries of one column
// that a lane's rows need are contiguous (wide scalar loads) and no row index needs clamping.
// The slice count P per order (rows per lane R = ceil(Np / P) <= 16) is the tile shape the autotuner picks.
// Items are split into 8 contiguous ranges of equal estimated cost, one per XCD, so that the neighbour
// traces an item reads were brought into the same XCD's L2 by the items next to it.
#include <hip/hip_runtime.h>

#include "ensemble.hpp"

namespace dg {

__host__ __device__ constexpr int ens_np(int N) { return (N + 1) * (N + 2) / 2; }
__host__ __device__ constexpr int ens_rows(int N, int P) { return (ens_np(N) + P - 1) / P; }
__host__ __device__ constexpr bool ens_split_ok(int N, int P) {
    return N >= 1 && N <= 8 && (P == 1 || P == 2 || P == 4 || P == 8) && ens_rows(N, P) <= 16 &&
           (P == 1 || ens_rows(N, P) >= 2);
}

bool ens_split_supported(int N, int P) { return ens_split_ok(N, P); }

// node of point k of face f (face f runs from vertex f to vertex f + 1; row j of the nodal set holds
// N + 1 - j nodes and starts at j (N + 1) - j (j - 1) / 2)
template <int N>
__device__ __forceinline__ int face_node(int f, int k) {
    if (f == 0) return k;
    if (f == 1) return k * (N + 1) - k * (k - 1) / 2 + (N - k);
    const int j = N - k;
    return j * (N + 1) - j * (j - 1) / 2;
}

__device__ __forceinline__ size_t nidx(int e, int n, int Np) { return (size_t(e >> 4) * Np + n) * 16 + (e & 15); }

template <int N, int P>
__device__ void ens_q_item(const EnsArgs& a, const EnsCase& cs, int blk, int p, int lane) {
    constexpr int Np = ens_np(N), Nfp = N + 1, R = ens_rows(N, P);
    const int Kp = cs.Kp;
    const int e0 = blk * kEnsLanes + lane;
    const bool live = e0 < cs.K;
    const int e = live ? e0 : cs.K - 1;
    const size_t arr = size_t(Kp) * Np;
    const double* __restrict__ u = a.data + cs.off_u + size_t(a.cur) * arr;
    double* __restrict__ qx = a.data + cs.off_u + ENS_QX * arr;
    double* __restrict__ qy = a.data + cs.off_u + ENS_QY * arr;
    const double* __restrict__ g = a.geo + cs.off_geo;
    const int* __restrict__ cn = a.conn + cs.off_conn;
    const int i0 = p * R;
    const double* __restrict__ DrT = a.ops[N] + i0;                 // DrT[j * kEnsRows + i]
    const double* __restrict__ DsT = DrT + Np * kEnsRows;
    const double* __restrict__ LfT = DsT + Np * kEnsRows;           // LIFT^T[(f Nfp + k) * kEnsRows + i]
    double ar[R], as[R];
#pragma unroll
    for (int r = 0; r < R; ++r) ar[r] = as[r] = 0.0;
#pragma unroll 2
    for (int j = 0; j < Np; ++j) {
        const double uj = u[nidx(e, j, Np)];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ar[r] += DrT[j * kEnsRows + r] * uj;
            as[r] += DsT[j * kEnsRows + r] * uj;
        }
    }
    const double rx = g[0 * Kp + e], ry = g[1 * Kp + e], sx = g[2 * Kp + e], sy = g[3 * Kp + e];
    double ox[R], oy[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        ox[r] = rx * ar[r] + sx * as[r];
        oy[r] = ry * ar[r] + sy * as[r];
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int c = cn[f * Kp + e];
        const int en = c >> 3, fn = (c >> 1) & 3;
        const bool own_side = c & 1;   // u^ = u-: no jump term
        const double nx = g[(4 + 4 * f) * Kp + e], ny = g[(5 + 4 * f) * Kp + e], fs = g[(6 + 4 * f) * Kp + e];
#pragma unroll 1
        for (int k = 0; k < Nfp; ++k) {
            const double um = u[nidx(e, face_node<N>(f, k), Np)];
            const double up = u[nidx(en, face_node<N>(fn, N - k), Np)];
            const double d = own_side ? 0.0 : fs * (um - up);
            const double gx = nx * d, gy = ny * d;
            const double* __restrict__ l = LfT + (f * Nfp + k) * kEnsRows;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                ox[r] -= l[r] * gx;
                oy[r] -= l[r] * gy;
            }
        }
    }
    if (live) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (i0 + r < Np) {
                qx[nidx(e0, i0 + r, Np)] = ox[r];
                qy[nidx(e0, i0 + r, Np)] = oy[r];
            }
    }
}

template <int N, int P>
__device__ void ens_rhs_item(const EnsArgs& a, const EnsCase& cs, int blk, int p, int lane) {
    constexpr int Np = ens_np(N), Nfp = N + 1, R = ens_rows(N, P);
    const int Kp = cs.Kp;
    const int e0 = blk * kEnsLanes + lane;
    const bool live = e0 < cs.K;
    const int e = live ? e0 : cs.K - 1;
    const size_t arr = size_t(Kp) * Np;
    double* __restrict__ base = a.data + cs.off_u;
    const double* __restrict__ u = base + size_t(a.cur) * arr;
    double* __restrict__ un = base + size_t(a.cur ^ 1) * arr;
    double* __restrict__ res = base + ENS_RES * arr;
    const double* __restrict__ qx = base + ENS_QX * arr;
    const double* __restrict__ qy = base + ENS_QY * arr;
    const double* __restrict__ g = a.geo + cs.off_geo;
    const int* __restrict__ cn = a.conn + cs.off_conn;
    const double ax = cs.ax, ay = cs.ay, nu = cs.nu;
    const int i0 = p * R;
    const double* __restrict__ DrT = a.ops[N] + i0;
    const double* __restrict__ DsT = DrT + Np * kEnsRows;
    const double* __restrict__ LfT = DsT + Np * kEnsRows;
    const double rx = g[0 * Kp + e], ry = g[1 * Kp + e], sx = g[2 * Kp + e], sy = g[3 * Kp + e];
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    // volume: -Dr (rx Fx + ry Fy) - Ds (sx Fx + sy Fy), F = a u - nu q
#pragma unroll 2
    for (int j = 0; j < Np; ++j) {
        const size_t id = nidx(e, j, Np);
        const double Fx = ax * u[id] - nu * qx[id], Fy = ay * u[id] - nu * qy[id];
        const double Fr = rx * Fx + ry * Fy, Fs = sx * Fx + sy * Fy;
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] -= DrT[j * kEnsRows + r] * Fr + DsT[j * kEnsRows + r] * Fs;
    }
    // faces: LIFT(fscale ((a.n u- - nu n.q-) - ((a.n u)^ - nu n.q^)))
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int c = cn[f * Kp + e];
        const int en = c >> 3, fn = (c >> 1) & 3;
        const bool own_side = c & 1;   // u^ = u-, so q^ = q+
        const double nx = g[(4 + 4 * f) * Kp + e], ny = g[(5 + 4 * f) * Kp + e];
        const double fs = g[(6 + 4 * f) * Kp + e], tf = g[(7 + 4 * f) * Kp + e];
        const double an = ax * nx + ay * ny;
#pragma unroll 1
        for (int k = 0; k < Nfp; ++k) {
            const size_t im = nidx(e, face_node<N>(f, k), Np);
            const size_t ip = nidx(en, face_node<N>(fn, N - k), Np);
            const double um = u[im], up = u[ip];
            const double qnm = nx * qx[im] + ny * qy[im];
            const double qnp = nx * qx[ip] + ny * qy[ip];
            const double fadv = an >= 0.0 ? an * um : an * up;
            const double qhat = (own_side ? qnp : qnm) - tf * (um - up);
            const double d = fs * ((an * um - nu * qnm) - (fadv - nu * qhat));
            const double* __restrict__ l = LfT + (f * Nfp + k) * kEnsRows;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += l[r] * d;
        }
    }
    if (!live) return;
    const double dt = cs.dt;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (i0 + r >= Np) continue;
        const size_t id = nidx(e0, i0 + r, Np);
        if (a.mode == 0) {
            res[id] = acc[r];
        } else {
            // the first stage (a = 0) does not read res, so it needs no clearing
            const double rn = (a.rk_a == 0.0 ? 0.0 : a.rk_a * res[id]) + dt * acc[r];
            res[id] = rn;
            un[id] = u[id] + a.rk_b * rn;
        }
    }
}

template <int PASS, int N, int P>
__device__ __forceinline__ void ens_run(const EnsArgs& a, const EnsCase& cs, const EnsItem& it, int lane) {
    if constexpr (ens_split_ok(N, P)) {
        if constexpr (PASS == 0) ens_q_item<N, P>(a, cs, it.blk, it.p, lane);
        else ens_rhs_item<N, P>(a, cs, it.blk, it.p, lane);
    }
}

template <int PASS, int N>
__device__ __forceinline__ void ens_split(const EnsArgs& a, const EnsCase& cs, const EnsItem& it, int lane) {
    switch (it.P) {
        case 1: ens_run<PASS, N, 1>(a, cs, it, lane); break;
        case 2: ens_run<PASS, N, 2>(a, cs, it, lane); break;
        case 4: ens_run<PASS, N, 4>(a, cs, it, lane); break;
        case 8: ens_run<PASS, N, 8>(a, cs, it, lane); break;
        default: break;
    }
}

template <int PASS>
__global__ __launch_bounds__(256) void ens_kernel(EnsArgs a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3, bx = blockIdx.x >> 3;
    const int lo = a.chunk[xcd], hi = a.chunk[xcd + 1];
    for (int i = lo + bx * 4 + wave; i < hi; i += nbx * 4) {
        const int it = __builtin_amdgcn_readfirstlane(i);
        const EnsItem item = a.items[it];
        const EnsCase cs = a.cases[item.c];
        switch (cs.N) {
            case 1: ens_split<PASS, 1>(a, cs, item, lane); break;
            case 2: ens_split<PASS, 2>(a, cs, item, lane); break;
            case 3: ens_split<PASS, 3>(a, cs, item, lane); break;
            case 4: ens_split<PASS, 4>(a, cs, item, lane); break;
            case 5: ens_split<PASS, 5>(a, cs, item, lane); break;
            case 6: ens_split<PASS, 6>(a, cs, item, lane); break;
            case 7: ens_split<PASS, 7>(a, cs, item, lane); break;
            case 8: ens_split<PASS, 8>(a, cs, item, lane); break;
            default: break;
        }
    }
}

hipError_t launch_ens_q(const EnsArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(ens_kernel<0>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ens_rhs(const EnsArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(ens_kernel<1>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

int ens_blocks_per_cu() {
    int b0 = 0, b1 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, ens_kernel<0>, 256, 0) != hipSuccess) b0 = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, ens_kernel<1>, 256, 0) != hipSuccess) b1 = 1;
    const int b = b0 < b1 ? b0 : b1;
    return b > 0 ? b : 1;
}

}  // namespace dg
