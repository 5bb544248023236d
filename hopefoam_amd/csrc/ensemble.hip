// Ensembles of independent LDG advection-diffusion cases (u_t + div(a u) = nu div(q), q = grad u; periodic
// triangle meshes of different sizes and orders) for gfx950.  The method is the local discontinuous
// Galerkin scheme of Cockburn & Shu (1998) with alternating fluxes (u^ from the side the switch vector
// points away from, q^ from the other, penalty tau on u jumps) in the nodal strong form; it has no
// counterpart in the reference, whose dgc::laplacian is empty.
//
// Each LSERK4 stage is two passes (the auxiliary q of an element's neighbours must be complete before the
// right-hand side reads it -- a grid-wide dependency that a kernel boundary settles without fences or spin
// waits), each pass one persistent launch per order present in the ensemble over all cases of that order:
//   ens_kernel<0, N, P>  q = grad u - LIFT(fscale n (u- - u^))                   -> qx, qy
//   ens_kernel<1, N, P>  rhs = -div(a u - nu q) + LIFT(fscale (n.F- - n.F^)), then res = a res + dt rhs,
//                        u_next = u + b res
// Orders (and row splits) get kernels of their own so that each is register-allocated for its own shape
// (one kernel holding every order needs the union of all their registers, and spills).
// A work item is a run of 64-element blocks of one case, taken by a whole workgroup of kEnsWaves waves: one
// element per lane, each block's node rows split into P slices, the (block, slice) tasks spread over the waves.
// All lanes of a wave need the same operator entry at a time, so the order's operators (transposed, rows
// padded to kEnsRows with zeros: the R entries of one column a slice needs are contiguous and need no
// clamping) sit in LDS and are read as broadcasts.  Element and face loads are issued in batches ahead of
// the arithmetic that consumes them.
// The slice count P per order (rows per lane R = ceil(Np / P) <= 8) is the tile shape the autotuner picks.
// Each order's items are split into 8 contiguous ranges of equal estimated cost, one per XCD, so that the
// neighbour traces an item reads were brought into the same XCD's L2 by the items next to it.
#include <hip/hip_runtime.h>

#include "ensemble.hpp"

namespace dg {

__host__ __device__ constexpr int ens_np(int N) { return (N + 1) * (N + 2) / 2; }
__host__ __device__ constexpr int ens_rows(int N, int P) { return (ens_np(N) + P - 1) / P; }
__host__ __device__ constexpr bool ens_split_ok(int N, int P) {
    return N >= 1 && N <= 8 && (P == 1 || P == 2 || P == 4 || P == 8) && ens_rows(N, P) <= 8 &&
           (P == 1 || ens_rows(N, P) >= 2);
}

bool ens_split_supported(int N, int P) { return P == kEnsMfma ? (N >= 2 && N <= 8) : ens_split_ok(N, P); }

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

using d4_t = double __attribute__((ext_vector_type(4)));
__device__ __forceinline__ d4_t mfma64(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// MFMA variant: a wave owns 16 elements (group grp of the case); lane (col, rg) = (lane % 16, lane / 16).
// B operands: K step s holds node 4 s + rg of element col.  A operands: the operator fragments in LDS.
// Accumulators: register r of row tile m holds row 16 m + rg + 4 r of element col.
template <int N>
__device__ void ens_q_mfma(const EnsArgs& a, const EnsCase& cs, const double* ops, int grp, int lane) {
    constexpr int Np = ens_np(N), Nfp = N + 1;
    constexpr int MT = ens_mfma_mt(N), KS = ens_mfma_ks(N), KSL = ens_mfma_ksl(N);
    const int Kp = cs.Kp, col = lane & 15, rg = lane >> 4;
    const int e0 = grp * 16 + col;
    const bool live = e0 < cs.K;
    const int e = live ? e0 : cs.K - 1;
    const size_t arr = size_t(Kp) * Np;
    const double* __restrict__ u = a.data + cs.off_u + size_t(a.cur) * arr;
    double* __restrict__ qx = a.data + cs.off_u + ENS_QX * arr;
    double* __restrict__ qy = a.data + cs.off_u + ENS_QY * arr;
    const double* __restrict__ g = a.geo + cs.off_geo;
    const int* __restrict__ cn = a.conn + cs.off_conn;
    const double* __restrict__ DrF = ops;
    const double* __restrict__ DsF = ops + MT * KS * 64;
    const double* __restrict__ LF = ops + 2 * MT * KS * 64;
    double ub[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) ub[s] = 4 * s + rg < Np ? u[nidx(e, 4 * s + rg, Np)] : 0.0;
    // jumps at face node 4 s + rg (face f = idx / Nfp, point k = idx % Nfp)
    double gxb[KSL], gyb[KSL];
#pragma unroll
    for (int s = 0; s < KSL; ++s) {
        const int idx = 4 * s + rg;
        gxb[s] = gyb[s] = 0.0;
        if (idx < 3 * Nfp) {
            const int f = idx / Nfp, k = idx - f * Nfp;
            const int c = cn[f * Kp + e];
            const int en = c >> 3, fn = (c >> 1) & 3;
            if (!(c & 1)) {
                const double um = u[nidx(e, face_node<N>(f, k), Np)];
                const double up = u[nidx(en, face_node<N>(fn, N - k), Np)];
                const double d = g[(6 + 4 * f) * Kp + e] * (um - up);
                gxb[s] = g[(4 + 4 * f) * Kp + e] * d;
                gyb[s] = g[(5 + 4 * f) * Kp + e] * d;
            }
        }
    }
    d4_t ar[MT], as[MT], lx[MT], ly[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) ar[m] = as[m] = lx[m] = ly[m] = d4_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            ar[m] = mfma64(DrF[(m * KS + s) * 64 + lane], ub[s], ar[m]);
            as[m] = mfma64(DsF[(m * KS + s) * 64 + lane], ub[s], as[m]);
        }
#pragma unroll
    for (int s = 0; s < KSL; ++s)
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const double L = LF[(m * KSL + s) * 64 + lane];
            lx[m] = mfma64(L, gxb[s], lx[m]);
            ly[m] = mfma64(L, gyb[s], ly[m]);
        }
    if (!live) return;
    const double rx = g[0 * Kp + e], ry = g[1 * Kp + e], sx = g[2 * Kp + e], sy = g[3 * Kp + e];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * m + rg + 4 * r;
            if (i < Np) {
                qx[nidx(e0, i, Np)] = rx * ar[m][r] + sx * as[m][r] - lx[m][r];
                qy[nidx(e0, i, Np)] = ry * ar[m][r] + sy * as[m][r] - ly[m][r];
            }
        }
}

template <int N>
__device__ void ens_rhs_mfma(const EnsArgs& a, const EnsCase& cs, const double* ops, int grp, int lane) {
    constexpr int Np = ens_np(N), Nfp = N + 1;
    constexpr int MT = ens_mfma_mt(N), KS = ens_mfma_ks(N), KSL = ens_mfma_ksl(N);
    const int Kp = cs.Kp, col = lane & 15, rg = lane >> 4;
    const int e0 = grp * 16 + col;
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
    const double* __restrict__ DrF = ops;
    const double* __restrict__ DsF = ops + MT * KS * 64;
    const double* __restrict__ LF = ops + 2 * MT * KS * 64;
    const double ax = cs.ax, ay = cs.ay, nu = cs.nu;
    const double rx = g[0 * Kp + e], ry = g[1 * Kp + e], sx = g[2 * Kp + e], sy = g[3 * Kp + e];
    // B operands: -Fr = -(rx Fx + ry Fy), -Fs = -(sx Fx + sy Fy) at node 4 s + rg, F = a u - nu q (negated so
    // that the volume and lift products share one accumulator)
    double frb[KS], fsb[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        frb[s] = fsb[s] = 0.0;
        if (4 * s + rg < Np) {
            const size_t id = nidx(e, 4 * s + rg, Np);
            const double Fx = ax * u[id] - nu * qx[id], Fy = ay * u[id] - nu * qy[id];
            frb[s] = -(rx * Fx + ry * Fy);
            fsb[s] = -(sx * Fx + sy * Fy);
        }
    }
    // flux differences at face node 4 s + rg
    double gb[KSL];
#pragma unroll
    for (int s = 0; s < KSL; ++s) {
        const int idx = 4 * s + rg;
        gb[s] = 0.0;
        if (idx < 3 * Nfp) {
            const int f = idx / Nfp, k = idx - f * Nfp;
            const int c = cn[f * Kp + e];
            const int en = c >> 3, fn = (c >> 1) & 3;
            const bool own_side = c & 1;
            const double nx = g[(4 + 4 * f) * Kp + e], ny = g[(5 + 4 * f) * Kp + e];
            const double fs = g[(6 + 4 * f) * Kp + e], tf = g[(7 + 4 * f) * Kp + e];
            const size_t im = nidx(e, face_node<N>(f, k), Np);
            const size_t ip = nidx(en, face_node<N>(fn, N - k), Np);
            const double um = u[im], up = u[ip];
            const double qnm = nx * qx[im] + ny * qy[im];
            const double qnp = nx * qx[ip] + ny * qy[ip];
            const double an = ax * nx + ay * ny;
            const double fadv = an >= 0.0 ? an * um : an * up;
            const double qhat = (own_side ? qnp : qnm) - tf * (um - up);
            gb[s] = fs * ((an * um - nu * qnm) - (fadv - nu * qhat));
        }
    }
    d4_t acc[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m] = d4_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < KSL; ++s)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m] = mfma64(LF[(m * KSL + s) * 64 + lane], gb[s], acc[m]);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            acc[m] = mfma64(DrF[(m * KS + s) * 64 + lane], frb[s], acc[m]);
            acc[m] = mfma64(DsF[(m * KS + s) * 64 + lane], fsb[s], acc[m]);
        }
    if (!live) return;
    const double dt = cs.dt;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * m + rg + 4 * r;
            if (i >= Np) continue;
            const size_t id = nidx(e0, i, Np);
            const double rhs = acc[m][r];
            if (a.mode == 0) {
                res[id] = rhs;
            } else {
                const double rn = (a.rk_a == 0.0 ? 0.0 : a.rk_a * res[id]) + dt * rhs;
                res[id] = rn;
                un[id] = u[id] + a.rk_b * rn;
            }
        }
}

template <int N, int P>
__device__ void ens_q_item(const EnsArgs& a, const EnsCase& cs, const double* ops, int blk, int p, int lane) {
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
    const double* __restrict__ DrT = ops + i0;                      // DrT[j * kEnsRows + i]
    const double* __restrict__ DsT = DrT + Np * kEnsRows;
    const double* __restrict__ LfT = DsT + Np * kEnsRows;           // LIFT^T[(f Nfp + k) * kEnsRows + i]
    double ar[R], as[R];
#pragma unroll
    for (int r = 0; r < R; ++r) ar[r] = as[r] = 0.0;
    // columns in chunks whose loads are all issued before the chunk's arithmetic (latency hiding)
    constexpr int CH = Np < 16 ? 4 : 8;
#pragma unroll 1
    for (int j0 = 0; j0 < Np; j0 += CH) {
        double uv[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) uv[c] = j0 + c < Np ? u[nidx(e, j0 + c, Np)] : 0.0;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            if (j0 + c >= Np) break;
            const int j = j0 + c;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                ar[r] += DrT[j * kEnsRows + r] * uv[c];
                as[r] += DsT[j * kEnsRows + r] * uv[c];
            }
        }
    }
    const double rx = g[0 * Kp + e], ry = g[1 * Kp + e], sx = g[2 * Kp + e], sy = g[3 * Kp + e];
    double ox[R], oy[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        ox[r] = rx * ar[r] + sx * as[r];
        oy[r] = ry * ar[r] + sy * as[r];
    }
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
        const int c = cn[f * Kp + e];
        const int en = c >> 3, fn = (c >> 1) & 3;
        const bool own_side = c & 1;   // u^ = u-: no jump term
        const double nx = g[(4 + 4 * f) * Kp + e], ny = g[(5 + 4 * f) * Kp + e], fs = g[(6 + 4 * f) * Kp + e];
        double um[Nfp], up[Nfp];
#pragma unroll
        for (int k = 0; k < Nfp; ++k) {
            um[k] = u[nidx(e, face_node<N>(f, k), Np)];
            up[k] = u[nidx(en, face_node<N>(fn, N - k), Np)];
        }
#pragma unroll
        for (int k = 0; k < Nfp; ++k) {
            const double d = own_side ? 0.0 : fs * (um[k] - up[k]);
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
__device__ void ens_rhs_item(const EnsArgs& a, const EnsCase& cs, const double* ops, int blk, int p, int lane) {
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
    const double* __restrict__ DrT = ops + i0;
    const double* __restrict__ DsT = DrT + Np * kEnsRows;
    const double* __restrict__ LfT = DsT + Np * kEnsRows;
    const double rx = g[0 * Kp + e], ry = g[1 * Kp + e], sx = g[2 * Kp + e], sy = g[3 * Kp + e];
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    // volume: -Dr (rx Fx + ry Fy) - Ds (sx Fx + sy Fy), F = a u - nu q
    constexpr int CH = Np < 4 ? Np : 4;
#pragma unroll 1
    for (int j0 = 0; j0 < Np; j0 += CH) {
        double uv[CH], xv[CH], yv[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const size_t id = nidx(e, j0 + c < Np ? j0 + c : Np - 1, Np);
            uv[c] = u[id];
            xv[c] = qx[id];
            yv[c] = qy[id];
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            if (j0 + c >= Np) break;
            const int j = j0 + c;
            const double Fx = ax * uv[c] - nu * xv[c], Fy = ay * uv[c] - nu * yv[c];
            const double Fr = rx * Fx + ry * Fy, Fs = sx * Fx + sy * Fy;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] -= DrT[j * kEnsRows + r] * Fr + DsT[j * kEnsRows + r] * Fs;
        }
    }
    // faces: LIFT(fscale ((a.n u- - nu n.q-) - ((a.n u)^ - nu n.q^)))
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
        const int c = cn[f * Kp + e];
        const int en = c >> 3, fn = (c >> 1) & 3;
        const bool own_side = c & 1;   // u^ = u-, so q^ = q+
        const double nx = g[(4 + 4 * f) * Kp + e], ny = g[(5 + 4 * f) * Kp + e];
        const double fs = g[(6 + 4 * f) * Kp + e], tf = g[(7 + 4 * f) * Kp + e];
        const double an = ax * nx + ay * ny;
        constexpr int KC = Nfp < 5 ? Nfp : 5;   // face points per batch of loads
#pragma unroll 1
        for (int k0 = 0; k0 < Nfp; k0 += KC) {
            double bu[KC], bq[KC], cu[KC], cq[KC];
#pragma unroll
            for (int t = 0; t < KC; ++t) {
                const int kk = k0 + t < Nfp ? k0 + t : Nfp - 1;
                const size_t im = nidx(e, face_node<N>(f, kk), Np);
                const size_t ip = nidx(en, face_node<N>(fn, N - kk), Np);
                bu[t] = u[im];
                bq[t] = nx * qx[im] + ny * qy[im];
                cu[t] = u[ip];
                cq[t] = nx * qx[ip] + ny * qy[ip];
            }
#pragma unroll
            for (int t = 0; t < KC; ++t) {
            const int k = k0 + t;
            if (k >= Nfp) break;
            const double um = bu[t], up = cu[t];
            const double qnm = bq[t], qnp = cq[t];
            const double fadv = an >= 0.0 ? an * um : an * up;
            const double qhat = (own_side ? qnp : qnm) - tf * (um - up);
            const double d = fs * ((an * um - nu * qnm) - (fadv - nu * qhat));
            const double* __restrict__ l = LfT + (f * Nfp + k) * kEnsRows;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += l[r] * d;
            }
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

// MFMA variant: items are runs of 16-element groups, one group per wave; fragments staged into LDS
template <int PASS, int N>
__global__ __launch_bounds__(64 * kEnsWaves) void ens_mfma_kernel(EnsArgs a) {
    constexpr int nop = (2 * ens_mfma_mt(N) * ens_mfma_ks(N) + ens_mfma_mt(N) * ens_mfma_ksl(N)) * 64;
    __shared__ double sop[nop];
    for (int i = threadIdx.x; i < nop; i += 64 * kEnsWaves) sop[i] = a.opsM[N][i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3, bx = blockIdx.x >> 3;
    const int lo = a.chunk[xcd], hi = a.chunk[xcd + 1];
    for (int it = lo + bx; it < hi; it += nbx) {
        const EnsItem item = a.items[it];
        const EnsCase cs = a.cases[item.c];
        const int ngrp = cs.Kp / 16;
        for (int t = wave; t < item.nb; t += kEnsWaves) {
            const int grp = item.b0 + t;
            if (grp >= ngrp) break;
            if constexpr (PASS == 0) ens_q_mfma<N>(a, cs, sop, grp, lane);
            else ens_rhs_mfma<N>(a, cs, sop, grp, lane);
        }
    }
}

// One persistent kernel per (pass, order, split) over that order's items: each is register-allocated for its
// own shape.  The order's operators are copied into LDS once per workgroup.
template <int PASS, int N, int P>
__global__ __launch_bounds__(64 * kEnsWaves) __attribute__((amdgpu_waves_per_eu(4))) void ens_kernel(EnsArgs a) {
    constexpr int nop = (2 * ens_np(N) + 3 * (N + 1)) * kEnsRows;
    __shared__ double sop[nop];
    for (int i = threadIdx.x; i < nop; i += 64 * kEnsWaves) sop[i] = a.ops[N][i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3, bx = blockIdx.x >> 3;
    const int lo = a.chunk[xcd], hi = a.chunk[xcd + 1];
    for (int it = lo + bx; it < hi; it += nbx) {
        const EnsItem item = a.items[it];
        const EnsCase cs = a.cases[item.c];
        const int nblk = cs.Kp / kEnsLanes;
        for (int t = wave; t < item.nb * P; t += kEnsWaves) {
            const int blk = item.b0 + t / P, p = t % P;
            if (blk >= nblk) break;
            if constexpr (PASS == 0) ens_q_item<N, P>(a, cs, sop, blk, p, lane);
            else ens_rhs_item<N, P>(a, cs, sop, blk, p, lane);
        }
    }
}

template <int PASS, int N, int P>
static hipError_t launch_np(const EnsArgs& a, int grid, hipStream_t s) {
    if constexpr (ens_split_ok(N, P)) {
        hipLaunchKernelGGL((ens_kernel<PASS, N, P>), dim3(grid), dim3(64 * kEnsWaves), 0, s, a);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

template <int PASS, int N>
static hipError_t launch_n(int P, const EnsArgs& a, int grid, hipStream_t s) {
    if constexpr (N >= 2) {
        if (P == kEnsMfma) {
            hipLaunchKernelGGL((ens_mfma_kernel<PASS, N>), dim3(grid), dim3(64 * kEnsWaves), 0, s, a);
            return hipGetLastError();
        }
    }
    switch (P) {
        case 1: return launch_np<PASS, N, 1>(a, grid, s);
        case 2: return launch_np<PASS, N, 2>(a, grid, s);
        case 4: return launch_np<PASS, N, 4>(a, grid, s);
        case 8: return launch_np<PASS, N, 8>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <int PASS>
static hipError_t launch_pass(int N, int P, const EnsArgs& a, int grid, hipStream_t s) {
    switch (N) {
        case 1: return launch_n<PASS, 1>(P, a, grid, s);
        case 2: return launch_n<PASS, 2>(P, a, grid, s);
        case 3: return launch_n<PASS, 3>(P, a, grid, s);
        case 4: return launch_n<PASS, 4>(P, a, grid, s);
        case 5: return launch_n<PASS, 5>(P, a, grid, s);
        case 6: return launch_n<PASS, 6>(P, a, grid, s);
        case 7: return launch_n<PASS, 7>(P, a, grid, s);
        case 8: return launch_n<PASS, 8>(P, a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ens_q(int N, int P, const EnsArgs& a, int grid, hipStream_t s) {
    return launch_pass<0>(N, P, a, grid, s);
}
hipError_t launch_ens_rhs(int N, int P, const EnsArgs& a, int grid, hipStream_t s) {
    return launch_pass<1>(N, P, a, grid, s);
}

// resident workgroups per CU: LDS (the order's operator table or MFMA fragments) and four waves per SIMD
int ens_blocks_per_cu(int N, int P) {
    const size_t lds = P == kEnsMfma
        ? size_t(2 * ens_mfma_mt(N) * ens_mfma_ks(N) + ens_mfma_mt(N) * ens_mfma_ksl(N)) * 64 * sizeof(double)
        : size_t(2 * ens_np(N) + 3 * (N + 1)) * kEnsRows * sizeof(double);
    const int by_lds = int((160u * 1024u) / lds);
    const int by_waves = 16 / kEnsWaves;
    const int b = by_lds < by_waves ? by_lds : by_waves;
    return b > 0 ? b : 1;
}

}  // namespace dg
