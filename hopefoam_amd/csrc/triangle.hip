// The reference's Triangle limiter (Trianglelimite::limite, Trianglelimite.C:61-852) for gfx950: a weighted
// Green-Gauss P1 reconstruction of every cell in primitive variables, with ghost cells on boundary faces.
//
// Three kernels, each HBM-bound (a handful of flops per byte), no atomics on data:
//   tri_cell_kernel     per cell: the four means (mean weights on the nodal values) and the four conserved
//                       variables at the three vertices -> cell[K/16][16][16]; one block of 256 threads per
//                       group of 16 cells, thread (cell, slot), reads of q coalesced over the 16 cells;
//   tri_grad_kernel     per cell, one thread: for each face the neighbour (or ghost) means and the face
//                       endpoint averages, the diamond gradients V_f of rho, u, v, p
//                         Vx =  ((phibar_n - phibar_o) ey + (phi_0 - phi_1) dyc) / (2 A)
//                         Vy = -((phibar_n - phibar_o) ex + (phi_0 - phi_1) dxc) / (2 A)
//                       and the cell gradient sum_f A_2 V_f / sum_f A_2 -> cv; V_f of boundary faces -> vg
//                       (the ghost cell's gradient).  Each interior face is evaluated from both sides (the
//                       formula is independent of the orientation), which removes the reference's separate
//                       face pass and its scatter;
//   tri_rebuild_kernel  per cell: nonlinear weights w_i = (g_j g_k + 1e-10) / (sum g^2 + 3e-10), g = |grad|^2,
//                       over the three neighbour gradients, then rho, rhou, rhov, E at every node from the
//                       limited primitive gradients, with the reference's density floor (du halved while
//                       rhobar + du < tol) -- bounded here: it stops once du is zero.
// Ghost cells: reflective wall (rho, E copied, momentum minus its normal part; endpoint states mirrored),
// fixedValue (the data at the face's first node; endpoint states from the data), otherwise a copy.
#include <hip/hip_runtime.h>

#include "dg_common.hpp"
#include "triangle.hpp"

namespace dg {

template <int N>
struct TShape {
    static constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
};

// node index of vertex i (0, 1, 2) of the reference triangle
template <int N>
__device__ __forceinline__ int vertex_node(int i) {
    return i == 0 ? 0 : (i == 1 ? N : TShape<N>::Np - 1);
}

template <int N>
__global__ __launch_bounds__(256) void tri_cell_kernel(TriArgs a, const double* __restrict__ q) {
    constexpr int Np = TShape<N>::Np;
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const size_t base = size_t(g) * 4 * Np * 16;
    double val;
    if (slot < 4) {
        val = 0.0;
        for (int n = 0; n < Np; ++n) val += a.pmean[n] * q[base + (size_t(slot) * Np + n) * 16 + el];
    } else {
        const int v = (slot - 4) / 3, i = (slot - 4) - 3 * v;
        val = q[base + (size_t(v) * Np + vertex_node<N>(i)) * 16 + el];
    }
    a.cell[(size_t(g) * kTriCell + slot) * 16 + el] = val;
}

__device__ __forceinline__ void primitives(const double* m, double gamma, double* p) {
    p[0] = m[0];
    p[1] = m[1] / m[0];
    p[2] = m[2] / m[0];
    p[3] = (gamma - 1.0) * (m[3] - 0.5 * (m[1] * m[1] + m[2] * m[2]) / m[0]);
}

template <int N>
__global__ __launch_bounds__(256) void tri_grad_kernel(TriArgs a) {
    constexpr int Nfp = TShape<N>::Nfp;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.K) return;
    const int g = e >> 4, l = e & 15;
    auto cell = [&](int c, int s) { return a.cell[(size_t(c >> 4) * kTriCell + s) * 16 + (c & 15)]; };
    auto geo = [&](int s) { return a.tgeo[(size_t(g) * kTriGeo + s) * 16 + l]; };
    double mo[4], po[4];
    for (int v = 0; v < 4; ++v) mo[v] = cell(e, v);
    primitives(mo, a.gamma, po);
    double V[3][8], A2[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int cn = a.conn[f * a.K + e];
        const double dxc = geo(4 + 8 * f), dyc = geo(5 + 8 * f), ex = geo(6 + 8 * f), ey = geo(7 + 8 * f);
        const double A = geo(8 + 8 * f);
        A2[f] = geo(9 + 8 * f);
        const int f1 = f == 2 ? 0 : f + 1;
        double q0[4], q1[4], n0[4], n1[4], mn[4];
        for (int v = 0; v < 4; ++v) {
            q0[v] = cell(e, 4 + 3 * v + f);
            q1[v] = cell(e, 4 + 3 * v + f1);
        }
        if (cn >= 0) {
            // the neighbour runs the shared edge the other way: its vertex fn + 1 is our first endpoint
            const int en = cn >> 2, fn = cn & 3, i0 = fn == 2 ? 0 : fn + 1;
            for (int v = 0; v < 4; ++v) {
                n0[v] = cell(en, 4 + 3 * v + i0);
                n1[v] = cell(en, 4 + 3 * v + fn);
                mn[v] = cell(en, v);
            }
        } else {
            const int tag = -cn - 1;
            if (tag == BC_FIXED && a.bcq) {
                // ghost mean: the first value of the face's patch (Trianglelimite.C:220-223) when the patch map is
                // set, else the face's own first node; the endpoints are the face's own data (:381-391)
                const int pf = a.pfirst ? a.pfirst[e * 3 + f] : -1;
                const int src = pf >= 0 ? pf : e * 3 + f;
                for (int v = 0; v < 4; ++v) {
                    const double* b = a.bcq + ((size_t(v) * a.K + e) * 3 + f) * Nfp;
                    n0[v] = b[0];
                    n1[v] = b[Nfp - 1];
                    mn[v] = a.bcq[(size_t(v) * a.K * 3 + src) * Nfp];
                }
            } else {
                for (int v = 0; v < 4; ++v) {
                    n0[v] = q0[v];
                    n1[v] = q1[v];
                    mn[v] = mo[v];
                }
                if (tag == BC_WALL) {
                    const double nx = geo(10 + 8 * f), ny = geo(11 + 8 * f);
                    const double u0 = q0[1] * nx + q0[2] * ny, u1 = q1[1] * nx + q1[2] * ny;
                    n0[1] = q0[1] - 2.0 * u0 * nx;
                    n0[2] = q0[2] - 2.0 * u0 * ny;
                    n1[1] = q1[1] - 2.0 * u1 * nx;
                    n1[2] = q1[2] - 2.0 * u1 * ny;
                    const double un = nx * mo[1] + ny * mo[2];
                    mn[1] = mo[1] - nx * un;
                    mn[2] = mo[2] - ny * un;
                }
            }
        }
        double pn[4], a0[4], a1[4], p0[4], p1[4];
        primitives(mn, a.gamma, pn);
        for (int v = 0; v < 4; ++v) {
            a0[v] = q0[v] * 0.5 + n0[v] * 0.5;
            a1[v] = q1[v] * 0.5 + n1[v] * 0.5;
        }
        // endpoint primitives as the reference forms them: p = (gamma-1)(E - rho (u^2 + v^2) / 2)
        p0[0] = a0[0];
        p0[1] = a0[1] / a0[0];
        p0[2] = a0[2] / a0[0];
        p0[3] = (a.gamma - 1.0) * (a0[3] - 0.5 * a0[0] * (p0[1] * p0[1] + p0[2] * p0[2]));
        p1[0] = a1[0];
        p1[1] = a1[1] / a1[0];
        p1[2] = a1[2] / a1[0];
        p1[3] = (a.gamma - 1.0) * (a1[3] - 0.5 * a1[0] * (p1[1] * p1[1] + p1[2] * p1[2]));
        for (int k = 0; k < 4; ++k) {
            const double dm = pn[k] - po[k], dv = p0[k] - p1[k];
            V[f][2 * k] = 0.5 * (dm * ey + dv * dyc) / A;
            V[f][2 * k + 1] = -0.5 * (dm * ex + dv * dxc) / A;
        }
        if (cn < 0)
            for (int k = 0; k < 8; ++k) a.vg[((size_t(g) * 3 + f) * 8 + k) * 16 + l] = V[f][k];
    }
    const double sa = A2[0] + A2[1] + A2[2];
    for (int k = 0; k < 8; ++k)
        a.cv[(size_t(g) * 8 + k) * 16 + l] = A2[0] * V[0][k] / sa + A2[1] * V[1][k] / sa + A2[2] * V[2][k] / sa;
}

template <int N>
__global__ __launch_bounds__(256) void tri_rebuild_kernel(TriArgs a, double* __restrict__ q) {
    constexpr int Np = TShape<N>::Np;
    __shared__ double sL[8][16];
    __shared__ int sfl[16];
    const int g = blockIdx.x, el = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int e0 = g * 16 + el;
    const bool live = e0 < a.K;
    const int e = live ? e0 : a.K - 1;
    const bool chosen = live && (!a.select || a.flags[e0] != 0);
    if (slot == 0) sfl[el] = 0;
    if (slot < 4) {
        const int k = slot;
        double cx[3], cy[3];
        for (int f = 0; f < 3; ++f) {
            const int cn = a.conn[f * a.K + e];
            if (cn >= 0) {
                const int en = cn >> 2;
                cx[f] = a.cv[(size_t(en >> 4) * 8 + 2 * k) * 16 + (en & 15)];
                cy[f] = a.cv[(size_t(en >> 4) * 8 + 2 * k + 1) * 16 + (en & 15)];
            } else {
                cx[f] = a.vg[((size_t(g) * 3 + f) * 8 + 2 * k) * 16 + el];
                cy[f] = a.vg[((size_t(g) * 3 + f) * 8 + 2 * k + 1) * 16 + el];
            }
        }
        const double g1 = cx[0] * cx[0] + cy[0] * cy[0];
        const double g2 = cx[1] * cx[1] + cy[1] * cy[1];
        const double g3 = cx[2] * cx[2] + cy[2] * cy[2];
        const double eps = 1e-10, fac = g1 * g1 + g2 * g2 + g3 * g3 + 3.0 * eps;
        const double w1 = (g2 * g3 + eps) / fac, w2 = (g1 * g3 + eps) / fac, w3 = (g2 * g1 + eps) / fac;
        sL[2 * k][el] = w1 * cx[0] + w2 * cx[1] + w3 * cx[2];
        sL[2 * k + 1][el] = w1 * cy[0] + w2 * cy[1] + w3 * cy[2];
    }
    __syncthreads();
    if (chosen) {
        auto cell = [&](int s) { return a.cell[(size_t(g) * kTriCell + s) * 16 + el]; };
        auto geo = [&](int s) { return a.tgeo[(size_t(g) * kTriGeo + s) * 16 + el]; };
        const double rb = cell(0), m1b = cell(1), m2b = cell(2), Eb = cell(3);
        const double ub = m1b / rb, vb = m2b / rb;
        const double xr = geo(0), xs = geo(1), yr = geo(2), ys = geo(3);
        const double L0 = sL[0][el], L1 = sL[1][el], L2 = sL[2][el], L3 = sL[3][el];
        const double L4 = sL[4][el], L5 = sL[5][el], L6 = sL[6][el], L7 = sL[7][el];
        const double gm1 = a.gamma - 1.0;
        int fl = 0;
        for (int n = slot; n < Np; n += 16) {
            const double r1 = a.rsn[2 * n], s1 = a.rsn[2 * n + 1];
            const double dx = xr * r1 + xs * s1, dy = yr * r1 + ys * s1;
            double du = dx * L0 + dy * L1;
            const double du1 = dx * L2 + dy * L3, du2 = dx * L4 + dy * L5, du3 = dx * L6 + dy * L7;
            double r = rb + du;
            // density floor; a finite du reaches zero within 2098 halvings
            for (int it = 0; it < 2100 && r < a.tol && du != 0.0; ++it) {
                fl |= 2;
                du *= 0.5;
                r = rb + du;
            }
            if (r < a.tol) fl |= 2;
            const double mu = m1b + rb * du1 + du * ub;
            const double mv = m2b + rb * du2 + du * vb;
            const double E = Eb + ((1.0 / gm1) * du3 + 0.5 * du * (ub * ub + vb * vb) + rb * (ub * du1 + vb * du2));
            q[dev_index(e0, 0, n, 4, Np)] = r;
            q[dev_index(e0, 1, n, 4, Np)] = mu;
            q[dev_index(e0, 2, n, 4, Np)] = mv;
            q[dev_index(e0, 3, n, 4, Np)] = E;
            if (gm1 * (E - 0.5 * (mu * mu + mv * mv) / r) < a.tol) fl |= 4;
        }
        if (fl) atomicOr(&sfl[el], fl);
    }
    __syncthreads();
    if (slot == 0 && live) a.flags[e0] = chosen ? (1 | sfl[el]) : 0;
}

template <int N>
static hipError_t triangle_n(const TriArgs& a, double* q, hipStream_t s) {
    const int groups = (a.K + 15) / 16;
    hipLaunchKernelGGL((tri_cell_kernel<N>), dim3(groups), dim3(256), 0, s, a, q);
    hipLaunchKernelGGL((tri_grad_kernel<N>), dim3((a.K + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL((tri_rebuild_kernel<N>), dim3(groups), dim3(256), 0, s, a, q);
    return hipGetLastError();
}

hipError_t launch_triangle(int N, const TriArgs& a, double* q, hipStream_t s) {
    switch (N) {
        case 1: return triangle_n<1>(a, q, s);
        case 2: return triangle_n<2>(a, q, s);
        case 3: return triangle_n<3>(a, q, s);
        case 4: return triangle_n<4>(a, q, s);
        case 5: return triangle_n<5>(a, q, s);
        case 6: return triangle_n<6>(a, q, s);
        case 7: return triangle_n<7>(a, q, s);
        case 8: return triangle_n<8>(a, q, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dg
