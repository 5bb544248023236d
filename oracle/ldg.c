/*
 * Oracle of the LDG advection-diffusion right-hand side (test infrastructure only).
 *
 * Not in the reference: HopeFOAM has no explicit diffusion (dgc::laplacian is an empty stub,
 * defaultLaplacianScheme.C:354-359).  This restates the published local discontinuous Galerkin method
 * (Cockburn & Shu, SIAM J. Numer. Anal. 35 (1998) 2440-2463, with the alternating fluxes C12 = beta/2 and
 * the penalty C11; nodal strong form as Hesthaven & Warburton, Nodal DG Methods (2008) section 7.1) for
 *     u_t + div(a u) = nu div(q),   q = grad u,
 * on a periodic triangle mesh, in the reference's element and geometry (the same nodes, Dr, Ds, fmask,
 * normals and Js / J as or_rhs_advect):
 *
 *   q   = grad_h u - LIFT( fscale n (u- - u^) )
 *   rhs = -div_h(a u - nu q) + LIFT( fscale [ (a.n u- - nu n.q-) - ((a.n u)^ - nu n.q^) ] )
 *
 *   div_h F = Dr (rx Fx + ry Fy) + Ds (sx Fx + sy Fy) (affine elements), LIFT = M^-1 E with E the face
 *   mass matrix on each face's nodes (the exact N+2-point Gauss rule), fscale = Js / J of the own side;
 *   face switch sigma = sign(beta . n-) (the lower element id wins a tie): u^ = u- if sigma > 0 else u+;
 *   n.q^ = n.q+ if sigma > 0 else n.q-, minus tau_f (u- - u+), tau_f = tau max(fscale-, fscale+);
 *   (a.n u)^ = upwind.
 *
 * With nu = 0 the right-hand side is the upwind advection operator in strong form, which equals the
 * reference's cubature weak form (or_rhs_advect) to round-off on affine elements.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

/* LIFT [Np x 3 Nfp]: M^-1 times the face mass matrix on face f's nodes (Vf^T diag(gw) Vf, exact) */
int or_ldg_lift(const or_mesh* m, double* lift)
{
    const int Np = m->Np, Nfp = m->Nfp, Nfq = m->Nfq;
    double* M1 = calloc((size_t)Nfp * Nfp, sizeof(double));
    for (int a = 0; a < Nfp; ++a)
        for (int b = 0; b < Nfp; ++b) {
            long double s = 0.0L;
            for (int g = 0; g < Nfq; ++g) s += (long double)m->Vf[g * Nfp + a] * m->gw[g] * m->Vf[g * Nfp + b];
            M1[a * Nfp + b] = (double)s;
        }
    for (int i = 0; i < Np; ++i)
        for (int f = 0; f < 3; ++f)
            for (int k = 0; k < Nfp; ++k) {
                long double s = 0.0L;
                for (int a = 0; a < Nfp; ++a) s += (long double)m->invM[i * Np + m->fmask[f * Nfp + a]] * M1[a * Nfp + k];
                lift[(size_t)i * 3 * Nfp + f * Nfp + k] = (double)s;
            }
    free(M1);
    return 0;
}

static int face_sigma(const or_mesh* m, int e, int f, double bx, double by)
{
    const double s = bx * m->nx[e * 3 + f] + by * m->ny[e * 3 + f];
    if (s > 0.0) return 1;
    if (s < 0.0) return -1;
    return e < m->EToE[e * 3 + f] ? 1 : -1;
}

/* q [2][K][Np] (may be NULL) and rhs [K][Np] of u [K][Np]; every face must have a neighbour (periodic) */
int or_ldg_rhs(const or_mesh* m, const double* u, double ax, double ay, double nu, double tau, double bx, double by,
               double* rhs, double* qout)
{
    const int K = m->K, Np = m->Np, Nfp = m->Nfp, N = m->N;
    for (int i = 0; i < 3 * K; ++i)
        if (m->EToE[i] < 0) return -1;
    double* lift = malloc(sizeof(double) * Np * 3 * Nfp);
    double* q = malloc(sizeof(double) * 2 * (size_t)K * Np);
    double *g = malloc(sizeof(double) * 3 * Nfp), *gx = malloc(sizeof(double) * 3 * Nfp);
    double* gy = malloc(sizeof(double) * 3 * Nfp);
    double *Fr = malloc(sizeof(double) * Np), *Fs = malloc(sizeof(double) * Np);
    or_ldg_lift(m, lift);
    /* q = grad u - LIFT(fscale n (u- - u^)) */
    for (int e = 0; e < K; ++e) {
        const double* ue = u + (size_t)e * Np;
        for (int f = 0; f < 3; ++f) {
            const int en = m->EToE[e * 3 + f], fn = m->EToF[e * 3 + f];
            const int sg = face_sigma(m, e, f, bx, by);
            const double fs = m->sJ[e * 3 + f] / m->J[e];
            for (int k = 0; k < Nfp; ++k) {
                const double um = ue[m->fmask[f * Nfp + k]];
                const double up = u[(size_t)en * Np + m->fmask[fn * Nfp + (N - k)]];
                const double d = sg > 0 ? 0.0 : um - up;
                gx[f * Nfp + k] = fs * m->nx[e * 3 + f] * d;
                gy[f * Nfp + k] = fs * m->ny[e * 3 + f] * d;
            }
        }
        for (int i = 0; i < Np; ++i) {
            double dr = 0.0, ds = 0.0, lx = 0.0, ly = 0.0;
            for (int j = 0; j < Np; ++j) {
                dr += m->Dr[i * Np + j] * ue[j];
                ds += m->Ds[i * Np + j] * ue[j];
            }
            for (int c = 0; c < 3 * Nfp; ++c) {
                lx += lift[(size_t)i * 3 * Nfp + c] * gx[c];
                ly += lift[(size_t)i * 3 * Nfp + c] * gy[c];
            }
            q[(size_t)e * Np + i] = m->rx[e] * dr + m->sx[e] * ds - lx;
            q[((size_t)K + e) * Np + i] = m->ry[e] * dr + m->sy[e] * ds - ly;
        }
    }
    /* rhs = -div(a u - nu q) + LIFT(fscale (n.F- - n.F^)) */
    for (int e = 0; e < K; ++e) {
        const double* ue = u + (size_t)e * Np;
        const double *qx = q + (size_t)e * Np, *qy = q + ((size_t)K + e) * Np;
        for (int j = 0; j < Np; ++j) {
            const double Fx = ax * ue[j] - nu * qx[j], Fy = ay * ue[j] - nu * qy[j];
            Fr[j] = m->rx[e] * Fx + m->ry[e] * Fy;
            Fs[j] = m->sx[e] * Fx + m->sy[e] * Fy;
        }
        for (int f = 0; f < 3; ++f) {
            const int en = m->EToE[e * 3 + f], fn = m->EToF[e * 3 + f];
            const int sg = face_sigma(m, e, f, bx, by);
            const double nx = m->nx[e * 3 + f], ny = m->ny[e * 3 + f];
            const double fs = m->sJ[e * 3 + f] / m->J[e];
            const double tf = tau * m->fscale_max[e * 3 + f];
            const double an = ax * nx + ay * ny;
            for (int k = 0; k < Nfp; ++k) {
                const int im = m->fmask[f * Nfp + k];
                const size_t ip = (size_t)en * Np + m->fmask[fn * Nfp + (N - k)];
                const double um = ue[im], up = u[ip];
                const double qnm = nx * qx[im] + ny * qy[im];
                const double qnp = nx * q[ip] + ny * q[(size_t)K * Np + ip];
                const double fadv = an >= 0.0 ? an * um : an * up;
                const double qhat = (sg > 0 ? qnp : qnm) - tf * (um - up);
                g[f * Nfp + k] = fs * ((an * um - nu * qnm) - (fadv - nu * qhat));
            }
        }
        for (int i = 0; i < Np; ++i) {
            double a = 0.0;
            for (int j = 0; j < Np; ++j) a += m->Dr[i * Np + j] * Fr[j] + m->Ds[i * Np + j] * Fs[j];
            double l = 0.0;
            for (int c = 0; c < 3 * Nfp; ++c) l += lift[(size_t)i * 3 * Nfp + c] * g[c];
            rhs[(size_t)e * Np + i] = -a + l;
        }
    }
    if (qout) memcpy(qout, q, sizeof(double) * 2 * (size_t)K * Np);
    free(lift); free(q); free(g); free(gx); free(gy); free(Fr); free(Fs);
    return 0;
}

/* nsteps LSERK4 steps of the LDG right-hand side on u [K][Np] in place */
int or_ldg_lserk4(const or_mesh* m, double* u, int nsteps, double dt, double ax, double ay, double nu, double tau,
                  double bx, double by)
{
    const size_t n = (size_t)m->K * m->Np;
    double *res = calloc(n, sizeof(double)), *rhs = malloc(sizeof(double) * n);
    double A[5], B[5], C[5];
    or_lserk4_coeffs(A, B, C);
    int rc = 0;
    for (int step = 0; step < nsteps && !rc; ++step)
        for (int s = 0; s < 5 && !rc; ++s) {
            rc = or_ldg_rhs(m, u, ax, ay, nu, tau, bx, by, rhs, NULL);
            for (size_t i = 0; i < n; ++i) {
                res[i] = A[s] * res[i] + dt * rhs[i];
                u[i] += B[s] * res[i];
            }
        }
    free(res); free(rhs);
    return rc;
}
