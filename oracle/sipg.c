/*
 * Oracle SIPG Laplacian and PETSc-semantics CG (test infrastructure only).
 *
 * Assembly restates HopeFOAM-0.1/src/DG/DG/laplacianSchemes/defaultLaplacianScheme/
 * defaultLaplacianScheme.C:62-288 (matrix = the NEGATIVE SIPG form):
 *   diag_e   -= sum_k WJ_k grad(phi_i).grad(phi_j)                         (:96-102, cubature)
 *   interior face, own side (:222-286):
 *     diag_e += -1/2 gtau Vf^T W Vf + 1/2 Vf^T W Dn_e + 1/2 Dn_e^T W Vf
 *     lu2_en += +1/2 gtau Vf^T W Vf' - 1/2 Vf^T W Dn_en' - 1/2 Dn_e^T W Vf'
 *   Dirichlet face (:153-209):
 *     diag_e += -gtau Vf^T W Vf + Vf^T W Dn_e + Dn_e^T W Vf
 *     b_e    += -gtau Vf^T W g + Dn_e^T W g
 * with gtau = fscale * 20 * (N+1)^2 (:114), fscale = max over the two sides of Js/J
 * (physicalElementData.C:247-259), W = diag(Js w_g) of the owner face, Dn_e the derivative
 * along e's outward normal at the face Gauss points (physicalElementData.C:262-343), primes
 * the neighbour's functions at the same physical points (rotated map).  The right-hand side of
 * dgm::laplacian(u) == f is b = M_e f + boundary terms (EquationEqual.H:84-92,
 * EquationGeometricDofField.H:84-102, dgMatrix.C:383-406).
 *
 * The CG restates PETSc 3.7 KSPCG + PCJACOBI with KSP_NORM_PRECONDITIONED and
 * KSPConvergedDefault (relative to the initial preconditioned residual, nonzero initial guess:
 * dgMatrix.C:163-283), from the published algorithm.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

/* gradient of element e's basis at its face f Gauss points: Dx[g][j], Dy[g][j] (Nfq x Np),
 * obtained like gaussOwnerFaceDn_: nodal derivative matrices restricted to the face nodes and
 * interpolated with Vf (node order reversed when `reversed`, i.e. in the other side's orientation). */
static void face_grad(const or_mesh* m, int e, int f, int reversed, double* Dx, double* Dy)
{
    int Np = m->Np, Nfp = m->Nfp, Nfq = m->Nfq, N = m->N;
    const int* fm = m->fmask + f * Nfp;
    double rx = m->rx[e], ry = m->ry[e], sx = m->sx[e], sy = m->sy[e];
    for (int g = 0; g < Nfq; ++g)
        for (int j = 0; j < Np; ++j) {
            double dx = 0.0, dy = 0.0;
            for (int k = 0; k < Nfp; ++k) {
                int node = fm[reversed ? N - k : k];
                double dr = m->Dr[node * Np + j], ds = m->Ds[node * Np + j];
                double w = m->Vf[g * Nfp + k];
                dx += w * (rx * dr + sx * ds);
                dy += w * (ry * dr + sy * ds);
            }
            Dx[g * Np + j] = dx;
            Dy[g * Np + j] = dy;
        }
}

/* face interpolation of element e onto the Gauss points of its face f (full Np columns) */
static void face_vand(const or_mesh* m, int f, int reversed, double* V)
{
    int Np = m->Np, Nfp = m->Nfp, Nfq = m->Nfq, N = m->N;
    const int* fm = m->fmask + f * Nfp;
    memset(V, 0, sizeof(double) * Nfq * Np);
    for (int g = 0; g < Nfq; ++g)
        for (int k = 0; k < Nfp; ++k) V[g * Np + fm[reversed ? N - k : k]] += m->Vf[g * Nfp + k];
}

int or_sipg_nnz(const or_mesh* m)
{
    int nnz = 0;
    for (int e = 0; e < m->K; ++e) {
        int blocks = 1;
        for (int f = 0; f < 3; ++f) blocks += m->EToE[e * 3 + f] >= 0;
        nnz += blocks * m->Np * m->Np;
    }
    return nnz;
}

/* penalty of face f of element e */
static double face_tau(const or_mesh* m, int e, int f, int penalty_mode)
{
    double fs = m->fscale_max[e * 3 + f];
    double n1 = m->N + 1.0;
    return penalty_mode == 0 ? fs * 20.0 * n1 * n1 : fs * n1 * n1;
}

int or_sipg_assemble(const or_mesh* m, int penalty_mode, const double* f, const double* gbc, int* rowptr, int* col,
                     double* val, double* b)
{
    int Np = m->Np, Nq = m->Nq, Nfq = m->Nfq, Nfp = m->Nfp, K = m->K;
    double* blk = malloc(sizeof(double) * 4 * Np * Np);   /* diag + up to 3 neighbour blocks */
    double *V1 = malloc(sizeof(double) * Nfq * Np), *V2 = malloc(sizeof(double) * Nfq * Np);
    double *D1x = malloc(sizeof(double) * Nfq * Np), *D1y = malloc(sizeof(double) * Nfq * Np);
    double *D2x = malloc(sizeof(double) * Nfq * Np), *D2y = malloc(sizeof(double) * Nfq * Np);
    double *Dn1 = malloc(sizeof(double) * Nfq * Np), *Dn2 = malloc(sizeof(double) * Nfq * Np);
    double *gx = malloc(sizeof(double) * Nq * Np), *gy = malloc(sizeof(double) * Nq * Np);
    double* bg = malloc(sizeof(double) * Nfq);
    int nz = 0;
    rowptr[0] = 0;
    for (int e = 0; e < K; ++e) {
        memset(blk, 0, sizeof(double) * 4 * Np * Np);
        double J = m->J[e], rx = m->rx[e], ry = m->ry[e], sx = m->sx[e], sy = m->sy[e];
        /* volume: diag -= sum_k WJ_k grad phi_i . grad phi_j */
        for (int q = 0; q < Nq; ++q)
            for (int j = 0; j < Np; ++j) {
                gx[q * Np + j] = m->Drq[q * Np + j] * rx + m->Dsq[q * Np + j] * sx;
                gy[q * Np + j] = m->Drq[q * Np + j] * ry + m->Dsq[q * Np + j] * sy;
            }
        for (int q = 0; q < Nq; ++q) {
            double wj = J * m->cw[q];
            for (int i = 0; i < Np; ++i)
                for (int j = 0; j < Np; ++j)
                    blk[i * Np + j] -= (gx[q * Np + i] * gx[q * Np + j] + gy[q * Np + i] * gy[q * Np + j]) * wj;
        }
        /* b = M_e f */
        for (int i = 0; i < Np; ++i) {
            double a = 0.0;
            for (int j = 0; j < Np; ++j) a += J * m->M[i * Np + j] * f[(size_t)e * Np + j];
            b[(size_t)e * Np + i] = a;
        }
        int nbl[3] = {-1, -1, -1};
        for (int fc = 0; fc < 3; ++fc) {
            int en = m->EToE[e * 3 + fc], fn = m->EToF[e * 3 + fc];
            double nx = m->nx[e * 3 + fc], ny = m->ny[e * 3 + fc];
            /* the face quadrature weights of the owner (smaller index) are used on both sides */
            int owner = (en < 0 || e < en) ? e : en, of = owner == e ? fc : fn;
            double Js = m->sJ[owner * 3 + of];
            double tau = face_tau(m, e, fc, penalty_mode);
            face_vand(m, fc, 0, V1);
            face_grad(m, e, fc, 0, D1x, D1y);
            for (int i = 0; i < Nfq * Np; ++i) Dn1[i] = nx * D1x[i] + ny * D1y[i];
            if (en >= 0) {
                face_vand(m, fn, 1, V2);
                face_grad(m, en, fn, 1, D2x, D2y);
                /* the other side's outward normal derivative is -n . grad, and enters as -1/2 Vf^T W Dn' */
                for (int i = 0; i < Nfq * Np; ++i) Dn2[i] = nx * D2x[i] + ny * D2y[i];
                int slot = 1;
                while (nbl[slot - 1] >= 0) ++slot;
                nbl[slot - 1] = en;
                double* L = blk + slot * Np * Np;
                for (int g = 0; g < Nfq; ++g) {
                    double w = Js * m->gw[g];
                    for (int i = 0; i < Np; ++i)
                        for (int j = 0; j < Np; ++j) {
                            blk[i * Np + j] += w * (-0.5 * tau * V1[g * Np + i] * V1[g * Np + j]
                                                    + 0.5 * V1[g * Np + i] * Dn1[g * Np + j]
                                                    + 0.5 * Dn1[g * Np + i] * V1[g * Np + j]);
                            L[i * Np + j] += w * (0.5 * tau * V1[g * Np + i] * V2[g * Np + j]
                                                  + 0.5 * V1[g * Np + i] * Dn2[g * Np + j]
                                                  - 0.5 * Dn1[g * Np + i] * V2[g * Np + j]);
                        }
                }
            } else if (m->bctag[e * 3 + fc] == 1) {
                /* Dirichlet: g interpolated from the owner's face nodes */
                for (int g = 0; g < Nfq; ++g) {
                    double a = 0.0;
                    for (int k = 0; k < Nfp; ++k)
                        if (gbc) a += m->Vf[g * Nfp + k] * gbc[((size_t)e * 3 + fc) * Nfp + k];
                    bg[g] = a;
                }
                for (int g = 0; g < Nfq; ++g) {
                    double w = Js * m->gw[g];
                    for (int i = 0; i < Np; ++i) {
                        for (int j = 0; j < Np; ++j)
                            blk[i * Np + j] += w * (-tau * V1[g * Np + i] * V1[g * Np + j]
                                                    + V1[g * Np + i] * Dn1[g * Np + j]
                                                    + Dn1[g * Np + i] * V1[g * Np + j]);
                        b[(size_t)e * Np + i] += w * (-tau * V1[g * Np + i] + Dn1[g * Np + i]) * bg[g];
                    }
                }
            }
            /* other tags: homogeneous Neumann, no face term (zero fixedGradient) */
        }
        /* CSR rows of element e: columns sorted by element index */
        int order[4] = {e, nbl[0], nbl[1], nbl[2]}, slot_of[4] = {0, 1, 2, 3};
        for (int a = 0; a < 4; ++a)
            for (int c = a + 1; c < 4; ++c)
                if (order[c] >= 0 && (order[a] < 0 || order[c] < order[a])) {
                    int t = order[a]; order[a] = order[c]; order[c] = t;
                    t = slot_of[a]; slot_of[a] = slot_of[c]; slot_of[c] = t;
                }
        for (int i = 0; i < Np; ++i) {
            for (int a = 0; a < 4; ++a) {
                if (order[a] < 0) continue;
                const double* B = blk + slot_of[a] * Np * Np;
                for (int j = 0; j < Np; ++j) {
                    col[nz] = order[a] * Np + j;
                    val[nz] = B[i * Np + j];
                    ++nz;
                }
            }
            rowptr[e * Np + i + 1] = nz;
        }
    }
    free(blk); free(V1); free(V2); free(D1x); free(D1y); free(D2x); free(D2y); free(Dn1); free(Dn2);
    free(gx); free(gy); free(bg);
    return nz;
}

int or_csr_matvec(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y)
{
    for (int i = 0; i < n; ++i) {
        double a = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) a += val[k] * x[col[k]];
        y[i] = a;
    }
    return 0;
}

static double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* PCG with point Jacobi, PETSc KSPCG order of operations and KSP_NORM_PRECONDITIONED. */
int or_cg_jacobi(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                 double rtol, double abstol, int maxit, double* rnorm0_out, double* rnorm_out, double* hist)
{
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
    double* dinv = malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
            if (col[k] == i) d = val[k];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    or_csr_matvec(n, rowptr, col, val, x, z);
    for (int i = 0; i < n; ++i) { r[i] = b[i] - z[i]; z[i] = dinv[i] * r[i]; }
    double dp = sqrt(dot(n, z, z));
    double rnorm0 = dp, ttol = fmax(rtol * rnorm0, abstol);
    if (hist) hist[0] = dp;
    int its = 0;
    if (dp > ttol) {
        double beta = dot(n, z, r), betaold = 0.0;
        for (int i = 0; i < maxit; ++i) {
            its = i + 1;
            if (beta == 0.0) break;
            if (i == 0) memcpy(p, z, sizeof(double) * n);
            else {
                double bb = beta / betaold;
                for (int k = 0; k < n; ++k) p[k] = z[k] + bb * p[k];
            }
            betaold = beta;
            or_csr_matvec(n, rowptr, col, val, p, z);
            double dpi = dot(n, p, z);
            double a = beta / dpi;
            for (int k = 0; k < n; ++k) { x[k] += a * p[k]; r[k] -= a * z[k]; }
            for (int k = 0; k < n; ++k) z[k] = dinv[k] * r[k];
            dp = sqrt(dot(n, z, z));
            beta = dot(n, z, r);
            if (hist) hist[i + 1] = dp;
            if (dp <= ttol) break;
        }
    }
    if (rnorm0_out) *rnorm0_out = rnorm0;
    if (rnorm_out) *rnorm_out = dp;
    free(r); free(z); free(p); free(dinv);
    return its;
}
