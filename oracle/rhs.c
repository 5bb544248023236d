/*
 * Oracle explicit right-hand sides and time integration (test infrastructure only).
 *
 * Weak form per element, exactly as the reference assembles it:
 *   b_j = sum_i WJ_i grad(phi_j)(x_i) . F(q(x_i))              (cubature, order 3(N+1))
 *       - sum_faces sum_g faceWJ_g F*_g phi_j(x_g)              (owner -, neighbour +)
 *   rhs = M_e^{-1} b,  M_e = Vq^T diag(WJ) Vq
 * following HopeFOAM-0.1/src/DG/DG/convectionSchemes/defaultConvectionScheme/
 * defaultConvectionScheme.C:47-129 (volume and face scatter),
 * src/DG/DG/gradSchemes/defaultGrad/defaultGrad.C:86-166 (pressure gradient, face flux none),
 * src/DG/DG/divSchemes/defaultDivScheme/defaultDivScheme.C:84-166 (div(U p)),
 * src/DG/fields/dgGaussField/dgGaussField.C:189-269 (cubature / face traces, neighbour
 * via the rotated map, boundary values interpolated from patch nodes),
 * src/DG/DG/simpleFlux/schemes/LFFlux/LFFlux.C:45-100 (LF: maxV over the face),
 * src/DG/DG/godunovFlux/fluxSchemes/scheme/RoeFlux/RoeFlux.C:46-191 (Roe, e = 0),
 * src/DG/Equation/Equation.C:42-79 + dgMesh.C:129-172 (exact block mass solve).
 * The Euler local Lax-Friedrichs flux is not in the reference; it is the textbook
 * Rusanov flux 1/2(F(qM)+F(qP)).n - 1/2 max(|uM.n|+cM, |uP.n|+cP)(qP-qM).
 */
#include <stdlib.h>
#include <tgmath.h>
#include <string.h>
#include "oracle_internal.h"

/* Arithmetic type of the RHS evaluation.  The default build keeps the reference's FP64 arithmetic
 * (and is the CPU baseline); the OR_EXTENDED build evaluates the same algorithm in long double
 * (inputs and outputs stay FP64), i.e. it returns the discretisation's value to well below FP64
 * round-off, which is what the GPU path is compared against. */
#ifdef OR_EXTENDED
typedef long double acc_t;
#else
typedef double acc_t;
#endif
typedef acc_t real;   /* pointwise arithmetic of the RHS evaluation (tgmath picks sqrt/fabs) */

#define ELEM_OWNS_FACE(m, e, f) \
    ((m)->EToE[(e) * 3 + (f)] < 0 || (e) < (m)->EToE[(e) * 3 + (f)] || \
     ((e) == (m)->EToE[(e) * 3 + (f)] && (f) < (m)->EToF[(e) * 3 + (f)]))

/* face trace at the Nfq Gauss points from Nfp face-node values given in the order of vals */
static void face_interp(const or_mesh* m, const real* vals, real* out)
{
    for (int g = 0; g < m->Nfq; ++g) {
        acc_t a = 0.0;
        for (int k = 0; k < m->Nfp; ++k) a += (acc_t)m->Vf[g * m->Nfp + k] * vals[k];
        out[g] = a;
    }
}

/* fixedValue ghost nodal values (FP64 input) interpolated to the face Gauss points */
static void face_interp_bc(const or_mesh* m, const double* vals, real* out)
{
    real tmp[16];
    for (int k = 0; k < m->Nfp; ++k) tmp[k] = vals[k];
    face_interp(m, tmp, out);
}

/* own face nodes of element e, face f (ownerDofMapping) or reversed (neighborDofMapping, rotate 1) */
static void gather_face(const or_mesh* m, const double* q, int e, int f, int reversed, real* out)
{
    const int* fm = m->fmask + f * m->Nfp;
    for (int k = 0; k < m->Nfp; ++k) out[k] = q[(size_t)e * m->Np + fm[reversed ? m->N - k : k]];
}

/* b_e[fmask] += sign * Vf^T t (owner uses the straight map, neighbour the reversed one) */
static void scatter_face(const or_mesh* m, acc_t* b, int e, int f, int reversed, real sign, const real* t)
{
    const int* fm = m->fmask + f * m->Nfp;
    for (int k = 0; k < m->Nfp; ++k) {
        acc_t a = 0.0;
        for (int g = 0; g < m->Nfq; ++g) a += (acc_t)t[g] * m->Vf[g * m->Nfp + k];
        b[(size_t)e * m->Np + fm[reversed ? m->N - k : k]] += sign * a;
    }
}

static void mass_solve(const or_mesh* m, const acc_t* b, double* rhs, int nvar)
{
    int Np = m->Np, K = m->K;
    for (int v = 0; v < nvar; ++v)
        for (int e = 0; e < K; ++e) {
            const acc_t* be = b + ((size_t)v * K + e) * Np;
            double* re = rhs + ((size_t)v * K + e) * Np;
            for (int i = 0; i < Np; ++i) {
                acc_t a = 0.0;
                for (int j = 0; j < Np; ++j) a += (acc_t)m->invM[i * Np + j] * be[j];
                re[i] = (double)(a / (acc_t)m->J[e]);
            }
        }
}

/* volume contribution b_j += sum_i (dphi_j/dx tx_i + dphi_j/dy ty_i) with WJ folded into tx, ty */
static void volume_scatter(const or_mesh* m, int e, const real* tx, const real* ty, acc_t* be)
{
    int Np = m->Np, Nq = m->Nq;
    acc_t rx = m->rx[e], ry = m->ry[e], sx = m->sx[e], sy = m->sy[e];
    for (int i = 0; i < Nq; ++i)
        for (int j = 0; j < Np; ++j) {
            acc_t dr = m->Drq[i * Np + j], ds = m->Dsq[i * Np + j];
            be[j] += (dr * rx + ds * sx) * tx[i] + (dr * ry + ds * sy) * ty[i];
        }
}

/* average != 0: the reference's averageFlux (averageFlux.C:45-90), F* = 1/2 (vO qO + vN qN); else its LFFlux
 * (LFFlux.C:45-100), the same plus 1/2 maxV (qO - qN), maxV the largest |v.n| over the face. */
int or_rhs_advect_flux(const or_mesh* m, const double* q, double ax, double ay, const double* bcq, int average,
                       double* rhs)
{
    int Np = m->Np, Nq = m->Nq, Nfp = m->Nfp, Nfq = m->Nfq, K = m->K;
    acc_t* b = calloc((size_t)K * Np, sizeof(acc_t));
    real *qc = malloc(sizeof(real) * Nq), *tx = malloc(sizeof(real) * Nq), *ty = malloc(sizeof(real) * Nq);
    real *fn = malloc(sizeof(real) * Nfp), *qO = malloc(sizeof(real) * Nfq), *qN = malloc(sizeof(real) * Nfq);
    real* t = malloc(sizeof(real) * Nfq);
    for (int e = 0; e < K; ++e) {
        const double* qe = q + (size_t)e * Np;
        for (int i = 0; i < Nq; ++i) {
            acc_t a = 0.0;
            for (int j = 0; j < Np; ++j) a += (acc_t)m->Vq[i * Np + j] * qe[j];
            qc[i] = a;
            real wj = m->J[e] * m->cw[i];
            tx[i] = ax * qc[i] * wj;
            ty[i] = ay * qc[i] * wj;
        }
        volume_scatter(m, e, tx, ty, b + (size_t)e * Np);
    }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            if (!ELEM_OWNS_FACE(m, e, f)) continue;
            int en = m->EToE[e * 3 + f], fnb = m->EToF[e * 3 + f];
            gather_face(m, q, e, f, 0, fn);
            face_interp(m, fn, qO);
            if (en >= 0) {
                gather_face(m, q, en, fnb, 1, fn);
                face_interp(m, fn, qN);
            } else if (m->bctag[e * 3 + f] == 1 && bcq) {
                face_interp_bc(m, bcq + ((size_t)e * 3 + f) * Nfp, qN);
            } else {
                memcpy(qN, qO, sizeof(real) * Nfq);
            }
            real nx = m->nx[e * 3 + f], ny = m->ny[e * 3 + f];
            real vO = nx * ax + ny * ay, vN = vO;
            real maxV = 0.0;
            for (int g = 0; g < Nfq && !average; ++g) {
                real a = fabs(vO) > fabs(vN) ? fabs(vO) : fabs(vN);
                if (a > maxV) maxV = a;
            }
            for (int g = 0; g < Nfq; ++g) {
                real F = (vO * qO[g] + vN * qN[g]) * 0.5 + maxV * (qO[g] - qN[g]) * 0.5;
                t[g] = m->sJ[e * 3 + f] * m->gw[g] * F;
            }
            scatter_face(m, b, e, f, 0, -1.0, t);
            if (en >= 0) scatter_face(m, b, en, fnb, 1, +1.0, t);
        }
    mass_solve(m, b, rhs, 1);
    free(b); free(qc); free(tx); free(ty); free(fn); free(qO); free(qN); free(t);
    return 0;
}

int or_rhs_advect(const or_mesh* m, const double* q, double ax, double ay, const double* bcq, double* rhs)
{
    return or_rhs_advect_flux(m, q, ax, ay, bcq, 0, rhs);
}

/* Roe flux in the face-normal frame, entropy fix e = 0 (RoeFlux.C:109-181).
 * Inputs are the owner (O) and neighbour (N) states; output F*.n in x/y frame. */
static void roe_flux(real g, real nx, real ny, const real* O, const real* Nb, real* F)
{
    real QP2 = nx * Nb[1] + ny * Nb[2], QM2 = nx * O[1] + ny * O[2];
    real QP3 = nx * Nb[2] - ny * Nb[1], QM3 = nx * O[2] - ny * O[1];
    real rhoP = Nb[0], rhoM = O[0];
    real uM = QM2 / rhoM, uP = QP2 / rhoP, vM = QM3 / rhoM, vP = QP3 / rhoP;
    real pM = (g - 1) * (O[3] - 0.5 * (QM2 * uM + QM3 * vM));
    real pP = (g - 1) * (Nb[3] - 0.5 * (QP2 * uP + QP3 * vP));
    real HM = (O[3] + pM) / rhoM, HP = (Nb[3] + pP) / rhoP;
    real fr = (QM2 + QP2) / 2;
    real fu = (QM2 * uM + pM + QP2 * uP + pP) / 2;
    real fv = (QM3 * uM + QP3 * uP) / 2;
    real fe = (uM * (O[3] + pM) + uP * (Nb[3] + pP)) / 2;
    real rMs = sqrt(rhoM), rPs = sqrt(rhoP);
    real rhob = rMs * rPs;
    real u = (rMs * uM + rPs * uP) / (rMs + rPs);
    real v = (rMs * vM + rPs * vP) / (rMs + rPs);
    real H = (rMs * HM + rPs * HP) / (rMs + rPs);
    real c2 = (g - 1) * (H - 0.5 * (u * u + v * v));
    real c = sqrt(fabs(c2) + 0.0);
    real dw1 = (-0.5 * rhob * (uP - uM) / c + 0.5 * (pP - pM) / c2) * fabs(u - c);
    real dw2 = ((rhoP - rhoM) - (pP - pM) / c2) * fabs(u);
    real dw3 = (rhob * (vP - vM)) * fabs(u);
    real dw4 = (0.5 * rhob * (uP - uM) / c + 0.5 * (pP - pM) / c2) * fabs(u + c);
    fr -= (dw1 + dw2 + dw4) / 2;
    fu -= (dw1 * (u - c) + dw2 * u + dw4 * (u + c)) / 2;
    fv -= (dw1 * v + dw2 * v + dw3 + dw4 * v) / 2;
    fe -= (dw1 * (H - u * c) + dw2 * (u * u + v * v) / 2 + dw3 * v + dw4 * (H + u * c)) / 2;
    F[0] = fr;
    F[1] = nx * fu - ny * fv;
    F[2] = ny * fu + nx * fv;
    F[3] = fe;
}

/* local Lax-Friedrichs (Rusanov) Euler flux, pointwise wave speed. */
static void llf_flux(real g, real nx, real ny, const real* O, const real* Nb, real* F)
{
    real uO = O[1] / O[0], vO = O[2] / O[0], uN = Nb[1] / Nb[0], vN = Nb[2] / Nb[0];
    real pO = (g - 1) * (O[3] - 0.5 * O[0] * (uO * uO + vO * vO));
    real pN = (g - 1) * (Nb[3] - 0.5 * Nb[0] * (uN * uN + vN * vN));
    real unO = uO * nx + vO * ny, unN = uN * nx + vN * ny;
    real cO = sqrt(fabs(g * pO / O[0])), cN = sqrt(fabs(g * pN / Nb[0]));
    real lO = fabs(unO) + cO, lN = fabs(unN) + cN;
    real lam = lO > lN ? lO : lN;
    real FO[4] = {O[0] * unO, O[1] * unO + pO * nx, O[2] * unO + pO * ny, (O[3] + pO) * unO};
    real FN[4] = {Nb[0] * unN, Nb[1] * unN + pN * nx, Nb[2] * unN + pN * ny, (Nb[3] + pN) * unN};
    for (int v = 0; v < 4; ++v) F[v] = 0.5 * (FO[v] + FN[v]) + 0.5 * lam * (O[v] - Nb[v]);
}

int or_rhs_euler(const or_mesh* m, const double* q, double gamma, int flux, const double* bcq, double* rhs)
{
    int Np = m->Np, Nq = m->Nq, Nfp = m->Nfp, Nfq = m->Nfq, K = m->K;
    size_t KN = (size_t)K * Np;
    acc_t* b = calloc(4 * KN, sizeof(acc_t));
    real *qc = malloc(sizeof(real) * 4 * Nq), *tx = malloc(sizeof(real) * Nq), *ty = malloc(sizeof(real) * Nq);
    real* fnode = malloc(sizeof(real) * Nfp);
    real *qO = malloc(sizeof(real) * 4 * Nfq), *qN = malloc(sizeof(real) * 4 * Nfq);
    real* t = malloc(sizeof(real) * 4 * Nfq);
    for (int e = 0; e < K; ++e) {
        for (int v = 0; v < 4; ++v) {
            const double* qe = q + v * KN + (size_t)e * Np;
            for (int i = 0; i < Nq; ++i) {
                acc_t a = 0.0;
                for (int j = 0; j < Np; ++j) a += (acc_t)m->Vq[i * Np + j] * qe[j];
                qc[v * Nq + i] = a;
            }
        }
        real *rho = qc, *ru = qc + Nq, *rv = qc + 2 * Nq, *E = qc + 3 * Nq;
        acc_t *brho = b + (size_t)e * Np, *bru = b + KN + (size_t)e * Np, *brv = b + 2 * KN + (size_t)e * Np,
              *bE = b + 3 * KN + (size_t)e * Np;
        /* div(gther_U, rho1): U_x rho WJ, U_y rho WJ */
        for (int i = 0; i < Nq; ++i) {
            real wj = m->J[e] * m->cw[i], Ux = ru[i] / rho[i], Uy = rv[i] / rho[i];
            tx[i] = Ux * rho[i] * wj; ty[i] = Uy * rho[i] * wj;
        }
        volume_scatter(m, e, tx, ty, brho);
        /* div(gther_U, rhoU1) + grad(gther_p) */
        for (int i = 0; i < Nq; ++i) {
            real wj = m->J[e] * m->cw[i], Ux = ru[i] / rho[i], Uy = rv[i] / rho[i];
            real p = (gamma - 1) * (E[i] - 0.5 * (rho[i] * (Ux * Ux + Uy * Uy)));
            tx[i] = Ux * ru[i] * wj + p * wj; ty[i] = Uy * ru[i] * wj;
        }
        volume_scatter(m, e, tx, ty, bru);
        for (int i = 0; i < Nq; ++i) {
            real wj = m->J[e] * m->cw[i], Ux = ru[i] / rho[i], Uy = rv[i] / rho[i];
            real p = (gamma - 1) * (E[i] - 0.5 * (rho[i] * (Ux * Ux + Uy * Uy)));
            tx[i] = Ux * rv[i] * wj; ty[i] = Uy * rv[i] * wj + p * wj;
        }
        volume_scatter(m, e, tx, ty, brv);
        /* div(gther_U, Ener1) + div(gther_U, gther_p) */
        for (int i = 0; i < Nq; ++i) {
            real wj = m->J[e] * m->cw[i], Ux = ru[i] / rho[i], Uy = rv[i] / rho[i];
            real p = (gamma - 1) * (E[i] - 0.5 * (rho[i] * (Ux * Ux + Uy * Uy)));
            tx[i] = Ux * E[i] * wj + Ux * p * wj; ty[i] = Uy * E[i] * wj + Uy * p * wj;
        }
        volume_scatter(m, e, tx, ty, bE);
    }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            if (!ELEM_OWNS_FACE(m, e, f)) continue;
            int en = m->EToE[e * 3 + f], fnb = m->EToF[e * 3 + f], tag = m->bctag[e * 3 + f];
            real nx = m->nx[e * 3 + f], ny = m->ny[e * 3 + f];
            for (int v = 0; v < 4; ++v) {
                gather_face(m, q + v * KN, e, f, 0, fnode);
                face_interp(m, fnode, qO + v * Nfq);
                if (en >= 0) {
                    gather_face(m, q + v * KN, en, fnb, 1, fnode);
                    face_interp(m, fnode, qN + v * Nfq);
                } else if (tag == 1 && bcq) {
                    face_interp_bc(m, bcq + ((size_t)v * K * 3 + (size_t)e * 3 + f) * Nfp, qN + v * Nfq);
                } else {
                    memcpy(qN + v * Nfq, qO + v * Nfq, sizeof(real) * Nfq);
                }
            }
            if (en < 0 && tag == 3) {     /* reflective: mirror momentum (reflectiveDgPatchField.C:111-154) */
                for (int g = 0; g < Nfq; ++g) {
                    real mu = qN[Nfq + g], mv = qN[2 * Nfq + g], mn = mu * nx + mv * ny;
                    qN[Nfq + g] = mu - 2 * mn * nx;
                    qN[2 * Nfq + g] = mv - 2 * mn * ny;
                }
            }
            for (int g = 0; g < Nfq; ++g) {
                real O[4] = {qO[g], qO[Nfq + g], qO[2 * Nfq + g], qO[3 * Nfq + g]};
                real Nb[4] = {qN[g], qN[Nfq + g], qN[2 * Nfq + g], qN[3 * Nfq + g]};
                real F[4];
                if (flux == 0) roe_flux(gamma, nx, ny, O, Nb, F);
                else llf_flux(gamma, nx, ny, O, Nb, F);
                real wj = m->sJ[e * 3 + f] * m->gw[g];
                for (int v = 0; v < 4; ++v) t[v * Nfq + g] = wj * F[v];
            }
            for (int v = 0; v < 4; ++v) {
                scatter_face(m, b + v * KN, e, f, 0, -1.0, t + v * Nfq);
                if (en >= 0) scatter_face(m, b + v * KN, en, fnb, 1, +1.0, t + v * Nfq);
            }
        }
    mass_solve(m, b, rhs, 4);
    free(b); free(qc); free(tx); free(ty); free(fnode); free(qO); free(qN); free(t);
    return 0;
}

/* Carpenter-Kennedy (2N-storage, 5 stages, order 4), createFields.H:119-138. */
void or_lserk4_coeffs(double* a, double* b, double* c)
{
    const double A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                         -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
    const double B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};
    const double C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                         2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};
    for (int i = 0; i < 5; ++i) { a[i] = A[i]; b[i] = B[i]; c[i] = C[i]; }
}

static int rhs_eval(const or_mesh* m, int eq, const double* q, double a0, double a1, int flux,
                    const double* bcq, double* rhs)
{
    if (eq == 0) return or_rhs_advect(m, q, a0, a1, bcq, rhs);
    return or_rhs_euler(m, q, a0, flux, bcq, rhs);
}

int or_lserk4(const or_mesh* m, int eq, double* q, int nsteps, double dt, double a0, double a1, int flux,
              const double* bcq)
{
    size_t n = (size_t)m->K * m->Np * (eq == 0 ? 1 : 4);
    double *res = calloc(n, sizeof(double)), *rhs = malloc(sizeof(double) * n);
    double A[5], B[5], C[5];
    or_lserk4_coeffs(A, B, C);
    for (int step = 0; step < nsteps; ++step)
        for (int s = 0; s < 5; ++s) {
            rhs_eval(m, eq, q, a0, a1, flux, bcq, rhs);
            for (size_t i = 0; i < n; ++i) {
                res[i] = A[s] * res[i] + dt * rhs[i];
                q[i] += B[s] * res[i];
            }
        }
    free(res); free(rhs);
    return 0;
}

/* SSP-RK2 as dgEulerFoam.C:64-123: two forward-Euler solves then the 1/2-1/2 average. */
int or_ssprk2(const or_mesh* m, int eq, double* q, int nsteps, double dt, double a0, double a1, int flux,
              const double* bcq)
{
    size_t n = (size_t)m->K * m->Np * (eq == 0 ? 1 : 4);
    double *q1 = malloc(sizeof(double) * n), *rhs = malloc(sizeof(double) * n);
    for (int step = 0; step < nsteps; ++step) {
        memcpy(q1, q, sizeof(double) * n);
        for (int stage = 0; stage < 2; ++stage) {
            rhs_eval(m, eq, q1, a0, a1, flux, bcq, rhs);
            for (size_t i = 0; i < n; ++i) q1[i] += dt * rhs[i];
        }
        for (size_t i = 0; i < n; ++i) q[i] = 0.5 * q[i] + 0.5 * q1[i];
    }
    free(q1); free(rhs);
    return 0;
}

/* Isentropic vortex of setNonUniformInlet.H / setBoundaryValues.H / eulerError.H:
 * r = |x - x0 - (t,0)|^2 (squared distance), free stream (1,0). */
void or_vortex_exact(int n, const double* x, const double* y, double t, double x0, double y0,
                     double beta, double gamma, double* rho, double* rhou, double* rhov, double* E)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double dx = x[i] - x0 - t, dy = y[i] - y0;
        double r = dx * dx + dy * dy;
        double d = pow(1.0 - (gamma - 1.0) * (beta * beta) * exp(2.0 * (1.0 - r)) / (16.0 * gamma * pi * pi),
                       1.0 / (gamma - 1.0));
        double u = 1 - beta * exp(1 - r) * dy / (2.0 * pi);
        double v = beta * exp(1 - r) * dx / (2.0 * pi);
        rho[i] = d;
        rhou[i] = u * d;
        rhov[i] = v * d;
        E[i] = pow(d, gamma) / (gamma - 1.0) + 0.5 * (rhou[i] * rhou[i] + rhov[i] * rhov[i]) / d;
    }
}

/* dgEulerFoam.C:46-131 on the vortex case: exact IC, fixedValue exact BC values set at t_n for
 * both SSP-RK2 stages (setBoundaryValues.H is called with runTime - deltaT), Roe flux, and the
 * mean-absolute errors of eulerError.H:37-38 at the final time. */
int or_vortex_ref_run(const or_mesh* m, double dt, int nsteps, double* err_rho, double* err_rhou)
{
    const double gamma = 1.4, beta = 5.0, x0 = 5.0, y0 = 0.0;
    int K = m->K, Np = m->Np, Nfp = m->Nfp;
    size_t KN = (size_t)K * Np;
    double* q = malloc(sizeof(double) * 4 * KN);
    or_vortex_exact((int)KN, m->x, m->y, 0.0, x0, y0, beta, gamma, q, q + KN, q + 2 * KN, q + 3 * KN);
    double* bcq = calloc((size_t)4 * K * 3 * Nfp, sizeof(double));
    double *bx = malloc(sizeof(double) * Nfp), *by = malloc(sizeof(double) * Nfp), *bv = malloc(sizeof(double) * 4 * Nfp);
    double *q1 = malloc(sizeof(double) * 4 * KN), *rhs = malloc(sizeof(double) * 4 * KN);
    for (int step = 0; step < nsteps; ++step) {
        double tn = step * dt;
        for (int e = 0; e < K; ++e)
            for (int f = 0; f < 3; ++f) {
                if (m->EToE[e * 3 + f] >= 0) continue;
                for (int k = 0; k < Nfp; ++k) {
                    int node = m->fmask[f * Nfp + k];
                    bx[k] = m->x[(size_t)e * Np + node];
                    by[k] = m->y[(size_t)e * Np + node];
                }
                or_vortex_exact(Nfp, bx, by, tn, x0, y0, beta, gamma, bv, bv + Nfp, bv + 2 * Nfp, bv + 3 * Nfp);
                for (int v = 0; v < 4; ++v)
                    memcpy(bcq + ((size_t)v * K * 3 + (size_t)e * 3 + f) * Nfp, bv + v * Nfp, sizeof(double) * Nfp);
            }
        memcpy(q1, q, sizeof(double) * 4 * KN);
        for (int stage = 0; stage < 2; ++stage) {
            or_rhs_euler(m, q1, gamma, 0, bcq, rhs);
            for (size_t i = 0; i < 4 * KN; ++i) q1[i] += dt * rhs[i];
        }
        for (size_t i = 0; i < 4 * KN; ++i) q[i] = 0.5 * q[i] + 0.5 * q1[i];
    }
    double* ex = malloc(sizeof(double) * 4 * KN);
    or_vortex_exact((int)KN, m->x, m->y, nsteps * dt, x0, y0, beta, gamma, ex, ex + KN, ex + 2 * KN, ex + 3 * KN);
    double er = 0.0, eu = 0.0;
    for (size_t i = 0; i < KN; ++i) {
        er += fabs(ex[i] - q[i]);
        double du = ex[KN + i] - q[KN + i], dv = ex[2 * KN + i] - q[2 * KN + i];
        eu += sqrt(du * du + dv * dv);
    }
    *err_rho = er / (double)KN;
    *err_rhou = eu / (double)KN;
    free(q); free(bcq); free(bx); free(by); free(bv); free(q1); free(rhs); free(ex);
    return 0;
}
