/*
 * Reference-triangle operators for the oracle (test infrastructure only).
 *
 * Restates HopeFOAM-0.1/src/DG/element/polynomials/Legendre/Legendre.C:39-540
 * (Jacobi polynomials, Gauss rules, Vandermonde matrices),
 * src/DG/element/baseFunctions/straightBaseFunctions/triangleBaseFunction/
 * triangleBaseFunction.C:46-276 (warp & blend nodes, Dr, face maps) and
 * src/DG/element/gaussIntegration/gaussTriangleIntegration/gaussTriangleIntegration.C:45-97
 * (cubature interpolation / derivative matrices, face Gauss rule).
 *
 * The Gauss rule is computed by the Golub-Welsch method as in Legendre.C:39-157,
 * with a cyclic Jacobi eigen-solver in place of SLEPc's EPS (both return the
 * eigenpairs of the same symmetric tridiagonal matrix), then sorted and
 * weight-symmetrised like Legendre.C:134-150.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "dgref.h"
#include "cubature_tri_table.h"

/* Gamma(x) for integer x >= 1 (Legendre.C:209-213). */
static double gamma_int(int x)
{
    double a = 1.0;
    for (int i = 2; i < x; ++i) a *= i;
    return a;
}

/* Orthonormal Jacobi polynomial P_n^{(alpha,beta)}(x) (Legendre.C:174-205). */
double or_jacobi_p(double x, double alpha, double beta, int n)
{
    double g0 = pow(2.0, alpha + beta + 1) / (alpha + beta + 1) * gamma_int((int)alpha + 1)
              * gamma_int((int)beta + 1) / gamma_int((int)(alpha + beta) + 1);
    double g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0;
    double p0 = 1.0 / sqrt(g0);
    if (n == 0) return p0;
    double p1 = ((alpha + beta + 2) / 2 * x + (alpha - beta) / 2) / sqrt(g1);
    if (n == 1) return p1;
    double aold = 2 / (2 + alpha + beta) * sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3));
    double pm = p0, pc = p1;
    for (int i = 1; i < n; ++i) {
        double h1 = 2.0 * i + alpha + beta;
        double anew = 2 / (h1 + 2) * sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha)
                                          * (i + 1 + beta) / (h1 + 1) / (h1 + 3));
        double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2);
        double pn = 1 / anew * (-aold * pm + (x - bnew) * pc);
        pm = pc;
        pc = pn;
        aold = anew;
    }
    return pc;
}

/* d/dx P_n^{(alpha,beta)} (Legendre.C:207-216). */
double or_grad_jacobi_p(double x, double alpha, double beta, int n)
{
    if (n == 0) return 0.0;
    return sqrt(n * (n + alpha + beta + 1)) * or_jacobi_p(x, alpha + 1, beta + 1, n - 1);
}

/* Symmetric eigen-decomposition by cyclic Jacobi rotations.  A (n x n, row-major)
 * is destroyed; eigenvalues in d, eigenvectors in the columns of Z. */
static void sym_eig(int n, double* A, double* d, double* Z)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {           /* A <- A J */
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {           /* A <- J^T A */
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double zkp = Z[k * n + p], zkq = Z[k * n + q];
                    Z[k * n + p] = c * zkp - s * zkq;
                    Z[k * n + q] = s * zkp + c * zkq;
                }
            }
        }
    }
    for (int i = 0; i < n; ++i) d[i] = A[i * n + i];
}

/* Gauss-Jacobi rule with n+1 points (Legendre.C:39-157). */
int or_jacobi_gq(double alpha, double beta, int n, double* x, double* w)
{
    if (n == 0) {
        x[0] = -(alpha - beta) / (alpha + beta + 2);
        w[0] = 2;
        return 1;
    }
    int m = n + 1;
    double* A = calloc((size_t)m * m, sizeof(double));
    double* Z = calloc((size_t)m * m, sizeof(double));
    double* d = calloc((size_t)m, sizeof(double));
    for (int i = 0; i < n; ++i) {
        int j = i + 1;
        double h = 2.0 * i + alpha + beta;
        double v = 2 / (h + 2) * sqrt(j * (j + alpha + beta) * (j + alpha) * (j + beta) / (h + 1) / (h + 3));
        A[i * m + j] = v;
        A[j * m + i] = v;
    }
    /* diagonal: -(a^2-b^2)/((h)(h+2)); zero for the symmetric rules the reference uses */
    for (int i = 0; i < m; ++i) {
        double h = 2.0 * i + alpha + beta;
        A[i * m + i] = (alpha + beta == 0.0 || alpha == beta) ? 0.0 : -(alpha * alpha - beta * beta) / h / (h + 2);
    }
    sym_eig(m, A, d, Z);
    double c = pow(2.0, alpha + beta + 1) / (alpha + beta + 1) * gamma_int((int)alpha + 1)
             * gamma_int((int)beta + 1) / gamma_int((int)(alpha + beta) + 1);
    for (int i = 0; i < m; ++i) {
        x[i] = d[i];
        w[i] = Z[0 * m + i] * Z[0 * m + i] * c;
    }
    for (int i = 0; i < m; ++i)          /* sort ascending (Legendre.C:134-145) */
        for (int j = i + 1; j < m; ++j)
            if (x[i] > x[j]) {
                double t = x[i]; x[i] = x[j]; x[j] = t;
                t = w[i]; w[i] = w[j]; w[j] = t;
            }
    for (int i = 0; i <= (n - 1) / 2; ++i) {   /* symmetrise weights (Legendre.C:146-149) */
        double t = 0.5 * (w[i] + w[n - i]);
        w[i] = w[n - i] = t;
    }
    free(A); free(Z); free(d);
    return m;
}

/* Gauss-Lobatto rule nodes (Legendre.C:160-172). */
int or_jacobi_gl(double alpha, double beta, int n, double* x)
{
    for (int i = 0; i <= n; ++i) x[i] = 0;
    x[0] = -1.0;
    x[n] = 1.0;
    if (n < 2) return n + 1;
    double* xi = malloc(sizeof(double) * (n - 1));
    double* wi = malloc(sizeof(double) * (n - 1));
    or_jacobi_gq(alpha + 1, beta + 1, n - 2, xi, wi);
    for (int i = 1; i < n; ++i) x[i] = xi[i - 1];
    free(xi); free(wi);
    return n + 1;
}

int or_np(int N) { return (N + 1) * (N + 2) / 2; }

int or_nq(int N)
{
    int order = 3 * (N + 1);
    if (order > or_cub_max_order) return -1;
    return or_cub_npts[order];
}

/* Gauss-Jordan inverse with partial pivoting. */
int or_invert(int n, const double* A, double* Ainv)
{
    double* a = malloc(sizeof(double) * n * n);
    memcpy(a, A, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ainv[i * n + j] = (i == j);
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[r * n + c]) > fabs(a[p * n + c])) p = r;
        if (a[p * n + c] == 0.0) { free(a); return -1; }
        if (p != c)
            for (int k = 0; k < n; ++k) {
                double t = a[c * n + k]; a[c * n + k] = a[p * n + k]; a[p * n + k] = t;
                t = Ainv[c * n + k]; Ainv[c * n + k] = Ainv[p * n + k]; Ainv[p * n + k] = t;
            }
        double inv = 1.0 / a[c * n + c];
        for (int k = 0; k < n; ++k) { a[c * n + k] *= inv; Ainv[c * n + k] *= inv; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = a[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[r * n + k] -= f * a[c * n + k];
                Ainv[r * n + k] -= f * Ainv[c * n + k];
            }
        }
    }
    free(a);
    return 0;
}

/* 1D Vandermonde V[i][j] = P_j(r_i) (Legendre.C:216-230). */
static void vandermonde1d(int N, int m, const double* r, double* V)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= N; ++j) V[i * (N + 1) + j] = or_jacobi_p(r[i], 0, 0, j);
}

/* warpFactor (triangleBaseFunction.C:189-234). */
static void warp_factor(int N, int m, const double* rout, double* warp)
{
    double* lgl = malloc(sizeof(double) * (N + 1));
    double* req = malloc(sizeof(double) * (N + 1));
    double* Veq = malloc(sizeof(double) * (N + 1) * (N + 1));
    double* VeqT = malloc(sizeof(double) * (N + 1) * (N + 1));
    double* invVeqT = malloc(sizeof(double) * (N + 1) * (N + 1));
    or_jacobi_gl(0, 0, N, lgl);
    for (int i = 0; i <= N; ++i) req[i] = 2.0 / N * i - 1;
    vandermonde1d(N, N + 1, req, Veq);
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) VeqT[i * (N + 1) + j] = Veq[j * (N + 1) + i];
    or_invert(N + 1, VeqT, invVeqT);
    for (int k = 0; k < m; ++k) {
        /* Lmat[:,k] = inv(Veq') * P(rout_k); warp_k = sum_i Lmat[i,k] (lgl_i - req_i) */
        double w = 0.0;
        for (int i = 0; i <= N; ++i) {
            double l = 0.0;
            for (int j = 0; j <= N; ++j) l += invVeqT[i * (N + 1) + j] * or_jacobi_p(rout[k], 0, 0, j);
            w += l * (lgl[i] - req[i]);
        }
        double zerof = fabs(rout[k]) < (1.0 - 1.0e-10) ? 1.0 : 0.0;
        double sf = 1.0 - (zerof * rout[k]) * (zerof * rout[k]);
        warp[k] = w / sf + w * (zerof - 1.0);
    }
    free(lgl); free(req); free(Veq); free(VeqT); free(invVeqT);
}

/* Warp & blend nodes on the reference triangle (triangleBaseFunction.C:122-186). */
int or_nodes2d(int N, double* r, double* s)
{
    static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999,
                                      1.2832, 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
    int Np = or_np(N);
    double alpha = N < 16 ? alpopt[N - 1] : 5.0 / 3.0;
    double *L1 = malloc(sizeof(double) * Np), *L2 = malloc(sizeof(double) * Np), *L3 = malloc(sizeof(double) * Np);
    double *x = malloc(sizeof(double) * Np), *y = malloc(sizeof(double) * Np);
    double *t1 = malloc(sizeof(double) * Np), *t2 = malloc(sizeof(double) * Np), *t3 = malloc(sizeof(double) * Np);
    double *w1 = malloc(sizeof(double) * Np), *w2 = malloc(sizeof(double) * Np), *w3 = malloc(sizeof(double) * Np);
    int sk = 0;
    for (int n = 1; n <= N + 1; ++n)
        for (int mm = 1; mm <= N + 2 - n; ++mm) {
            L1[sk] = (double)(n - 1) / N;
            L3[sk] = (double)(mm - 1) / N;
            ++sk;
        }
    for (int i = 0; i < Np; ++i) {
        L2[i] = 1.0 - L1[i] - L3[i];
        x[i] = L3[i] - L2[i];
        y[i] = (2 * L1[i] - L2[i] - L3[i]) / sqrt(3.0);
        t1[i] = L3[i] - L2[i];
        t2[i] = L1[i] - L3[i];
        t3[i] = L2[i] - L1[i];
    }
    warp_factor(N, Np, t1, w1);
    warp_factor(N, Np, t2, w2);
    warp_factor(N, Np, t3, w3);
    const double pi = 3.14159265358979323846;
    double cos1 = cos(2 * pi / 3), cos2 = cos(4 * pi / 3), sin1 = sin(2 * pi / 3), sin2 = sin(4 * pi / 3);
    for (int i = 0; i < Np; ++i) {
        double b1 = 4 * L2[i] * L3[i], b2 = 4 * L1[i] * L3[i], b3 = 4 * L1[i] * L2[i];
        double a1 = alpha * L1[i], a2 = alpha * L2[i], a3 = alpha * L3[i];
        double wa = b1 * w1[i] * (1 + a1 * a1);
        double wb = b2 * w2[i] * (1 + a2 * a2);
        double wc = b3 * w3[i] * (1 + a3 * a3);
        x[i] += 1 * wa + cos1 * wb + cos2 * wc;
        y[i] += 0 * wa + sin1 * wb + sin2 * wc;
    }
    for (int i = 0; i < Np; ++i) {   /* equilateral (x,y) -> (r,s) */
        double l1 = (sqrt(3.0) * y[i] + 1.0) / 3.0;
        double l2 = (-3.0 * x[i] - sqrt(3.0) * y[i] + 2.0) / 6.0;
        double l3 = (3.0 * x[i] - sqrt(3.0) * y[i] + 2.0) / 6.0;
        r[i] = l3 - l2 - l1;
        s[i] = l1 - l2 - l3;
    }
    free(L1); free(L2); free(L3); free(x); free(y); free(t1); free(t2); free(t3);
    free(w1); free(w2); free(w3);
    return Np;
}

/* Vandermonde2D (Legendre.C:246-275). */
int or_vandermonde2d(int N, int m, const double* r, const double* s, double* V)
{
    int Np = or_np(N);
    for (int n = 0; n < m; ++n) {
        double a = s[n] != 1.0 ? 2 * (1 + r[n]) / (1 - s[n]) - 1 : -1.0;
        double b = s[n];
        int sk = 0;
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N - i; ++j) {
                V[n * Np + sk] = sqrt(2.0) * or_jacobi_p(a, 0, 0, i) * or_jacobi_p(b, 2 * i + 1, 0, j)
                               * pow(1 - b, i);
                ++sk;
            }
    }
    return 0;
}

/* GradVandermonde2D (Legendre.C:371-438). */
int or_grad_vandermonde2d(int N, int m, const double* r, const double* s, double* Vr, double* Vs)
{
    int Np = or_np(N);
    for (int n = 0; n < m; ++n) {
        double a = s[n] != 1.0 ? 2 * (1 + r[n]) / (1 - s[n]) - 1 : -1.0;
        double b = s[n];
        int sk = 0;
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N - i; ++j) {
                double fa = or_jacobi_p(a, 0, 0, i), gb = or_jacobi_p(b, 2 * i + 1, 0, j);
                double dfa = or_grad_jacobi_p(a, 0, 0, i), dgb = or_grad_jacobi_p(b, 2 * i + 1, 0, j);
                double c = pow(2.0, i + 0.5);
                double vr, vs;
                if (i > 0) {
                    vr = dfa * gb * c * pow(0.5 * (1.0 - b), i - 1);
                    double t = dgb * pow(0.5 * (1 - b), i) - 0.5 * i * gb * pow(0.5 * (1 - b), i - 1);
                    vs = (dfa * gb * 0.5 * (1 + a) * pow(0.5 * (1 - b), i - 1) + fa * t) * c;
                } else {
                    vr = dfa * gb * c;
                    double t = dgb * pow(0.5 * (1 - b), i);
                    vs = (dfa * gb * 0.5 * (1 + a) + fa * t) * c;
                }
                Vr[n * Np + sk] = vr;
                Vs[n * Np + sk] = vs;
                ++sk;
            }
    }
    return 0;
}

int or_cubature(int order, double* r, double* s, double* w)
{
    if (order < 1 || order > or_cub_max_order) return -1;
    int n = or_cub_npts[order];
    memcpy(r, or_cub_r[order], sizeof(double) * n);
    memcpy(s, or_cub_s[order], sizeof(double) * n);
    memcpy(w, or_cub_w[order], sizeof(double) * n);
    return n;
}

/* faceToCellIndex_[f][0] (triangleBaseFunction.C:75-94): face f runs from vertex f to f+1. */
int or_fmask(int N, int* fm)
{
    int Nfp = N + 1, Np = or_np(N);
    for (int i = 0; i < Nfp; ++i) fm[0 * Nfp + i] = i;
    fm[1 * Nfp + 0] = N;
    fm[2 * Nfp + 0] = Np - 1;
    for (int i = 1; i < Nfp; ++i) {
        fm[1 * Nfp + i] = fm[1 * Nfp + i - 1] + (Nfp - i);
        fm[2 * Nfp + i] = fm[2 * Nfp + i - 1] - i - 1;
    }
    return 0;
}

static void matmul(int m, int k, int n, const double* A, const double* B, double* C)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int l = 0; l < k; ++l) acc += A[i * k + l] * B[l * n + j];
            C[i * n + j] = acc;
        }
}

int or_refelem_ops(int N, double* r, double* s, double* V, double* invV, double* Dr, double* Ds,
                   double* Vq, double* Drq, double* Dsq, double* cw, double* Vf, double* gw, double* M)
{
    int Np = or_np(N), Nq = or_nq(N), Nfp = N + 1, Nfq = N + 2;
    if (Nq < 0) return -1;
    or_nodes2d(N, r, s);
    or_vandermonde2d(N, Np, r, s, V);
    if (or_invert(Np, V, invV)) return -2;
    double* Vr = malloc(sizeof(double) * Np * Np);
    double* Vs = malloc(sizeof(double) * Np * Np);
    or_grad_vandermonde2d(N, Np, r, s, Vr, Vs);
    matmul(Np, Np, Np, Vr, invV, Dr);
    matmul(Np, Np, Np, Vs, invV, Ds);
    free(Vr); free(Vs);
    /* cubature (gaussTriangleIntegration.C:49-78) */
    double* cr = malloc(sizeof(double) * Nq);
    double* cs = malloc(sizeof(double) * Nq);
    or_cubature(3 * (N + 1), cr, cs, cw);
    double* Vc = malloc(sizeof(double) * Nq * Np);
    double* Vcr = malloc(sizeof(double) * Nq * Np);
    double* Vcs = malloc(sizeof(double) * Nq * Np);
    or_vandermonde2d(N, Nq, cr, cs, Vc);
    or_grad_vandermonde2d(N, Nq, cr, cs, Vcr, Vcs);
    matmul(Nq, Np, Np, Vc, invV, Vq);
    matmul(Nq, Np, Np, Vcr, invV, Drq);
    matmul(Nq, Np, Np, Vcs, invV, Dsq);
    /* face rule: JacobiGQ(0,0,faceIntOrder/2) with faceIntOrder = 2(N+1) (gaussTriangleIntegration.C:80-94) */
    double* gr = malloc(sizeof(double) * Nfq);
    or_jacobi_gq(0, 0, N + 1, gr, gw);
    double* gl = malloc(sizeof(double) * Nfp);
    or_jacobi_gl(0, 0, N, gl);
    double* V1 = malloc(sizeof(double) * Nfp * Nfp);
    double* iV1 = malloc(sizeof(double) * Nfp * Nfp);
    double* Vg = malloc(sizeof(double) * Nfq * Nfp);
    vandermonde1d(N, Nfp, gl, V1);
    or_invert(Nfp, V1, iV1);
    vandermonde1d(N, Nfq, gr, Vg);
    matmul(Nfq, Nfp, Nfp, Vg, iV1, Vf);
    /* cubature mass matrix with unit Jacobian (physicalCellElement.C:99-108 with WJ = w) */
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            double acc = 0.0;
            for (int q = 0; q < Nq; ++q) acc += Vq[q * Np + i] * cw[q] * Vq[q * Np + j];
            M[i * Np + j] = acc;
        }
    free(cr); free(cs); free(Vc); free(Vcr); free(Vcs); free(gr); free(gl); free(V1); free(iV1); free(Vg);
    return 0;
}
