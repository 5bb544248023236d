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
 * weight-symmetrised like Legendre.C:134-150 (nodes are symmetrised as well).
 */
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#include "dgref.h"
#include "cubature_tri_table.h"

/* All operators are built in long double and rounded once to FP64: they are then the correctly
 * rounded exact operators (within an ulp) whatever order of operations produced them. */
typedef long double ld;

/* Gamma(x) for integer x >= 1 (Legendre.C:209-213). */
static ld gamma_int(int x)
{
    ld a = 1.0L;
    for (int i = 2; i < x; ++i) a *= i;
    return a;
}

/* Orthonormal Jacobi polynomial P_n^{(alpha,beta)}(x) (Legendre.C:174-205). */
static ld jacobi_p_ld(ld x, ld alpha, ld beta, int n)
{
    ld g0 = pow(2.0L, alpha + beta + 1) / (alpha + beta + 1) * gamma_int((int)alpha + 1)
          * gamma_int((int)beta + 1) / gamma_int((int)(alpha + beta) + 1);
    ld g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0;
    ld p0 = 1.0L / sqrt(g0);
    if (n == 0) return p0;
    ld p1 = ((alpha + beta + 2) / 2 * x + (alpha - beta) / 2) / sqrt(g1);
    if (n == 1) return p1;
    ld aold = 2 / (2 + alpha + beta) * sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3));
    ld pm = p0, pc = p1;
    for (int i = 1; i < n; ++i) {
        ld h1 = 2.0L * i + alpha + beta;
        ld anew = 2 / (h1 + 2) * sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha)
                                      * (i + 1 + beta) / (h1 + 1) / (h1 + 3));
        ld bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2);
        ld pn = 1 / anew * (-aold * pm + (x - bnew) * pc);
        pm = pc;
        pc = pn;
        aold = anew;
    }
    return pc;
}

/* d/dx P_n^{(alpha,beta)} (Legendre.C:207-216). */
static ld grad_jacobi_p_ld(ld x, ld alpha, ld beta, int n)
{
    if (n == 0) return 0.0L;
    return sqrt(n * (n + alpha + beta + 1)) * jacobi_p_ld(x, alpha + 1, beta + 1, n - 1);
}

double or_jacobi_p(double x, double alpha, double beta, int n) { return (double)jacobi_p_ld(x, alpha, beta, n); }
double or_grad_jacobi_p(double x, double alpha, double beta, int n)
{
    return (double)grad_jacobi_p_ld(x, alpha, beta, n);
}

/* Symmetric eigen-decomposition by cyclic Jacobi rotations.  A (n x n, row-major)
 * is destroyed; eigenvalues in d, eigenvectors in the columns of Z. */
static void sym_eig(int n, ld* A, ld* d, ld* Z)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        ld off = 0.0L;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-300L) break;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                ld apq = A[p * n + q];
                if (fabs(apq) < 1e-300L) continue;
                ld app = A[p * n + p], aqq = A[q * n + q];
                ld theta = (aqq - app) / (2 * apq);
                ld t = (theta >= 0 ? 1.0L : -1.0L) / (fabs(theta) + sqrt(theta * theta + 1));
                ld c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {           /* A <- A J */
                    ld akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {           /* A <- J^T A */
                    ld apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    ld zkp = Z[k * n + p], zkq = Z[k * n + q];
                    Z[k * n + p] = c * zkp - s * zkq;
                    Z[k * n + q] = s * zkp + c * zkq;
                }
            }
        }
    }
    for (int i = 0; i < n; ++i) d[i] = A[i * n + i];
}

/* Gauss-Jacobi rule with n+1 points (Legendre.C:39-157). */
static int jacobi_gq_ld(ld alpha, ld beta, int n, ld* x, ld* w)
{
    if (n == 0) {
        x[0] = -(alpha - beta) / (alpha + beta + 2);
        w[0] = 2;
        return 1;
    }
    int m = n + 1;
    ld* A = calloc((size_t)m * m, sizeof(ld));
    ld* Z = calloc((size_t)m * m, sizeof(ld));
    ld* d = calloc((size_t)m, sizeof(ld));
    for (int i = 0; i < n; ++i) {
        int j = i + 1;
        ld h = 2.0L * i + alpha + beta;
        ld v = 2 / (h + 2) * sqrt(j * (j + alpha + beta) * (j + alpha) * (j + beta) / (h + 1) / (h + 3));
        A[i * m + j] = v;
        A[j * m + i] = v;
    }
    /* diagonal: -(a^2-b^2)/((h)(h+2)); zero for the symmetric rules the reference uses */
    for (int i = 0; i < m; ++i) {
        ld h = 2.0L * i + alpha + beta;
        A[i * m + i] = (alpha == beta) ? 0.0L : -(alpha * alpha - beta * beta) / h / (h + 2);
    }
    sym_eig(m, A, d, Z);
    ld c = pow(2.0L, alpha + beta + 1) / (alpha + beta + 1) * gamma_int((int)alpha + 1)
         * gamma_int((int)beta + 1) / gamma_int((int)(alpha + beta) + 1);
    for (int i = 0; i < m; ++i) {
        x[i] = d[i];
        w[i] = Z[0 * m + i] * Z[0 * m + i] * c;
    }
    for (int i = 0; i < m; ++i)          /* sort ascending (Legendre.C:134-145) */
        for (int j = i + 1; j < m; ++j)
            if (x[i] > x[j]) {
                ld t = x[i]; x[i] = x[j]; x[j] = t;
                t = w[i]; w[i] = w[j]; w[j] = t;
            }
    for (int i = 0; i <= (n - 1) / 2; ++i) {   /* symmetrise weights (Legendre.C:146-149) */
        ld t = 0.5L * (w[i] + w[n - i]);
        w[i] = w[n - i] = t;
    }
    if (alpha == beta) {   /* the rule is symmetric: remove the eigen-solver's last-bit asymmetry of the nodes */
        for (int i = 0; i <= (n - 1) / 2; ++i) {
            ld t = 0.5L * (x[n - i] - x[i]);
            x[i] = -t;
            x[n - i] = t;
        }
        if (n % 2 == 0) x[n / 2] = 0.0L;
    }
    free(A); free(Z); free(d);
    return m;
}

/* Gauss-Lobatto rule nodes (Legendre.C:160-172). */
static int jacobi_gl_ld(ld alpha, ld beta, int n, ld* x)
{
    for (int i = 0; i <= n; ++i) x[i] = 0;
    x[0] = -1.0L;
    x[n] = 1.0L;
    if (n < 2) return n + 1;
    ld* xi = malloc(sizeof(ld) * (n - 1));
    ld* wi = malloc(sizeof(ld) * (n - 1));
    jacobi_gq_ld(alpha + 1, beta + 1, n - 2, xi, wi);
    for (int i = 1; i < n; ++i) x[i] = xi[i - 1];
    free(xi); free(wi);
    return n + 1;
}

int or_jacobi_gq(double alpha, double beta, int n, double* x, double* w)
{
    ld xl[64], wl[64];
    if (n + 1 > 64) return -1;
    int m = jacobi_gq_ld(alpha, beta, n, xl, wl);
    for (int i = 0; i < m; ++i) { x[i] = (double)xl[i]; w[i] = (double)wl[i]; }
    return m;
}

int or_jacobi_gl(double alpha, double beta, int n, double* x)
{
    ld xl[64];
    if (n + 1 > 64) return -1;
    int m = jacobi_gl_ld(alpha, beta, n, xl);
    for (int i = 0; i < m; ++i) x[i] = (double)xl[i];
    return m;
}

int or_np(int N) { return (N + 1) * (N + 2) / 2; }

int or_nq(int N)
{
    int order = 3 * (N + 1);
    if (order > or_cub_max_order) return -1;
    return or_cub_npts[order];
}

/* Gauss-Jordan inverse with partial pivoting. */
static int invert_ld(int n, const ld* A, ld* Ainv)
{
    ld* a = malloc(sizeof(ld) * n * n);
    memcpy(a, A, sizeof(ld) * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ainv[i * n + j] = (i == j);
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[r * n + c]) > fabs(a[p * n + c])) p = r;
        if (a[p * n + c] == 0.0L) { free(a); return -1; }
        if (p != c)
            for (int k = 0; k < n; ++k) {
                ld t = a[c * n + k]; a[c * n + k] = a[p * n + k]; a[p * n + k] = t;
                t = Ainv[c * n + k]; Ainv[c * n + k] = Ainv[p * n + k]; Ainv[p * n + k] = t;
            }
        ld inv = 1.0L / a[c * n + c];
        for (int k = 0; k < n; ++k) { a[c * n + k] *= inv; Ainv[c * n + k] *= inv; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            ld f = a[r * n + c];
            if (f == 0.0L) continue;
            for (int k = 0; k < n; ++k) {
                a[r * n + k] -= f * a[c * n + k];
                Ainv[r * n + k] -= f * Ainv[c * n + k];
            }
        }
    }
    free(a);
    return 0;
}

int or_invert(int n, const double* A, double* Ainv)
{
    if (n <= 0) return -1;
    ld* a = malloc(sizeof(ld) * n * n);
    ld* b = malloc(sizeof(ld) * n * n);
    if (!a || !b) {
        free(a); free(b);
        return -1;
    }
    for (int i = 0; i < n * n; ++i) a[i] = A[i];
    int rc = invert_ld(n, a, b);
    for (int i = 0; i < n * n; ++i) Ainv[i] = (double)b[i];
    free(a); free(b);
    return rc;
}

/* 1D Vandermonde V[i][j] = P_j(r_i) (Legendre.C:216-230). */
static void vandermonde1d(int N, int m, const ld* r, ld* V)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= N; ++j) V[i * (N + 1) + j] = jacobi_p_ld(r[i], 0, 0, j);
}

/* warpFactor (triangleBaseFunction.C:189-234). */
static void warp_factor(int N, int m, const ld* rout, ld* warp)
{
    ld* lgl = malloc(sizeof(ld) * (N + 1));
    ld* req = malloc(sizeof(ld) * (N + 1));
    ld* Veq = malloc(sizeof(ld) * (N + 1) * (N + 1));
    ld* VeqT = malloc(sizeof(ld) * (N + 1) * (N + 1));
    ld* invVeqT = malloc(sizeof(ld) * (N + 1) * (N + 1));
    jacobi_gl_ld(0, 0, N, lgl);
    for (int i = 0; i <= N; ++i) req[i] = 2.0L / N * i - 1;
    vandermonde1d(N, N + 1, req, Veq);
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) VeqT[i * (N + 1) + j] = Veq[j * (N + 1) + i];
    invert_ld(N + 1, VeqT, invVeqT);
    for (int k = 0; k < m; ++k) {
        /* Lmat[:,k] = inv(Veq') * P(rout_k); warp_k = sum_i Lmat[i,k] (lgl_i - req_i) */
        ld w = 0.0L;
        for (int i = 0; i <= N; ++i) {
            ld l = 0.0L;
            for (int j = 0; j <= N; ++j) l += invVeqT[i * (N + 1) + j] * jacobi_p_ld(rout[k], 0, 0, j);
            w += l * (lgl[i] - req[i]);
        }
        ld zerof = fabs(rout[k]) < (1.0L - 1.0e-10L) ? 1.0L : 0.0L;
        ld sf = 1.0L - (zerof * rout[k]) * (zerof * rout[k]);
        warp[k] = w / sf + w * (zerof - 1.0L);
    }
    free(lgl); free(req); free(Veq); free(VeqT); free(invVeqT);
}

/* Warp & blend nodes on the reference triangle (triangleBaseFunction.C:122-186). */
static int nodes2d_ld(int N, ld* r, ld* s)
{
    static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999,
                                      1.2832, 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
    int Np = or_np(N);
    ld alpha = N < 16 ? (ld)alpopt[N - 1] : 5.0L / 3.0L;
    ld *L1 = malloc(sizeof(ld) * Np), *L2 = malloc(sizeof(ld) * Np), *L3 = malloc(sizeof(ld) * Np);
    ld *x = malloc(sizeof(ld) * Np), *y = malloc(sizeof(ld) * Np);
    ld *t1 = malloc(sizeof(ld) * Np), *t2 = malloc(sizeof(ld) * Np), *t3 = malloc(sizeof(ld) * Np);
    ld *w1 = malloc(sizeof(ld) * Np), *w2 = malloc(sizeof(ld) * Np), *w3 = malloc(sizeof(ld) * Np);
    int sk = 0;
    for (int n = 1; n <= N + 1; ++n)
        for (int mm = 1; mm <= N + 2 - n; ++mm) {
            L1[sk] = (ld)(n - 1) / N;
            L3[sk] = (ld)(mm - 1) / N;
            ++sk;
        }
    const ld sq3 = sqrt(3.0L);
    for (int i = 0; i < Np; ++i) {
        L2[i] = 1.0L - L1[i] - L3[i];
        x[i] = L3[i] - L2[i];
        y[i] = (2 * L1[i] - L2[i] - L3[i]) / sq3;
        t1[i] = L3[i] - L2[i];
        t2[i] = L1[i] - L3[i];
        t3[i] = L2[i] - L1[i];
    }
    warp_factor(N, Np, t1, w1);
    warp_factor(N, Np, t2, w2);
    warp_factor(N, Np, t3, w3);
    const ld pi = 3.14159265358979323846264338327950288L;
    ld cos1 = cos(2 * pi / 3), cos2 = cos(4 * pi / 3), sin1 = sin(2 * pi / 3), sin2 = sin(4 * pi / 3);
    for (int i = 0; i < Np; ++i) {
        ld b1 = 4 * L2[i] * L3[i], b2 = 4 * L1[i] * L3[i], b3 = 4 * L1[i] * L2[i];
        ld a1 = alpha * L1[i], a2 = alpha * L2[i], a3 = alpha * L3[i];
        ld wa = b1 * w1[i] * (1 + a1 * a1);
        ld wb = b2 * w2[i] * (1 + a2 * a2);
        ld wc = b3 * w3[i] * (1 + a3 * a3);
        x[i] += 1 * wa + cos1 * wb + cos2 * wc;
        y[i] += 0 * wa + sin1 * wb + sin2 * wc;
    }
    for (int i = 0; i < Np; ++i) {   /* equilateral (x,y) -> (r,s) */
        ld l1 = (sq3 * y[i] + 1.0L) / 3.0L;
        ld l2 = (-3.0L * x[i] - sq3 * y[i] + 2.0L) / 6.0L;
        ld l3 = (3.0L * x[i] - sq3 * y[i] + 2.0L) / 6.0L;
        r[i] = l3 - l2 - l1;
        s[i] = l1 - l2 - l3;
    }
    free(L1); free(L2); free(L3); free(x); free(y); free(t1); free(t2); free(t3);
    free(w1); free(w2); free(w3);
    return Np;
}

int or_nodes2d(int N, double* r, double* s)
{
    int Np = or_np(N);
    ld* rl = malloc(sizeof(ld) * Np);
    ld* sl = malloc(sizeof(ld) * Np);
    nodes2d_ld(N, rl, sl);
    for (int i = 0; i < Np; ++i) { r[i] = (double)rl[i]; s[i] = (double)sl[i]; }
    free(rl); free(sl);
    return Np;
}

/* Vandermonde2D (Legendre.C:246-275). */
static void vandermonde2d_ld(int N, int m, const ld* r, const ld* s, ld* V)
{
    int Np = or_np(N);
    for (int n = 0; n < m; ++n) {
        ld a = s[n] != 1.0L ? 2 * (1 + r[n]) / (1 - s[n]) - 1 : -1.0L;
        ld b = s[n];
        int sk = 0;
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N - i; ++j) {
                V[n * Np + sk] = sqrt(2.0L) * jacobi_p_ld(a, 0, 0, i) * jacobi_p_ld(b, 2 * i + 1, 0, j) * pow(1 - b, i);
                ++sk;
            }
    }
}

/* GradVandermonde2D (Legendre.C:371-438). */
static void grad_vandermonde2d_ld(int N, int m, const ld* r, const ld* s, ld* Vr, ld* Vs)
{
    int Np = or_np(N);
    for (int n = 0; n < m; ++n) {
        ld a = s[n] != 1.0L ? 2 * (1 + r[n]) / (1 - s[n]) - 1 : -1.0L;
        ld b = s[n];
        int sk = 0;
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N - i; ++j) {
                ld fa = jacobi_p_ld(a, 0, 0, i), gb = jacobi_p_ld(b, 2 * i + 1, 0, j);
                ld dfa = grad_jacobi_p_ld(a, 0, 0, i), dgb = grad_jacobi_p_ld(b, 2 * i + 1, 0, j);
                ld c = pow(2.0L, i + 0.5L);
                ld vr, vs;
                if (i > 0) {
                    vr = dfa * gb * c * pow(0.5L * (1.0L - b), i - 1);
                    ld t = dgb * pow(0.5L * (1 - b), i) - 0.5L * i * gb * pow(0.5L * (1 - b), i - 1);
                    vs = (dfa * gb * 0.5L * (1 + a) * pow(0.5L * (1 - b), i - 1) + fa * t) * c;
                } else {
                    vr = dfa * gb * c;
                    ld t = dgb * pow(0.5L * (1 - b), i);
                    vs = (dfa * gb * 0.5L * (1 + a) + fa * t) * c;
                }
                Vr[n * Np + sk] = vr;
                Vs[n * Np + sk] = vs;
                ++sk;
            }
    }
}

static ld* to_ld(int n, const double* v)
{
    ld* o = malloc(sizeof(ld) * n);
    for (int i = 0; i < n; ++i) o[i] = v[i];
    return o;
}

static void to_double(int n, const ld* v, double* o)
{
    for (int i = 0; i < n; ++i) o[i] = (double)v[i];
}

int or_vandermonde2d(int N, int m, const double* r, const double* s, double* V)
{
    int Np = or_np(N);
    ld *rl = to_ld(m, r), *sl = to_ld(m, s), *Vl = malloc(sizeof(ld) * m * Np);
    vandermonde2d_ld(N, m, rl, sl, Vl);
    to_double(m * Np, Vl, V);
    free(rl); free(sl); free(Vl);
    return 0;
}

int or_grad_vandermonde2d(int N, int m, const double* r, const double* s, double* Vr, double* Vs)
{
    int Np = or_np(N);
    ld *rl = to_ld(m, r), *sl = to_ld(m, s), *a = malloc(sizeof(ld) * m * Np), *b = malloc(sizeof(ld) * m * Np);
    grad_vandermonde2d_ld(N, m, rl, sl, a, b);
    to_double(m * Np, a, Vr);
    to_double(m * Np, b, Vs);
    free(rl); free(sl); free(a); free(b);
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

static void matmul(int m, int k, int n, const ld* A, const ld* B, ld* C)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            ld acc = 0.0L;
            for (int l = 0; l < k; ++l) acc += A[i * k + l] * B[l * n + j];
            C[i * n + j] = acc;
        }
}

/* inverse of the cubature mass matrix (unit Jacobian), formed and inverted in long double */
int or_mass_inverse(int N, double* invM)
{
    int Np = or_np(N), Nq = or_nq(N);
    if (Nq < 0) return -1;
    ld *rl = malloc(sizeof(ld) * Np), *sl = malloc(sizeof(ld) * Np);
    ld *Vl = malloc(sizeof(ld) * Np * Np), *iVl = malloc(sizeof(ld) * Np * Np);
    ld *Vc = malloc(sizeof(ld) * Nq * Np), *Vql = malloc(sizeof(ld) * Nq * Np);
    ld *Ml = malloc(sizeof(ld) * Np * Np), *iMl = malloc(sizeof(ld) * Np * Np);
    ld *cr = to_ld(Nq, or_cub_r[3 * (N + 1)]), *cs = to_ld(Nq, or_cub_s[3 * (N + 1)]);
    ld *cwl = to_ld(Nq, or_cub_w[3 * (N + 1)]);
    nodes2d_ld(N, rl, sl);
    vandermonde2d_ld(N, Np, rl, sl, Vl);
    invert_ld(Np, Vl, iVl);
    vandermonde2d_ld(N, Nq, cr, cs, Vc);
    matmul(Nq, Np, Np, Vc, iVl, Vql);
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            ld acc = 0.0L;
            for (int q = 0; q < Nq; ++q) acc += Vql[q * Np + i] * cwl[q] * Vql[q * Np + j];
            Ml[i * Np + j] = acc;
        }
    int rc = invert_ld(Np, Ml, iMl);
    to_double(Np * Np, iMl, invM);
    free(rl); free(sl); free(Vl); free(iVl); free(Vc); free(Vql); free(Ml); free(iMl); free(cr); free(cs);
    free(cwl);
    return rc;
}

int or_refelem_ops(int N, double* r, double* s, double* V, double* invV, double* Dr, double* Ds,
                   double* Vq, double* Drq, double* Dsq, double* cw, double* Vf, double* gw, double* M)
{
    int Np = or_np(N), Nq = or_nq(N), Nfp = N + 1, Nfq = N + 2;
    if (Nq < 0) return -1;
    ld *rl = malloc(sizeof(ld) * Np), *sl = malloc(sizeof(ld) * Np);
    ld *Vl = malloc(sizeof(ld) * Np * Np), *iVl = malloc(sizeof(ld) * Np * Np);
    ld *Vr = malloc(sizeof(ld) * Np * Np), *Vs = malloc(sizeof(ld) * Np * Np), *D = malloc(sizeof(ld) * Np * Np);
    nodes2d_ld(N, rl, sl);
    vandermonde2d_ld(N, Np, rl, sl, Vl);
    if (invert_ld(Np, Vl, iVl)) return -2;
    grad_vandermonde2d_ld(N, Np, rl, sl, Vr, Vs);
    to_double(Np, rl, r);
    to_double(Np, sl, s);
    to_double(Np * Np, Vl, V);
    to_double(Np * Np, iVl, invV);
    matmul(Np, Np, Np, Vr, iVl, D);
    to_double(Np * Np, D, Dr);
    matmul(Np, Np, Np, Vs, iVl, D);
    to_double(Np * Np, D, Ds);
    /* cubature (gaussTriangleIntegration.C:49-78) */
    ld *cr = to_ld(Nq, or_cub_r[3 * (N + 1)]), *cs = to_ld(Nq, or_cub_s[3 * (N + 1)]);
    ld *cwl = to_ld(Nq, or_cub_w[3 * (N + 1)]);
    memcpy(cw, or_cub_w[3 * (N + 1)], sizeof(double) * Nq);
    ld *Vc = malloc(sizeof(ld) * Nq * Np), *Vcr = malloc(sizeof(ld) * Nq * Np), *Vcs = malloc(sizeof(ld) * Nq * Np);
    ld *Q = malloc(sizeof(ld) * Nq * Np), *Vql = malloc(sizeof(ld) * Nq * Np);
    vandermonde2d_ld(N, Nq, cr, cs, Vc);
    grad_vandermonde2d_ld(N, Nq, cr, cs, Vcr, Vcs);
    matmul(Nq, Np, Np, Vc, iVl, Vql);
    to_double(Nq * Np, Vql, Vq);
    matmul(Nq, Np, Np, Vcr, iVl, Q);
    to_double(Nq * Np, Q, Drq);
    matmul(Nq, Np, Np, Vcs, iVl, Q);
    to_double(Nq * Np, Q, Dsq);
    /* face rule: JacobiGQ(0,0,faceIntOrder/2) with faceIntOrder = 2(N+1) (gaussTriangleIntegration.C:80-94) */
    ld *gr = malloc(sizeof(ld) * Nfq), *gwl = malloc(sizeof(ld) * Nfq), *gl = malloc(sizeof(ld) * Nfp);
    jacobi_gq_ld(0, 0, N + 1, gr, gwl);
    to_double(Nfq, gwl, gw);
    jacobi_gl_ld(0, 0, N, gl);
    ld *V1 = malloc(sizeof(ld) * Nfp * Nfp), *iV1 = malloc(sizeof(ld) * Nfp * Nfp);
    ld *Vg = malloc(sizeof(ld) * Nfq * Nfp), *Vfl = malloc(sizeof(ld) * Nfq * Nfp);
    vandermonde1d(N, Nfp, gl, V1);
    invert_ld(Nfp, V1, iV1);
    vandermonde1d(N, Nfq, gr, Vg);
    matmul(Nfq, Nfp, Nfp, Vg, iV1, Vfl);
    to_double(Nfq * Nfp, Vfl, Vf);
    /* cubature mass matrix with unit Jacobian (physicalCellElement.C:99-108 with WJ = w) */
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            ld acc = 0.0L;
            for (int q = 0; q < Nq; ++q) acc += Vql[q * Np + i] * cwl[q] * Vql[q * Np + j];
            M[i * Np + j] = (double)acc;
        }
    free(rl); free(sl); free(Vl); free(iVl); free(Vr); free(Vs); free(D); free(cr); free(cs); free(cwl);
    free(Vc); free(Vcr); free(Vcs); free(Q); free(Vql); free(gr); free(gwl); free(gl); free(V1); free(iV1);
    free(Vg); free(Vfl);
    return 0;
}
