/*
 * Oracle TVB minmod limiter for triangles (test infrastructure only).
 *
 * Not in the reference: HopeFOAM's own triangle limiter (Trianglelimite.C:61-852) is a different,
 * gradient-weighting scheme.  This restates the published TVB limiter of Cockburn & Shu
 * (J. Comput. Phys. 141 (1998) 199-224, section 2.4, the triangle limiter), component-wise on the
 * conserved variables, with the troubled-cell test as the detection step:
 *
 *   for element 0 with edge midpoints m_i (edge i joins vertices i and i+1) and face neighbours i:
 *     ubar_0        cell mean;  ut(m_i) value of the P1 part (L2 projection) at m_i
 *     m_i - b_0 = a1 (b_i - b_0) + a2 (b_j - b_0)   with a1, a2 >= 0, j one of the two other faces
 *     D(m_i)        = a1 (ubar_i - ubar_0) + a2 (ubar_j - ubar_0)
 *     d_i           = mbar(ut(m_i) - ubar_0, nu D(m_i)),   mbar(a, b) = a if |a| <= M h^2 else minmod(a, b)
 *   the cell is troubled when some variable has d_i != ut(m_i) - ubar_0; then every variable becomes
 *   ubar_0 + sum_i dh_i phi_i (phi_i the P1 function with phi_i(m_k) = delta_ik) where dh = d when
 *   sum d = 0, else pos = sum max(0, d_i), neg = sum max(0, -d_i),
 *   dh_i = min(1, neg/pos) max(0, d_i) - min(1, pos/neg) max(0, -d_i).
 *
 * Barycentres of the neighbours are taken through the shared edge (b_i = m_i + b_n - m_n), so
 * periodic images need no special case.  A boundary face gets a ghost cell mirrored through the edge
 * (b = 2 m_i - b_0) whose mean is: the edge average of the fixedValue data (tag 1, when given), the
 * own mean with the normal momentum reversed (tag 3 with 4 variables, as reflectiveDgPatchField.C:111-154
 * mirrors the state), or the own mean (every other tag, the zeroGradient copy).  h = the element's
 * longest edge.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

/* Mean weights (pmean, Np), P1-part values at the three edge midpoints (pmid, 3 x Np) and the P1 edge
 * functions at the nodes (phi, Np x 3). */
int or_tvb_ops(int N, double* pmean, double* pmid, double* phi)
{
    int Np = or_np(N);
    double *r = malloc(sizeof(double) * Np), *s = malloc(sizeof(double) * Np);
    double *V = malloc(sizeof(double) * Np * Np), *invV = malloc(sizeof(double) * Np * Np);
    double mr[3] = {0.0, 0.0, -1.0}, ms[3] = {-1.0, 0.0, 0.0}, Vm[3 * 45], c0[45];
    or_nodes2d(N, r, s);
    or_vandermonde2d(N, Np, r, s, V);
    or_invert(Np, V, invV);
    or_vandermonde2d(N, 3, mr, ms, Vm);
    /* the degree <= 1 modes of the (i, j) ordering: (0,0) -> 0, (0,1) -> 1, (1,0) -> N+1 */
    int p1[3] = {0, 1, N + 1};
    or_vandermonde2d(N, 1, mr, ms, c0);   /* constant mode value */
    for (int j = 0; j < Np; ++j) {
        pmean[j] = c0[0] * invV[0 * Np + j];
        for (int i = 0; i < 3; ++i) {
            double a = 0.0;
            for (int k = 0; k < 3; ++k) a += Vm[i * Np + p1[k]] * invV[p1[k] * Np + j];
            pmid[i * Np + j] = a;
        }
    }
    for (int n = 0; n < Np; ++n) {
        double l0 = -(r[n] + s[n]) * 0.5, l1 = (r[n] + 1.0) * 0.5, l2 = (s[n] + 1.0) * 0.5;
        phi[n * 3 + 0] = 1.0 - 2.0 * l2;   /* edge 0 = (v0, v1), opposite v2 */
        phi[n * 3 + 1] = 1.0 - 2.0 * l0;   /* edge 1 = (v1, v2), opposite v0 */
        phi[n * 3 + 2] = 1.0 - 2.0 * l1;   /* edge 2 = (v2, v0), opposite v1 */
    }
    free(r); free(s); free(V); free(invV);
    return Np;
}

/* per element and edge: the neighbour pair coefficients (a1, a2, j) and h */
int or_tvb_geometry(const or_mesh* m, double* coef, int* jsel, double* h)
{
    for (int e = 0; e < m->K; ++e) {
        const double* v = m->vxy + e * 6;
        double bx = (v[0] + v[2] + v[4]) / 3.0, by = (v[1] + v[3] + v[5]) / 3.0;
        double dx[3], dy[3], tx[3], ty[3], hmax = 0.0;
        for (int f = 0; f < 3; ++f) {
            int a = f, b = (f + 1) % 3;
            double mx = 0.5 * (v[2 * a] + v[2 * b]), my = 0.5 * (v[2 * a + 1] + v[2 * b + 1]);
            double ex = v[2 * b] - v[2 * a], ey = v[2 * b + 1] - v[2 * a + 1];
            hmax = fmax(hmax, sqrt(ex * ex + ey * ey));
            tx[f] = mx - bx;
            ty[f] = my - by;
            int en = m->EToE[e * 3 + f];
            if (en >= 0) {
                int fn = m->EToF[e * 3 + f];
                const double* w = m->vxy + en * 6;
                double cx = (w[0] + w[2] + w[4]) / 3.0, cy = (w[1] + w[3] + w[5]) / 3.0;
                int p = fn, q = (fn + 1) % 3;
                double nmx = 0.5 * (w[2 * p] + w[2 * q]), nmy = 0.5 * (w[2 * p + 1] + w[2 * q + 1]);
                dx[f] = tx[f] + (cx - nmx);
                dy[f] = ty[f] + (cy - nmy);
            } else {
                dx[f] = 2.0 * tx[f];
                dy[f] = 2.0 * ty[f];
            }
        }
        h[e] = hmax;
        for (int f = 0; f < 3; ++f) {
            double best = -1e300, ba1 = 0.0, ba2 = 0.0;
            int bj = (f + 1) % 3;
            for (int k = 1; k <= 2; ++k) {
                int j = (f + k) % 3;
                double det = dx[f] * dy[j] - dy[f] * dx[j];
                double a1 = (tx[f] * dy[j] - ty[f] * dx[j]) / det;
                double a2 = (dx[f] * ty[f] - dy[f] * tx[f]) / det;
                double score = fmin(a1, a2);
                if (score > best + 1e-12) { best = score; ba1 = a1; ba2 = a2; bj = j; }
            }
            coef[(e * 3 + f) * 2 + 0] = ba1;
            coef[(e * 3 + f) * 2 + 1] = ba2;
            jsel[e * 3 + f] = bj;
        }
    }
    return 0;
}

static double minmod2(double a, double b)
{
    if (a > 0.0 && b > 0.0) return fmin(a, b);
    if (a < 0.0 && b < 0.0) return fmax(a, b);
    return 0.0;
}

int or_tvb_limit(const or_mesh* m, double* q, int nvar, double M, double nu, const double* bcq, int* troubled)
{
    int K = m->K, Np = m->Np, Nfp = m->Nfp;
    double *pmean = malloc(sizeof(double) * Np), *pmid = malloc(sizeof(double) * 3 * Np);
    double *phi = malloc(sizeof(double) * Np * 3), *coef = malloc(sizeof(double) * K * 6);
    double *h = malloc(sizeof(double) * K), *mean = malloc(sizeof(double) * nvar * K);
    double* um = malloc(sizeof(double) * nvar * K * 3);
    int* jsel = malloc(sizeof(int) * K * 3);
    or_tvb_ops(m->N, pmean, pmid, phi);
    or_tvb_geometry(m, coef, jsel, h);
    for (int v = 0; v < nvar; ++v)
        for (int e = 0; e < K; ++e) {
            const double* u = q + ((size_t)v * K + e) * Np;
            double a = 0.0, b[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < Np; ++j) {
                a += pmean[j] * u[j];
                for (int i = 0; i < 3; ++i) b[i] += pmid[i * Np + j] * u[j];
            }
            mean[(size_t)v * K + e] = a;
            for (int i = 0; i < 3; ++i) um[((size_t)v * K + e) * 3 + i] = b[i];
        }
    int ntroubled = 0;
    double* nb = malloc(sizeof(double) * nvar * 3);
    double* d = malloc(sizeof(double) * nvar * 3);
    for (int e = 0; e < K; ++e) {
        /* neighbour / ghost means */
        for (int f = 0; f < 3; ++f) {
            int en = m->EToE[e * 3 + f], tag = m->bctag[e * 3 + f];
            for (int v = 0; v < nvar; ++v) nb[v * 3 + f] = en >= 0 ? mean[(size_t)v * K + en] : mean[(size_t)v * K + e];
            if (en >= 0) continue;
            if (tag == 1 && bcq) {
                for (int v = 0; v < nvar; ++v) {
                    const double* g = bcq + (((size_t)v * K + e) * 3 + f) * Nfp;
                    double a = 0.0;
                    for (int gq = 0; gq < m->Nfq; ++gq) {
                        double t = 0.0;
                        for (int k = 0; k < Nfp; ++k) t += m->Vf[gq * Nfp + k] * g[k];
                        a += 0.5 * m->gw[gq] * t;
                    }
                    nb[v * 3 + f] = a;
                }
            } else if (tag == 3 && nvar == 4) {
                double nx = m->nx[e * 3 + f], ny = m->ny[e * 3 + f];
                double un = nb[1 * 3 + f] * nx + nb[2 * 3 + f] * ny;
                nb[1 * 3 + f] -= 2.0 * un * nx;
                nb[2 * 3 + f] -= 2.0 * un * ny;
            }
        }
        int bad = 0;
        double tol = M * h[e] * h[e];
        for (int v = 0; v < nvar; ++v) {
            double u0 = mean[(size_t)v * K + e];
            for (int f = 0; f < 3; ++f) {
                double a = um[((size_t)v * K + e) * 3 + f] - u0;
                int j = jsel[e * 3 + f];
                double D = coef[(e * 3 + f) * 2] * (nb[v * 3 + f] - u0) + coef[(e * 3 + f) * 2 + 1] * (nb[v * 3 + j] - u0);
                double r = fabs(a) <= tol ? a : minmod2(a, nu * D);
                d[v * 3 + f] = r;
                bad |= r != a;
            }
        }
        if (troubled) troubled[e] = bad;
        if (!bad) continue;
        ++ntroubled;
        for (int v = 0; v < nvar; ++v) {
            double* dv = d + v * 3;
            double sum = dv[0] + dv[1] + dv[2];
            double dh[3] = {dv[0], dv[1], dv[2]};
            if (sum != 0.0) {
                double pos = 0.0, neg = 0.0;
                for (int i = 0; i < 3; ++i) { pos += fmax(0.0, dv[i]); neg += fmax(0.0, -dv[i]); }
                double tp = pos > 0.0 ? fmin(1.0, neg / pos) : 0.0, tn = neg > 0.0 ? fmin(1.0, pos / neg) : 0.0;
                for (int i = 0; i < 3; ++i) dh[i] = tp * fmax(0.0, dv[i]) - tn * fmax(0.0, -dv[i]);
            }
            double u0 = mean[(size_t)v * K + e];
            double* u = q + ((size_t)v * K + e) * Np;
            for (int n = 0; n < Np; ++n) u[n] = u0 + dh[0] * phi[n * 3] + dh[1] * phi[n * 3 + 1] + dh[2] * phi[n * 3 + 2];
        }
    }
    free(pmean); free(pmid); free(phi); free(coef); free(h); free(mean); free(um); free(jsel); free(nb); free(d);
    return ntroubled;
}

/*
 * Positivity-preserving scaling (Zhang & Shu, J. Comput. Phys. 229 (2010) 8918-8934; on triangles
 * Zhang, Xia & Shu, J. Sci. Comput. 50 (2012) 29-62), applied to every cell of a 4-variable Euler state
 * after the TVB step.  The admissible set is checked at the points where the right-hand side evaluates
 * the state: the Np nodes, the Nq cubature points and the 3 Nfq face Gauss points.  With eps_c =
 * min(eps, rhobar, p(Ubar)):
 *   rho <- rhobar + t1 (rho - rhobar),   t1 = (rhobar - eps_c) / (rhobar - rho_min) when rho_min < eps_c;
 *   U   <- Ubar + t2 (U - Ubar),         t2 = min over points with p < eps_c of the t in [0, 1] with
 *                                        p(Ubar + t (U_x - Ubar)) = eps_c (60 bisection steps, keeping the
 *                                        admissible end).
 * Cell means are unchanged.  Returns the number of cells scaled.
 */
static double pressure(double r, double mx, double my, double E, double gamma)
{
    return (gamma - 1.0) * (E - 0.5 * (mx * mx + my * my) / r);
}

/* point-evaluation rows: nodes, cubature points, face Gauss points ((Np + Nq + 3 Nfq) x Np) */
int or_positivity_points(const or_mesh* m, double* V)
{
    int Np = m->Np, Nq = m->Nq, Nfq = m->Nfq, Nfp = m->Nfp, np = Np + Nq + 3 * Nfq;
    memset(V, 0, sizeof(double) * np * Np);
    for (int i = 0; i < Np; ++i) V[i * Np + i] = 1.0;
    for (int q = 0; q < Nq; ++q)
        for (int j = 0; j < Np; ++j) V[(Np + q) * Np + j] = m->Vq[q * Np + j];
    for (int f = 0; f < 3; ++f)
        for (int g = 0; g < Nfq; ++g)
            for (int k = 0; k < Nfp; ++k) V[(Np + Nq + f * Nfq + g) * Np + m->fmask[f * Nfp + k]] += m->Vf[g * Nfp + k];
    return np;
}

int or_positivity(const or_mesh* m, double* q, double gamma, double eps, int* scaled)
{
    int K = m->K, Np = m->Np, np = Np + m->Nq + 3 * m->Nfq, nscaled = 0;
    double *V = malloc(sizeof(double) * np * Np), *pmean = malloc(sizeof(double) * Np);
    double *pmid = malloc(sizeof(double) * 3 * Np), *phi = malloc(sizeof(double) * 3 * Np);
    or_positivity_points(m, V);
    or_tvb_ops(m->N, pmean, pmid, phi);
    for (int e = 0; e < K; ++e) {
        double* u[4];
        double ub[4];
        for (int v = 0; v < 4; ++v) {
            u[v] = q + ((size_t)v * K + e) * Np;
            double a = 0.0;
            for (int j = 0; j < Np; ++j) a += pmean[j] * u[v][j];
            ub[v] = a;
        }
        double pb = pressure(ub[0], ub[1], ub[2], ub[3], gamma);
        int did = 0;
        if (ub[0] > 0.0 && pb > 0.0) {
            double ec = fmin(eps, fmin(ub[0], pb));
            double rmin = 1e300;
            for (int x = 0; x < np; ++x) {
                double r = 0.0;
                for (int j = 0; j < Np; ++j) r += V[x * Np + j] * u[0][j];
                rmin = fmin(rmin, r);
            }
            if (rmin < ec) {
                double t1 = (ub[0] - ec) / (ub[0] - rmin);
                for (int j = 0; j < Np; ++j) u[0][j] = ub[0] + t1 * (u[0][j] - ub[0]);
                did = 1;
            }
            double t2 = 1.0;
            for (int x = 0; x < np; ++x) {
                double ux[4];
                for (int v = 0; v < 4; ++v) {
                    double a = 0.0;
                    for (int j = 0; j < Np; ++j) a += V[x * Np + j] * u[v][j];
                    ux[v] = a;
                }
                if (pressure(ux[0], ux[1], ux[2], ux[3], gamma) >= ec) continue;
                double lo = 0.0, hi = 1.0;
                for (int it = 0; it < 60; ++it) {
                    double t = 0.5 * (lo + hi);
                    double r = ub[0] + t * (ux[0] - ub[0]), mx = ub[1] + t * (ux[1] - ub[1]);
                    double my = ub[2] + t * (ux[2] - ub[2]), E = ub[3] + t * (ux[3] - ub[3]);
                    if (pressure(r, mx, my, E, gamma) >= ec) lo = t; else hi = t;
                }
                t2 = fmin(t2, lo);
            }
            if (t2 < 1.0) {
                for (int v = 0; v < 4; ++v)
                    for (int j = 0; j < Np; ++j) u[v][j] = ub[v] + t2 * (u[v][j] - ub[v]);
                did = 1;
            }
        }
        if (scaled) scaled[e] = did;
        nscaled += did;
    }
    free(V); free(pmean); free(pmid); free(phi);
    return nscaled;
}
