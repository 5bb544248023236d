/*
 * Oracle of the reference's own limiter, Trianglelimite::limite (test infrastructure only).
 *
 * Restates HopeFOAM-0.1 src/DG/DG/godunovFlux/limiteSchemes/scheme/Trianglelimite/Trianglelimite.C:61-852,
 * a weighted Green-Gauss P1 reconstruction of every cell in primitive variables (the EulerLimiter2D
 * scheme of Hesthaven & Warburton's book, section 5.6), step by step:
 *
 *   1 (:92-133)   cell means with the weights w_j = sum_i M_ij / 2 (M the reference mass matrix), the
 *                 centroid with the same weights, A0 = (2/3) J (a third of the cell's area);
 *   2 (:150-262)  one ghost cell per boundary face: centroid = the owner's centroid mirrored through
 *                 the face line (normal n through the face's first node); means = reflective: rho, E
 *                 copied, momentum minus its normal part (m - n (n.m)); fixedValue: the boundary data
 *                 at the face's first node; otherwise the owner's means;
 *   3 (:295-303)  primitive means u = m1/rho, v = m2/rho, p = (gamma-1)(E - |m|^2/(2 rho)), ghosts too;
 *   4 (:307-458)  per face (visited from its owner): conserved values at the two face endpoints averaged
 *                 over the two sides (a boundary side takes the patch value: fixedValue data, the mirror
 *                 state m - 2 n (n.m) on a reflective wall, the owner's value otherwise), turned into
 *                 primitives; A = half the cross product of (centroid_n - centroid_o) and (P_1 - P_0),
 *                 the signed area of the diamond; A_2 = A0_o + A0_n (2 A0_o on the boundary);
 *   5 (:486-520)  face gradients (Green-Gauss over the diamond) of rho (conserved mean), u, v, p:
 *                   Vx =  ((phibar_n - phibar_o)(y_1 - y_0) + (phi_0 - phi_1)(yc_n - yc_o)) / (2 A)
 *                   Vy = -((phibar_n - phibar_o)(x_1 - x_0) + (phi_0 - phi_1)(xc_n - xc_o)) / (2 A)
 *                 CellA2 = sum of A_2 over the cell's faces;
 *   6 (:527-603)  cell gradients CV = sum_faces A_2 V / CellA2; a ghost cell's gradient is its face's V;
 *   7 (:695-788)  per cell, from the gradients of its three face neighbours (ghosts included)
 *                 g_i = |CV_i|^2, w_1 = (g_2 g_3 + eps) / (g_1^2 + g_2^2 + g_3^2 + 3 eps) (cyclic),
 *                 eps = 1e-10, limited gradient L = sum_i w_i CV_i, for each of rho, u, v, p;
 *   8 (:791-848)  at every node X, du = (X - centroid).L_rho (du1, du2, du3 likewise for u, v, p):
 *                 rho = rhobar + du with du halved while rho < tol; rhou = m1bar + rhobar du1 + du ubar;
 *                 rhov likewise; E = Ebar + du3/(gamma-1) + du (ubar^2 + vbar^2)/2 + rhobar (ubar du1 + vbar du2);
 *                 a node whose pressure is below tol is reported (the reference prints a warning).
 *
 * Choices where the reference leaves the behaviour open for these data:
 *   - the density loop of step 8 never ends in the reference when rhobar < tol; here it stops when du
 *     has underflowed to zero, or after 2100 halvings (flags bit 1 is set either way);
 *   - a fixedValue ghost mean is the first value of the face's patch (Trianglelimite.C:220-224) when the
 *     patch map pfirst is given (or_triangle_limit_pf: pfirst[e 3 + f] = e' 3 + f', the patch's first face,
 *     whose first face node holds that value), else the data at its own face's first node (the same for
 *     uniform patch data);
 *   - the patch value of step 4 is the boundary condition evaluated on the current state (the reference
 *     reads whatever its boundary field last held);
 *   - periodic neighbours: the neighbour's centroid is moved through the shared edge (centroid_n +
 *     midpoint_o - midpoint_n), which leaves it unchanged for every other face;
 *   - sel (may be NULL): cells with sel[e] == 0 keep their values (the troubled-cell argument of the
 *     limite overload at :45-57, empty in the reference; NULL = every cell, as the reference).
 * flags[e] (may be NULL): bit 0 rebuilt, bit 1 the density floor shortened du at some node, bit 2 some
 * node's pressure < tol.  Boundary tags: 1 fixedValue (bcq: 4 x K x 3 x Nfp, may be NULL = copy),
 * 3 reflective wall, others copy.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

/* node index of vertex i of the reference triangle (row 0 of the nodal set runs from v0 to v1, the last
 * node is v2) */
static int vertex_node(int N, int i)
{
    int Np = (N + 1) * (N + 2) / 2;
    return i == 0 ? 0 : (i == 1 ? N : Np - 1);
}

typedef struct {
    int o, n;          /* owner cell, neighbour cell (interior index or ghost index >= K) */
    int interior;
    double A_2;
    double V[8];       /* face gradients (x, y) of rho, u, v, p */
} tri_face;

int or_triangle_limit_pf(const or_mesh* m, double* q, double gamma, double tol, const double* bcq, const int* sel,
                         const int* pfirst, int* flags)
{
    const int K = m->K, Np = m->Np, Nfp = m->Nfp, N = m->N;
    const double epse = 1e-10;
    int nb = 0, nf = 0;
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            const int en = m->EToE[e * 3 + f], fn = m->EToF[e * 3 + f];
            nb += en < 0;
            nf += en < 0 || en > e || (en == e && fn > f);
        }
    const int T = K + nb;
    double* w = calloc(Np, sizeof(double));
    double *ave = calloc((size_t)4 * T, sizeof(double)), *prim = calloc((size_t)3 * T, sizeof(double));
    double *cx = calloc(T, sizeof(double)), *cy = calloc(T, sizeof(double)), *A0 = calloc(K, sizeof(double));
    int* ghost = malloc(sizeof(int) * 3 * K);   /* ghost cell of boundary face (e, f), else -1 */
    double *CV = calloc((size_t)8 * T, sizeof(double)), *CellA2 = calloc(T, sizeof(double));
    tri_face* F = calloc(nf ? nf : 1, sizeof(tri_face));
    /* 1. means, centroids, A0 (:92-133) */
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) w[j] += m->M[i * Np + j] / 2;
    for (int e = 0; e < K; ++e) {
        for (int i = 0; i < Np; ++i) A0[e] += w[i] * m->J[e] * 2 / 3;
        for (int j = 0; j < Np; ++j) {
            for (int v = 0; v < 4; ++v) ave[(size_t)v * T + e] += q[((size_t)v * K + e) * Np + j] * w[j];
            cx[e] += m->x[(size_t)e * Np + j] * w[j];
            cy[e] += m->y[(size_t)e * Np + j] * w[j];
        }
    }
    /* 2. ghost cells (:150-262) */
    int gi = K;
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            ghost[e * 3 + f] = -1;
            if (m->EToE[e * 3 + f] >= 0) continue;
            const int g = gi++, tag = m->bctag[e * 3 + f];
            ghost[e * 3 + f] = g;
            const double A = m->nx[e * 3 + f], B = m->ny[e * 3 + f];
            const int n0 = m->fmask[f * Nfp];
            const double C = -m->x[(size_t)e * Np + n0] * A - m->y[(size_t)e * Np + n0] * B;
            cx[g] = (B * B - A * A) * cx[e] - 2 * A * B * cy[e] - 2 * A * C;
            cy[g] = (-B * B + A * A) * cy[e] - 2 * A * B * cx[e] - 2 * B * C;
            if (tag == 3) {
                const double m1 = ave[(size_t)1 * T + e], m2 = ave[(size_t)2 * T + e];
                ave[g] = ave[e];
                ave[(size_t)1 * T + g] = m1 - A * (A * m1 + B * m2);
                ave[(size_t)2 * T + g] = m2 - B * (A * m1 + B * m2);
                ave[(size_t)3 * T + g] = ave[(size_t)3 * T + e];
            } else if (tag == 1 && bcq) {
                /* the patch's first value (Trianglelimite.C:220-223) when the patch map is given, else the own
                   face's first node */
                const int src = (pfirst && pfirst[e * 3 + f] >= 0) ? pfirst[e * 3 + f] : e * 3 + f;
                for (int v = 0; v < 4; ++v) ave[(size_t)v * T + g] = bcq[((size_t)v * K * 3 + src) * Nfp];
            } else {
                for (int v = 0; v < 4; ++v) ave[(size_t)v * T + g] = ave[(size_t)v * T + e];
            }
        }
    /* 3. primitive means (:295-303) */
    for (int c = 0; c < T; ++c) {
        const double r = ave[c], m1 = ave[(size_t)T + c], m2 = ave[(size_t)2 * T + c], E = ave[(size_t)3 * T + c];
        prim[c] = m1 / r;
        prim[(size_t)T + c] = m2 / r;
        prim[(size_t)2 * T + c] = (gamma - 1) * (E - 0.5 * (pow(m1, 2) + pow(m2, 2)) / r);
    }
    /* 4.-5. faces from their owner: endpoint averages, diamond area, face gradients (:307-520) */
    int fi = 0;
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            const int en = m->EToE[e * 3 + f], fn = m->EToF[e * 3 + f];
            if (en >= 0 && (en < e || (en == e && fn < f))) continue;
            const int own[2] = {m->fmask[f * Nfp], m->fmask[f * Nfp + Nfp - 1]};
            double qo[2][4], qn[2][4];
            for (int s = 0; s < 2; ++s)
                for (int v = 0; v < 4; ++v) qo[s][v] = q[((size_t)v * K + e) * Np + own[s]];
            int nbc;
            double ncx, ncy;
            if (en >= 0) {
                /* the neighbour's node at the same point: its vertex with the same vertex id */
                for (int s = 0; s < 2; ++s) {
                    const int id = m->vid[e * 3 + (s == 0 ? f : (f + 1) % 3)];
                    int vn = 0;
                    for (int k = 0; k < 3; ++k)
                        if (m->vid[en * 3 + k] == id) vn = k;
                    const int node = vertex_node(N, vn);
                    for (int v = 0; v < 4; ++v) qn[s][v] = q[((size_t)v * K + en) * Np + node];
                }
                nbc = en;
                const double *vo = m->vxy + (size_t)e * 6, *vn = m->vxy + (size_t)en * 6;
                const int a = f, b = (f + 1) % 3, p = fn, r = (fn + 1) % 3;
                ncx = cx[en] + (0.5 * (vo[2 * a] + vo[2 * b]) - 0.5 * (vn[2 * p] + vn[2 * r]));
                ncy = cy[en] + (0.5 * (vo[2 * a + 1] + vo[2 * b + 1]) - 0.5 * (vn[2 * p + 1] + vn[2 * r + 1]));
            } else {
                const int tag = m->bctag[e * 3 + f];
                const double nx = m->nx[e * 3 + f], ny = m->ny[e * 3 + f];
                for (int s = 0; s < 2; ++s) {
                    if (tag == 1 && bcq) {
                        for (int v = 0; v < 4; ++v)
                            qn[s][v] = bcq[(((size_t)v * K + e) * 3 + f) * Nfp + (s == 0 ? 0 : Nfp - 1)];
                    } else {
                        for (int v = 0; v < 4; ++v) qn[s][v] = qo[s][v];
                        if (tag == 3) {   /* reflectiveDgPatchField.C:111-154: (I - 2 n n) m */
                            const double un = qo[s][1] * nx + qo[s][2] * ny;
                            qn[s][1] = qo[s][1] - 2.0 * un * nx;
                            qn[s][2] = qo[s][2] - 2.0 * un * ny;
                        }
                    }
                }
                nbc = ghost[e * 3 + f];
                ncx = cx[nbc];
                ncy = cy[nbc];
            }
            tri_face* t = &F[fi++];
            t->o = e;
            t->n = nbc;
            t->interior = en >= 0;
            t->A_2 = en >= 0 ? A0[e] + A0[en] : A0[e] + A0[e];
            const double x1v = m->x[(size_t)e * Np + own[0]], y1v = m->y[(size_t)e * Np + own[0]];
            const double x2v = m->x[(size_t)e * Np + own[1]], y2v = m->y[(size_t)e * Np + own[1]];
            const double x0c = cx[e], y0c = cy[e];
            const double A = ((ncx - x0c) * (y2v - y1v) - (x2v - x1v) * (ncy - y0c)) * 0.5;
            double fv[2][4];
            for (int s = 0; s < 2; ++s) {
                double a[4];
                for (int v = 0; v < 4; ++v) a[v] = qo[s][v] * 0.5 + qn[s][v] * 0.5;
                fv[s][0] = a[0];
                fv[s][1] = a[1] / a[0];
                fv[s][2] = a[2] / a[0];
                fv[s][3] = (gamma - 1) * (a[3] - 0.5 * a[0] * (pow(fv[s][1], 2) + pow(fv[s][2], 2)));
            }
            const double cm_o[4] = {ave[e], prim[e], prim[(size_t)T + e], prim[(size_t)2 * T + e]};
            const double cm_n[4] = {ave[nbc], prim[nbc], prim[(size_t)T + nbc], prim[(size_t)2 * T + nbc]};
            const double temp1 = x2v - x1v, temp2 = ncx - x0c;
            for (int k = 0; k < 4; ++k) {
                t->V[2 * k] = 0.5 * ((cm_n[k] - cm_o[k]) * (y2v - y1v) + (fv[0][k] - fv[1][k]) * (ncy - y0c)) / A;
                t->V[2 * k + 1] = -0.5 * ((cm_n[k] - cm_o[k]) * temp1 + (fv[0][k] - fv[1][k]) * temp2) / A;
            }
            CellA2[e] += t->A_2;
            if (t->interior) CellA2[en] += t->A_2;
        }
    /* 6. cell gradients (:527-603) */
    for (int i = 0; i < nf; ++i) {
        const tri_face* t = &F[i];
        for (int k = 0; k < 8; ++k) {
            CV[(size_t)8 * t->o + k] += t->A_2 * t->V[k] / CellA2[t->o];
            if (t->interior) CV[(size_t)8 * t->n + k] += t->A_2 * t->V[k] / CellA2[t->n];
            else CV[(size_t)8 * t->n + k] = t->V[k];
        }
    }
    /* 7.-8. limited gradients and the reconstruction (:695-848) */
    for (int e = 0; e < K; ++e) {
        if (flags) flags[e] = 0;
        if (sel && !sel[e]) continue;
        int c[3];
        for (int f = 0; f < 3; ++f) c[f] = m->EToE[e * 3 + f] >= 0 ? m->EToE[e * 3 + f] : ghost[e * 3 + f];
        double Lg[8];
        for (int k = 0; k < 4; ++k) {
            const double *c1 = CV + (size_t)8 * c[0] + 2 * k, *c2 = CV + (size_t)8 * c[1] + 2 * k;
            const double* c3 = CV + (size_t)8 * c[2] + 2 * k;
            const double g1 = c1[0] * c1[0] + c1[1] * c1[1];
            const double g2 = c2[0] * c2[0] + c2[1] * c2[1];
            const double g3 = c3[0] * c3[0] + c3[1] * c3[1];
            const double fac = g1 * g1 + g2 * g2 + g3 * g3;
            const double w1 = (g2 * g3 + epse) / (fac + 3 * epse);
            const double w2 = (g1 * g3 + epse) / (fac + 3 * epse);
            const double w3 = (g2 * g1 + epse) / (fac + 3 * epse);
            Lg[2 * k] = w1 * c1[0] + w2 * c2[0] + w3 * c3[0];
            Lg[2 * k + 1] = w1 * c1[1] + w2 * c2[1] + w3 * c3[1];
        }
        const double rb = ave[e], m1b = ave[(size_t)T + e], m2b = ave[(size_t)2 * T + e], Eb = ave[(size_t)3 * T + e];
        const double ub = prim[e], vb = prim[(size_t)T + e];
        int fl = 1;
        for (int n = 0; n < Np; ++n) {
            const double dx = m->x[(size_t)e * Np + n] - cx[e], dy = m->y[(size_t)e * Np + n] - cy[e];
            double du = dx * Lg[0] + dy * Lg[1];
            const double du1 = dx * Lg[2] + dy * Lg[3];
            const double du2 = dx * Lg[4] + dy * Lg[5];
            const double du3 = dx * Lg[6] + dy * Lg[7];
            double Lrho = rb + du;
            for (int it = 0; it < 2100 && Lrho < tol && du != 0.0; ++it) {   /* finite du: zero within 2098 */
                fl |= 2;
                du = du * 0.5;
                Lrho = rb + du;
            }
            if (Lrho < tol) fl |= 2;
            const double r = Lrho;
            const double mu = m1b + rb * du1 + du * ub;
            const double mv = m2b + rb * du2 + du * vb;
            const double dE = (1 / (gamma - 1)) * du3 + 0.5 * du * (pow(ub, 2) + pow(vb, 2)) + rb * (ub * du1 + vb * du2);
            const double E = Eb + dE;
            q[((size_t)0 * K + e) * Np + n] = r;
            q[((size_t)1 * K + e) * Np + n] = mu;
            q[((size_t)2 * K + e) * Np + n] = mv;
            q[((size_t)3 * K + e) * Np + n] = E;
            const double Lp = (gamma - 1) * (E - 0.5 * (pow(mu, 2) + pow(mv, 2)) / r);
            if (Lp < tol) fl |= 4;
        }
        if (flags) flags[e] = fl;
    }
    free(w); free(ave); free(prim); free(cx); free(cy); free(A0); free(ghost); free(CV); free(CellA2); free(F);
    return 0;
}

int or_triangle_limit(const or_mesh* m, double* q, double gamma, double tol, const double* bcq, const int* sel,
                      int* flags)
{
    return or_triangle_limit_pf(m, q, gamma, tol, bcq, sel, NULL, flags);
}
