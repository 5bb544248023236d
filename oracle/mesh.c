/*
 * Oracle mesh: connectivity and affine geometric factors (test infrastructure only).
 *
 * Restates HopeFOAM-0.1/src/DG/dgMesh/dgPolyMesh.C:412-530 (face k of a triangle
 * joins vertices k and k+1, anticlockwise), physicalFaceElement.C:80-93 and
 * triangleBaseFunction.C:96-107 (the neighbour traverses a shared edge in the
 * opposite direction, faceRotate = 1), physicalCellElement.C:72-110
 * (x = -(r+s)/2 v0 + (r+1)/2 v1 + (s+1)/2 v2, dx/dr, dr/dx, J) and
 * triangleBaseFunction.C:330-419 / gaussTriangleIntegration.C:153-192
 * (face normals, edge Jacobian Js, fscale = Js/J, max over the two sides).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

typedef struct { int a, b, e, f; } face_key;

static int cmp_face(const void* p, const void* q)
{
    const face_key* x = p;
    const face_key* y = q;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->b != y->b) return x->b < y->b ? -1 : 1;
    if (x->e != y->e) return x->e < y->e ? -1 : 1;
    return x->f - y->f;
}

#define ALLOC(p, n) (p) = calloc((size_t)(n), sizeof(*(p)))

or_mesh* or_mesh_create(int N, int K, const double* vxy, const int* vid, const int* bctag)
{
    if (N < 1 || N > 8 || K < 1) return NULL;
    or_mesh* m = calloc(1, sizeof(or_mesh));
    m->N = N; m->Np = or_np(N); m->Nfp = N + 1; m->Nq = or_nq(N); m->Nfq = N + 2; m->K = K;
    int Np = m->Np, Nq = m->Nq, Nfp = m->Nfp, Nfq = m->Nfq;
    ALLOC(m->r, Np); ALLOC(m->s, Np); ALLOC(m->V, Np * Np); ALLOC(m->invV, Np * Np);
    ALLOC(m->Dr, Np * Np); ALLOC(m->Ds, Np * Np); ALLOC(m->Vq, Nq * Np); ALLOC(m->Drq, Nq * Np);
    ALLOC(m->Dsq, Nq * Np); ALLOC(m->cw, Nq); ALLOC(m->Vf, Nfq * Nfp); ALLOC(m->gw, Nfq);
    ALLOC(m->M, Np * Np); ALLOC(m->invM, Np * Np); ALLOC(m->fmask, 3 * Nfp);
    or_refelem_ops(N, m->r, m->s, m->V, m->invV, m->Dr, m->Ds, m->Vq, m->Drq, m->Dsq, m->cw, m->Vf, m->gw, m->M);
    or_mass_inverse(N, m->invM);
    or_fmask(N, m->fmask);

    ALLOC(m->vxy, K * 6); memcpy(m->vxy, vxy, sizeof(double) * K * 6);
    ALLOC(m->vid, K * 3); memcpy(m->vid, vid, sizeof(int) * K * 3);
    ALLOC(m->bctag, K * 3);
    if (bctag) memcpy(m->bctag, bctag, sizeof(int) * K * 3);
    ALLOC(m->EToE, K * 3); ALLOC(m->EToF, K * 3);
    ALLOC(m->J, K); ALLOC(m->rx, K); ALLOC(m->ry, K); ALLOC(m->sx, K); ALLOC(m->sy, K);
    ALLOC(m->nx, K * 3); ALLOC(m->ny, K * 3); ALLOC(m->sJ, K * 3); ALLOC(m->fscale_max, K * 3);
    ALLOC(m->x, K * Np); ALLOC(m->y, K * Np);

    /* connectivity by matching sorted vertex-id pairs */
    face_key* fk = malloc(sizeof(face_key) * 3 * K);
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            int a = vid[e * 3 + f], b = vid[e * 3 + (f + 1) % 3];
            face_key k = {a < b ? a : b, a < b ? b : a, e, f};
            fk[e * 3 + f] = k;
        }
    qsort(fk, (size_t)3 * K, sizeof(face_key), cmp_face);
    for (int i = 0; i < 3 * K; ++i) { m->EToE[i] = -1; m->EToF[i] = -1; }
    for (int i = 0; i < 3 * K;) {
        int j = i + 1;
        while (j < 3 * K && fk[j].a == fk[i].a && fk[j].b == fk[i].b) ++j;
        if (j - i == 2) {
            face_key p = fk[i], q = fk[i + 1];
            m->EToE[p.e * 3 + p.f] = q.e; m->EToF[p.e * 3 + p.f] = q.f;
            m->EToE[q.e * 3 + q.f] = p.e; m->EToF[q.e * 3 + q.f] = p.f;
        } else if (j - i > 2) {
            free(fk);
            or_mesh_destroy(m);
            return NULL;             /* non-manifold */
        }
        i = j;
    }
    free(fk);
    m->nbfaces = 0;
    for (int i = 0; i < 3 * K; ++i) m->nbfaces += m->EToE[i] < 0;

    /* geometry */
    for (int e = 0; e < K; ++e) {
        const double* v = vxy + e * 6;
        double xr = 0.5 * (v[2] - v[0]), yr = 0.5 * (v[3] - v[1]);
        double xs = 0.5 * (v[4] - v[0]), ys = 0.5 * (v[5] - v[1]);
        double J = xr * ys - yr * xs;
        m->J[e] = J;
        m->rx[e] = ys / J; m->sx[e] = -yr / J; m->ry[e] = -xs / J; m->sy[e] = xr / J;
        double fn[3][2] = {{yr, -xr}, {ys - yr, xr - xs}, {-ys, xs}};
        for (int f = 0; f < 3; ++f) {
            double Js = sqrt(fn[f][0] * fn[f][0] + fn[f][1] * fn[f][1]);
            m->nx[e * 3 + f] = fn[f][0] / Js;
            m->ny[e * 3 + f] = fn[f][1] / Js;
            m->sJ[e * 3 + f] = Js;
        }
        for (int i = 0; i < Np; ++i) {
            double r = m->r[i], s = m->s[i];
            m->x[e * Np + i] = -(r + s) * 0.5 * v[0] + (r + 1) * 0.5 * v[2] + (s + 1) * 0.5 * v[4];
            m->y[e * Np + i] = -(r + s) * 0.5 * v[1] + (r + 1) * 0.5 * v[3] + (s + 1) * 0.5 * v[5];
        }
    }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            double fs = m->sJ[e * 3 + f] / m->J[e];
            int en = m->EToE[e * 3 + f];
            if (en >= 0) {
                double fo = m->sJ[en * 3 + m->EToF[e * 3 + f]] / m->J[en];
                if (fo > fs) fs = fo;
            }
            m->fscale_max[e * 3 + f] = fs;
        }
    return m;
}

void or_mesh_destroy(or_mesh* m)
{
    if (!m) return;
    void* ptrs[] = {m->r, m->s, m->V, m->invV, m->Dr, m->Ds, m->Vq, m->Drq, m->Dsq, m->cw, m->Vf, m->gw,
                    m->M, m->invM, m->fmask, m->vxy, m->vid, m->bctag, m->EToE, m->EToF, m->J, m->rx,
                    m->ry, m->sx, m->sy, m->nx, m->ny, m->sJ, m->fscale_max, m->x, m->y};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(m);
}

int or_mesh_info(const or_mesh* m, int* o)
{
    o[0] = m->N; o[1] = m->Np; o[2] = m->Nfp; o[3] = m->Nq; o[4] = m->Nfq; o[5] = m->K; o[6] = m->nbfaces;
    return 0;
}

int or_mesh_conn(const or_mesh* m, int* EToE, int* EToF)
{
    memcpy(EToE, m->EToE, sizeof(int) * 3 * m->K);
    memcpy(EToF, m->EToF, sizeof(int) * 3 * m->K);
    return 0;
}

int or_mesh_nodes(const or_mesh* m, double* x, double* y)
{
    memcpy(x, m->x, sizeof(double) * m->K * m->Np);
    memcpy(y, m->y, sizeof(double) * m->K * m->Np);
    return 0;
}

int or_mesh_geo(const or_mesh* m, double* g)
{
    for (int e = 0; e < m->K; ++e) {
        double* o = g + e * 14;
        o[0] = m->rx[e]; o[1] = m->ry[e]; o[2] = m->sx[e]; o[3] = m->sy[e]; o[4] = m->J[e];
        for (int f = 0; f < 3; ++f) {
            o[5 + 3 * f] = m->nx[e * 3 + f];
            o[6 + 3 * f] = m->ny[e * 3 + f];
            o[7 + 3 * f] = m->sJ[e * 3 + f];
        }
    }
    return 0;
}
