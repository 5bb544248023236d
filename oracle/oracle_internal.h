/* Internal mesh/operator state of the oracle (test infrastructure only). */
#ifndef ORACLE_INTERNAL_H
#define ORACLE_INTERNAL_H
#include "dgref.h"

struct or_mesh {
    int N, Np, Nfp, Nq, Nfq, K, nbfaces;
    /* reference element */
    double *r, *s, *V, *invV, *Dr, *Ds, *Vq, *Drq, *Dsq, *cw, *Vf, *gw, *M, *invM;
    int* fmask;                   /* 3 x Nfp */
    /* mesh */
    double* vxy;                  /* K x 3 x 2 */
    int *vid, *bctag;             /* K x 3 */
    int *EToE, *EToF;             /* K x 3, EToE = -1 on boundary faces */
    double *J, *rx, *ry, *sx, *sy;        /* K (affine elements) */
    double *nx, *ny, *sJ;                 /* K x 3: outward normal, edge Jacobian |edge|/2 */
    double* fscale_max;                   /* K x 3: max(sJ/J) over the two sides (physicalElementData.C:247-259) */
    double *x, *y;                        /* K x Np */
};

#endif
