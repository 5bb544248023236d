/* A C caller of libdghip.so (what a reference-side binding would be): one Euler right-hand side on a structured
 * periodic mesh through dgh_create / dgh_mesh_structured / dgh_field_* / dgh_rhs, written to stdout as the K x Np
 * x 4 values in the reference's [var][K][Np] order.  Needs a GPU. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "dgh.h"

#define CK(call)                                                                  \
    do {                                                                          \
        int rc_ = (call);                                                         \
        if (rc_ != DGH_OK) {                                                      \
            fprintf(stderr, "%s failed: %d (%s)\n", #call, rc_, dgh_last_error(ctx)); \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main(void) {
    dgh_ctx ctx = NULL;
    if (dgh_create(&ctx, 0) != DGH_OK) {
        fprintf(stderr, "dgh_create failed\n");
        return 1;
    }
    const int N = 3, nx = 5, ny = 4;
    int mesh, info[7];
    CK(dgh_mesh_structured(ctx, N, nx, ny, 0.0, 10.0, 0.0, 10.0, 1, 1, DGH_BC_ZERO_GRADIENT, &mesh));
    CK(dgh_mesh_info(ctx, mesh, info));
    const int Np = info[1], K = info[5];
    double* x = malloc(sizeof(double) * K * Np);
    double* y = malloc(sizeof(double) * K * Np);
    double* q = malloc(sizeof(double) * 4 * K * Np);
    double* r = malloc(sizeof(double) * 4 * K * Np);
    CK(dgh_mesh_nodes(ctx, mesh, x, y));
    for (int i = 0; i < K * Np; ++i) {
        const double rho = 1.0 + 0.2 * sin(0.6 * x[i]) * cos(0.4 * y[i]);
        const double u = 0.5 + 0.1 * cos(0.3 * y[i]), v = -0.2 + 0.1 * sin(0.5 * x[i]), p = 1.0 + 0.1 * rho;
        q[i] = rho;
        q[K * Np + i] = rho * u;
        q[2 * K * Np + i] = rho * v;
        q[3 * K * Np + i] = p / 0.4 + 0.5 * rho * (u * u + v * v);
    }
    int fq, fr;
    CK(dgh_field_create(ctx, mesh, 4, &fq));
    CK(dgh_field_create(ctx, mesh, 4, &fr));
    CK(dgh_field_upload(ctx, fq, q));
    dgh_rhs_desc d = {DGH_EQ_EULER, DGH_FLUX_ROE, DGH_PREC_FP64, 0, 0.0, 0.0, 1.4};
    CK(dgh_rhs(ctx, mesh, &d, fq, fr));
    CK(dgh_field_download(ctx, fr, r));
    printf("%d %d\n", K, Np);
    for (int i = 0; i < 4 * K * Np; ++i) printf("%.17g\n", r[i]);
    free(x);
    free(y);
    free(q);
    free(r);
    return dgh_destroy(ctx) == DGH_OK ? 0 : 1;
}
