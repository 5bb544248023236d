/* Compiled against include/dgh.h: prints the size and field offsets of every descriptor and statistics struct of
 * the C ABI as JSON, so that the Python binding's ctypes mirrors can be checked against the header; and, with the
 * library linked in, exercises the host-only entry point dgh_refelem_get. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "dgh.h"

#define FIELD(T, f) printf("%s\"%s\": [%zu, %zu]", first++ ? ", " : "", #f, offsetof(T, f), sizeof(((T*)0)->f))
#define BEGIN(T) do { int first = 0; printf("%s\"%s\": {\"size\": %zu, \"fields\": {", nstruct++ ? ", " : "", #T, sizeof(T));
#define END printf("}}"); } while (0)

int main(int argc, char** argv) {
    int nstruct = 0;
    printf("{\"structs\": {");
    BEGIN(dgh_rhs_desc); FIELD(dgh_rhs_desc, eq); FIELD(dgh_rhs_desc, flux); FIELD(dgh_rhs_desc, prec);
    FIELD(dgh_rhs_desc, reserved); FIELD(dgh_rhs_desc, ax); FIELD(dgh_rhs_desc, ay); FIELD(dgh_rhs_desc, gamma); END;
    BEGIN(dgh_lim_desc); FIELD(dgh_lim_desc, kind); FIELD(dgh_lim_desc, detect_only); FIELD(dgh_lim_desc, M);
    FIELD(dgh_lim_desc, nu); FIELD(dgh_lim_desc, mlp); FIELD(dgh_lim_desc, positivity); FIELD(dgh_lim_desc, threshold);
    FIELD(dgh_lim_desc, gamma); FIELD(dgh_lim_desc, eps); END;
    BEGIN(dgh_mlp_train_desc); FIELD(dgh_mlp_train_desc, epochs); FIELD(dgh_mlp_train_desc, batch);
    FIELD(dgh_mlp_train_desc, lr); FIELD(dgh_mlp_train_desc, pos_weight); FIELD(dgh_mlp_train_desc, seed);
    FIELD(dgh_mlp_train_desc, reserved); END;
    BEGIN(dgh_sipg_desc); FIELD(dgh_sipg_desc, penalty); FIELD(dgh_sipg_desc, pc); FIELD(dgh_sipg_desc, mode);
    FIELD(dgh_sipg_desc, maxit); FIELD(dgh_sipg_desc, rtol); FIELD(dgh_sipg_desc, abstol);
    FIELD(dgh_sipg_desc, inner_rtol); FIELD(dgh_sipg_desc, max_outer); FIELD(dgh_sipg_desc, reserved);
    FIELD(dgh_sipg_desc, mass_shift); END;
    BEGIN(dgh_sipg_stats); FIELD(dgh_sipg_stats, iters); FIELD(dgh_sipg_stats, outer); FIELD(dgh_sipg_stats, reason);
    FIELD(dgh_sipg_stats, reserved); FIELD(dgh_sipg_stats, rnorm0); FIELD(dgh_sipg_stats, rnorm);
    FIELD(dgh_sipg_stats, relres); FIELD(dgh_sipg_stats, ms); END;
    BEGIN(dgh_ens_case); FIELD(dgh_ens_case, N); FIELD(dgh_ens_case, nx); FIELD(dgh_ens_case, ny);
    FIELD(dgh_ens_case, x0); FIELD(dgh_ens_case, y0); FIELD(dgh_ens_case, Lx); FIELD(dgh_ens_case, Ly);
    FIELD(dgh_ens_case, ax); FIELD(dgh_ens_case, ay); FIELD(dgh_ens_case, nu); FIELD(dgh_ens_case, dt); END;
    BEGIN(dgh_ens_desc); FIELD(dgh_ens_desc, tau); FIELD(dgh_ens_desc, bx); FIELD(dgh_ens_desc, by);
    FIELD(dgh_ens_desc, autotune); FIELD(dgh_ens_desc, reserved); END;
    BEGIN(dgh_bcfn_desc); FIELD(dgh_bcfn_desc, kind); FIELD(dgh_bcfn_desc, reserved); FIELD(dgh_bcfn_desc, gamma);
    FIELD(dgh_bcfn_desc, x0); FIELD(dgh_bcfn_desc, y0); FIELD(dgh_bcfn_desc, beta); FIELD(dgh_bcfn_desc, u_inf);
    FIELD(dgh_bcfn_desc, v_inf); END;
    printf("}, \"enums\": {\"DGH_ERR_NOT_CONVERGED\": %d, \"DGH_FLUX_AVERAGE\": %d, \"DGH_PREC_BF16X3\": %d, "
           "\"DGH_BC_WALL\": %d, \"DGH_LIM_TRIANGLE_MLP\": %d, \"DGH_SIPG_PENALTY_H\": %d, \"DGH_PC_BLOCK_JACOBI\": %d, "
           "\"DGH_SOLVE_MIXED\": %d, \"DGH_MLP_BF16\": %d, \"DGH_BCFN_DOUBLE_MACH\": %d}",
           DGH_ERR_NOT_CONVERGED, DGH_FLUX_AVERAGE, DGH_PREC_BF16X3, DGH_BC_WALL, DGH_LIM_TRIANGLE_MLP,
           DGH_SIPG_PENALTY_H, DGH_PC_BLOCK_JACOBI, DGH_SOLVE_MIXED, DGH_MLP_BF16, DGH_BCFN_DOUBLE_MACH);
#ifdef WITH_LIBRARY
    /* host-only entry point: the order-3 mass matrix (10 x 10) through the library */
    double M[100];
    const int n = dgh_refelem_get(3, "M", M, 100);
    double tr = 0.0;
    for (int i = 0; i < 10 && n == 100; ++i) tr += M[i * 11];
    printf(", \"refelem_M3_count\": %d, \"refelem_M3_trace\": %.17g", n, tr);
#endif
    printf("}\n");
    (void)argc;
    (void)argv;
    return 0;
}
