/*
 * dgh.h -- C ABI of libdghip.so, the MI355X (gfx950) implementation of HopeFOAM's 2D
 * triangle DG hot path: explicit right-hand side with fused low-storage RK stage, and the
 * implicit SIPG (dgm::laplacian) solve.
 *
 * Conventions
 *   - Every call returns DGH_OK (0) or a negative dgh_status; the message of the last
 *     failure on a context is available from dgh_last_error().  Nothing throws across the ABI.
 *   - Objects (meshes, fields) are integer handles owned by the context.
 *   - Host field arrays are FP64 in the reference's DOF order [var][K][Np] (element-major,
 *     one block per variable; HopeFOAM-0.1/src/DG/fields/GeometricDofField/GeometricDofField.H:81-116).
 *   - All device work is queued on the context's single HIP stream; calls that return data to
 *     the host synchronise that stream.
 *
 * The reference exposes this path through per-cell C++ virtuals selected from the case
 * dictionary; each entry point below names the reference interface it replaces.
 */
#ifndef DGH_H
#define DGH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgh_ctx_s* dgh_ctx;

enum dgh_status {
    DGH_OK = 0,
    DGH_ERR_ARG = -1,          /* bad argument or handle */
    DGH_ERR_HIP = -2,          /* HIP runtime error */
    DGH_ERR_UNSUPPORTED = -3,  /* order / scheme / precision not available */
    DGH_ERR_NOMEM = -4,
    DGH_ERR_NOT_CONVERGED = -5
};

enum dgh_eq { DGH_EQ_ADVECT = 0, DGH_EQ_EULER = 1 };
/* flux words of the case dictionary: upwind == the reference "LF" scheme for constant velocity
 * (LFFlux.C:45-100), LF == local Lax-Friedrichs (Rusanov) for Euler, Roe == RoeFlux.C:46-191,
 * average == the reference "average" scheme (averageFlux.C:45-90, 1/2 (n.a qO + n.a qN), advection only): with
 * a = (1, 0) or (0, 1) the right-hand side is minus the DG x or y derivative with central fluxes, the
 * dgc::grad(p) and dgc::div(U) of the projection solver's gradSchemes / divSchemes "default average" */
enum dgh_flux { DGH_FLUX_UPWIND = 0, DGH_FLUX_LF = 1, DGH_FLUX_ROE = 2, DGH_FLUX_AVERAGE = 3 };
/* FP64: everything in FP64.  FP32: contractions and fluxes in FP32.  BF16X3: fluxes in FP32, the three
 * contractions (cubature interpolation, projection, lift) as bf16x3 split products on the bf16 matrix cores
 * (x = hi + lo in bf16; a b = a_lo b_hi + a_hi b_lo + a_hi b_hi, FP32 accumulation; about 2^-16 relative per
 * product).  In every precision the LSERK4 registers (q, res) stay FP64. */
enum dgh_prec { DGH_PREC_FP64 = 0, DGH_PREC_FP32 = 1, DGH_PREC_BF16X3 = 2 };
/* boundary tags for faces without a neighbour */
enum dgh_bc { DGH_BC_DEFAULT = 0, DGH_BC_FIXED_VALUE = 1, DGH_BC_ZERO_GRADIENT = 2, DGH_BC_WALL = 3 };

typedef struct {
    int eq;        /* dgh_eq */
    int flux;      /* dgh_flux */
    int prec;      /* dgh_prec: arithmetic of contractions and fluxes; RK registers stay FP64 */
    int reserved;
    double ax, ay; /* advection velocity (DGH_EQ_ADVECT) */
    double gamma;  /* ratio of specific heats (DGH_EQ_EULER) */
} dgh_rhs_desc;

/* ---- context ------------------------------------------------------------------------- */
/* replaces the process-wide PETSc/OpenFOAM set-up of dgMesh.C:82-105 */
int dgh_create(dgh_ctx* ctx, int device);
int dgh_destroy(dgh_ctx ctx);
const char* dgh_last_error(dgh_ctx ctx);
int dgh_sync(dgh_ctx ctx);
/* the context's HIP stream (hipStream_t), for callers that time or order work on it */
void* dgh_stream(dgh_ctx ctx);

/* ---- mesh -------------------------------------------------------------------------- */
/* dgMesh + physicalElementData construction (dgMesh.C:71-116, physicalElementData.C:60-161,
 * physicalCellElement.C:72-110).  vxy: K x 3 x 2 anticlockwise vertex coordinates (unwrapped
 * across periodic boundaries); vid: K x 3 vertex ids (periodic images share an id);
 * bctag: K x 3 dgh_bc for faces without a neighbour, or NULL. */
int dgh_mesh_create(dgh_ctx ctx, int N, int K, const double* vxy, const int* vid, const int* bctag, int* mesh);
/* nx x ny quads on [x0,x1]x[y0,y1], each split into two triangles (the benchmark meshes);
 * non-periodic sides get bc_tag. */
int dgh_mesh_structured(dgh_ctx ctx, int N, int nx, int ny, double x0, double x1, double y0, double y1,
                        int periodic_x, int periodic_y, int bc_tag, int* mesh);
/* info[7] = {N, Np, Nfp, Nq, Nfq, K, boundary faces} */
int dgh_mesh_info(dgh_ctx ctx, int mesh, int* info);
int dgh_mesh_nodes(dgh_ctx ctx, int mesh, double* x, double* y);      /* K x Np each */
int dgh_mesh_conn(dgh_ctx ctx, int mesh, int* EToE, int* EToF);       /* K x 3 each, -1 on boundary */
/* K x 14: rx ry sx sy J, then per face nx ny Js */
int dgh_mesh_geo(dgh_ctx ctx, int mesh, double* geo);
/* fixedValue ghost states: nvar x K x 3 x Nfp nodal values in the owner's face-node order
 * (setBoundaryValues.H:9-50 + dgGaussField.C:245-267); NULL clears them */
int dgh_mesh_set_bc_values(dgh_ctx ctx, int mesh, int nvar, const double* bcq);

/* fixedValue ghost states as a function of time, evaluated on the device at the fixedValue face nodes: the
 * reference's setBoundaryValues.H of the Euler tutorials, which the driver calls at the start of every time
 * step with t^n (dgEulerFoam.C:73).  Sets the function and the mesh's time t and evaluates the ghost states at t
 * (4 variables; replaces data given by dgh_mesh_set_bc_values, which in turn clears the function).  The run
 * loops then re-evaluate them on the device: dgh_ssprk2_run at t^n before each step (both stages see t^n, as in
 * the reference), dgh_lserk4_run(_limited) at t^n + c_s dt before each stage; they advance the mesh's time by
 * nsteps dt and leave the ghost states evaluated at the new time.  The per-stage entries use the ghost states
 * as they are: call this again with the next time between steps.  kind DGH_BCFN_NONE (or fn NULL) removes the
 * function and only sets the time. */
enum dgh_bcfn {
    DGH_BCFN_NONE = 0,
    /* isentropicVortex/dgEulerFoam/setBoundaryValues.H:20-46: centre (x0 + u_inf t, y0 + v_inf t), strength
     * beta, r = squared distance; the tutorial: x0 = 5, y0 = 0, beta = 5, u_inf = 1, v_inf = 0 */
    DGH_BCFN_ISENTROPIC_VORTEX = 1,
    /* doubleMach/dgEulerFoam/setBoundaryValues.H:18-60: the Mach 10 shock through (1/6, 0), position
     * (x - 1/6) (sqrt 3 / (1 + 20 t)) - y; behind it (position < 0) rho 8, rho U = 66 (cos 30, -sin 30), p 116.5,
     * ahead rho 1.4, U 0, p 1 */
    DGH_BCFN_DOUBLE_MACH = 2
};
typedef struct {
    int kind;                            /* dgh_bcfn */
    int reserved;
    double gamma;
    double x0, y0, beta, u_inf, v_inf;   /* DGH_BCFN_ISENTROPIC_VORTEX */
} dgh_bcfn_desc;
int dgh_mesh_set_bc_function(dgh_ctx ctx, int mesh, const dgh_bcfn_desc* fn, double t);
/* patch structure for the Triangle limiter's fixedValue ghost means: first[e 3 + f] = e' 3 + f', the first face
 * of the patch face (e, f) belongs to (its first face node's value is the patch's first value, which the
 * reference takes as the ghost mean of every face of the patch, Trianglelimite.C:220-223), -1 for other faces;
 * NULL (the default) uses each face's own first node.  DGH_ERR_ARG when an entry is not a boundary face. */
int dgh_mesh_set_patch_first(dgh_ctx ctx, int mesh, const int* first);
/* the mesh's time (advanced by the run loops) */
int dgh_mesh_get_time(dgh_ctx ctx, int mesh, double* t);

/* reference-element operators of order N for inspection (host only, no GPU needed):
 * name in {r, s, Dr, Ds, V, M, Vq, Drq, Dsq, cw, Vf, gw, Pr, Ps, Lift}; returns the number of
 * values written (at most cap) or a negative status. */
int dgh_refelem_get(int N, const char* name, double* out, int cap);

/* ---- fields -------------------------------------------------------------------------- */
/* device-resident DOF fields (GeometricDofField<scalar|vector>, GeometricDofField.H:81-116) */
int dgh_field_create(dgh_ctx ctx, int mesh, int nvar, int* field);
int dgh_field_destroy(dgh_ctx ctx, int field);
int dgh_field_upload(dgh_ctx ctx, int field, const double* host);
int dgh_field_download(dgh_ctx ctx, int field, double* host);
int dgh_field_copy(dgh_ctx ctx, int dst, int src);

/* ---- explicit right-hand side ------------------------------------------------------------
 * rhs = M^-1 [ (grad phi, F(q))_cubature - <phi, F*.n>_faces ]: the whole-mesh replacement of
 * convectionScheme::dgcDivCalculateCell (convectionScheme.H:141-163) + fluxSchemes::evaluateFlux
 * (fluxScheme.H:141-153) / fluxCalcScheme::fluxCalculateWeak (fluxCalcScheme.H:140-158) +
 * dg::solveEquation's mass solve (Equation.C:42-79). */
int dgh_rhs(dgh_ctx ctx, int mesh, const dgh_rhs_desc* d, int q, int rhs);
/* one stage of Carpenter-Kennedy LSERK4 fused into the RHS kernel (the coefficients of
 * createFields.H:119-138): res = a_s res + dt rhs(q_in); q_out = q_in + b_s res.
 * q_in, q_out and res must be three different fields (DGH_ERR_ARG otherwise). */
int dgh_lserk4_stage(dgh_ctx ctx, int mesh, const dgh_rhs_desc* d, int q_in, int q_out, int res, int stage,
                     double dt);
/* one stage of the reference's SSP-RK2 (dgEulerFoam.C:64-123) on the same fused kernel:
 * stage 0: res = dt rhs(q_in), q_out = q_in + res;
 * stage 1: q_out = (q_out + q_in + dt rhs(q_in)) / 2, i.e. q_out must hold q^n on entry (the reference's
 *          rho = 0.5 rho + 0.5 rho1); res is scratch.
 * Run stage 0 from q^n into q1, (limit q1,) then stage 1 from q1 into the q^n field, (limit it).
 * q_in, q_out and res must be three different fields (DGH_ERR_ARG otherwise). */
int dgh_ssprk2_stage(dgh_ctx ctx, int mesh, const dgh_rhs_desc* d, int q_in, int q_out, int res, int stage,
                     double dt);
/* nsteps full LSERK4 steps on q (res is zeroed first); *ms = device time of the loop measured
 * with HIP events on the context stream (may be NULL). */
int dgh_lserk4_run(dgh_ctx ctx, int mesh, const dgh_rhs_desc* d, int q, int nsteps, double dt, float* ms);

/* ---- limiter ----------------------------------------------------------------------------
 * The whole-field replacement of limiteSchemes::limite(rho, rhoU, E, ... [, troubledCell])
 * (limiteScheme.H:139-159), two schemes:
 *  - TVB minmod for triangles (Cockburn & Shu 1998), component-wise on the conserved variables, with the
 *    troubled-cell test as detection step (not in the reference);
 *  - the reference's Triangle scheme (Trianglelimite::limite, Trianglelimite.C:61-852): every cell rebuilt
 *    as a P1 field from nonlinearly weighted Green-Gauss gradients of rho, u, v, p, with the density floor
 *    eps (the reference's tol = 1e-2) and gamma (1.4 there); optionally only the cells a detection step
 *    flags (the troubledCell overload, Trianglelimite.C:45-57, empty in the reference).
 * Boundary ghosts: fixedValue data (the mesh's bc values, when set for this variable count), reflective wall
 * (momentum mirrored; the Triangle scheme's ghost mean keeps the tangential part, as the reference), else a
 * copy.  Triangle ghost means on fixedValue faces take the patch's first value as the reference does when the
 * mesh has a patch map (dgh_mesh_set_patch_first), else the data at the face's own first node. */
enum dgh_lim_kind {
    DGH_LIM_TVB = 0,            /* TVB troubled-cell test decides */
    DGH_LIM_MLP = 1,            /* the learned indicator decides (dgh_mlp_*; 4 variables, N <= 6) */
    DGH_LIM_ALL = 2,            /* no detection step: every cell is limited */
    DGH_LIM_TRIANGLE = 3,       /* the reference's Triangle scheme on every cell (4 variables) */
    DGH_LIM_TRIANGLE_TVB = 4,   /* Triangle scheme on the cells the TVB test flags */
    DGH_LIM_TRIANGLE_MLP = 5    /* Triangle scheme on the cells the learned indicator flags (N <= 6) */
};
typedef struct {
    int kind;          /* dgh_lim_kind */
    int detect_only;   /* 1: flag troubled cells only, q unchanged */
    double M;          /* TVB constant: deviations up to M h^2 (h = longest edge) are left alone */
    double nu;         /* minmod factor on the neighbour differences (1.5 in Cockburn-Shu) */
    int mlp;           /* DGH_LIM_MLP: network handle */
    int positivity;    /* 1 (4 variables): Zhang-Shu scaling keeps rho, p >= eps at every evaluation point */
    double threshold;  /* DGH_LIM_MLP: a cell is troubled when its predicted probability >= threshold */
    double gamma;      /* positivity and Triangle: ratio of specific heats */
    double eps;        /* positivity: floor (the cell's own mean density and pressure if smaller);
                          Triangle: the density floor of the reconstruction (positivity then uses it too) */
} dgh_lim_desc;
/* limits q (1 or 4 variables) in place; troubled: K flags (host, may be NULL); ntroubled: number of
 * troubled (Triangle: rebuilt) cells (may be NULL; when both are NULL the call is asynchronous).
 * Triangle kinds: troubled[e] bit 0 = rebuilt, bit 1 = the density floor shortened the slope at a node,
 * bit 2 = a node's pressure is below eps (the reference's two warnings). */
int dgh_limit(dgh_ctx ctx, int mesh, const dgh_lim_desc* d, int q, int* troubled, int* ntroubled);
/* dgh_limit without device allocation.  No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The same kernels and results
 * as dgh_limit for the kinds DGH_LIM_TVB, DGH_LIM_MLP and DGH_LIM_ALL; the moments, flags and counters live in a
 * static device array of the library (16.5 MiB per GPU, held until the process ends) instead of per-mesh
 * buffers.  Limits: K <= 131072 cells (DGH_ERR_ARG above), the Triangle kinds are refused (DGH_ERR_UNSUPPORTED),
 * and the mesh's limiter tables must already exist (built by the first dgh_limit call on the mesh, which
 * allocates; DGH_ERR_ARG otherwise).  Every call on one GPU shares that array: calls that overlap on one GPU
 * (other contexts or threads) race.  With troubled and ntroubled NULL the call is asynchronous. */
int dgh_limit_noalloc(dgh_ctx ctx, int mesh, const dgh_lim_desc* d, int q, int* troubled, int* ntroubled);
/* dgh_lserk4_run with the limiter ld applied to every stage's new state, as the reference's Euler driver
 * calls Godunov.limite after each Runge-Kutta stage (dgEulerFoam.C:92, :119); the same kernels as one
 * dgh_lserk4_stage + dgh_limit per stage, without host round trips.  ld must not be detect_only; an invalid
 * ld is reported at the first stage, q then unchanged. */
int dgh_lserk4_run_limited(dgh_ctx ctx, int mesh, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int q, int nsteps,
                           double dt, float* ms);
/* nsteps steps of the reference's Euler driver (dgEulerFoam.C:64-123): SSP-RK2 as two dgh_ssprk2_stage calls,
 * each new state followed by the limiter ld (Godunov.limite, dgEulerFoam.C:92, :119; NULL: none), inside the
 * library; q advances in place.  The same kernels, in the same order, as the per-stage calls.  ld must not be
 * detect_only; an invalid ld is reported at the first stage, q then unchanged. */
int dgh_ssprk2_run(dgh_ctx ctx, int mesh, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int q, int nsteps, double dt,
                   float* ms);

/* ---- learned troubled-cell indicator -----------------------------------------------------
 * No counterpart in the reference: the detection step of limiteSchemes::limite (the troubledCell
 * argument of limiteScheme.H:139-159) as a small MLP, inputs -> 64 -> 64 -> 1 (ReLU, sigmoid), trained on
 * the GPU in FP32 and evaluated in BF16 on MFMA inside the limiter pass.  Inputs per cell (2 Np + 8,
 * orders N <= 6): the modal coefficients of rho and E divided by |mean mode|, the jumps to the three
 * neighbour (or ghost) means of rho and E divided by |own mean|, and for rho and E the TVB scale
 * log10(max edge-midpoint deviation / (M h^2) + 1e-30) clipped to [-10, 10]. */
typedef struct {
    int epochs;
    int batch;          /* samples per Adam step */
    float lr;           /* Adam learning rate */
    float pos_weight;   /* loss weight of troubled samples */
    unsigned seed;      /* shuffling */
    int reserved;
} dgh_mlp_train_desc;
/* network for order N, parameters He-uniform from seed */
int dgh_mlp_create(dgh_ctx ctx, int N, unsigned seed, int* mlp);
int dgh_mlp_destroy(dgh_ctx ctx, int mlp);
/* info[3] = {N, inputs, parameters}; parameters are W1 [64][in], b1, W2 [64][64], b2, w3 [64], b3 */
int dgh_mlp_info(dgh_ctx ctx, int mlp, int* info);
int dgh_mlp_get_params(dgh_ctx ctx, int mlp, float* params);
int dgh_mlp_set_params(dgh_ctx ctx, int mlp, const float* params);
/* input features of every cell of q (4 variables) for TVB constant M: X [K][inputs], host */
int dgh_mlp_features(dgh_ctx ctx, int mesh, int q, double M, float* X);
enum dgh_mlp_prec { DGH_MLP_FP32 = 0, DGH_MLP_BF16 = 1 };
/* logits of n samples X [n][inputs]: DGH_MLP_FP32 (the training arithmetic) or DGH_MLP_BF16 (the MFMA
 * inference of the limiter pass) */
int dgh_mlp_forward(dgh_ctx ctx, int mlp, const float* X, int n, int prec, float* z);
/* loss and gradient of the weighted cross-entropy on n samples (no update) */
int dgh_mlp_grad(dgh_ctx ctx, int mlp, const float* X, const float* y, int n, float pos_weight, float* grad,
                 float* loss);
/* Adam training on n samples; epoch_loss[epochs] (may be NULL) = mean batch loss of each epoch */
int dgh_mlp_train(dgh_ctx ctx, int mlp, const float* X, const float* y, int n, const dgh_mlp_train_desc* d,
                  float* epoch_loss);

/* ---- implicit SIPG solve: dgm::laplacian(u) == f ----------------------------------------
 * Matrix-free replacement of defaultLaplacianScheme::laplacianFunc (defaultLaplacianScheme.C:62-288,
 * selected as laplacianSchemes{default default 20;}, defaultLaplacianScheme.H:94-108), of the PETSc
 * assembly in dgMatrix::{petscPreAllocate,assembleMatrix,finalAssemble} (dgMatrix.C:296-446) and
 * of the KSP solve (dgMatrix::kspSolver, dgMatrix.C:163-283; dgScalarMatrix::solve,
 * dgScalarMatrix.C:67-188).  A is the reference's negative SIPG matrix; b = M f + Dirichlet terms
 * (EquationGeometricDofField.H:84-102). */
enum dgh_sipg_penalty {
    DGH_SIPG_PENALTY_REFERENCE = 0, /* gtau = fscale * 20 (N+1)^2 (defaultLaplacianScheme.C:114) */
    DGH_SIPG_PENALTY_H = 1          /* tau = (N+1)^2 / h, h = 1 / fscale */
};
enum dgh_pc { DGH_PC_JACOBI = 0, DGH_PC_BLOCK_JACOBI = 1 };   /* kspPC jacobi | element block Jacobi */
enum dgh_solve_mode {
    DGH_SOLVE_FP64 = 0,  /* PETSc KSPCG semantics: preconditioned norm, rtol relative to the initial residual */
    DGH_SOLVE_MIXED = 1  /* FP32 PCG corrections inside FP64 iterative refinement; rtol on |b - A u| / |b| */
};

typedef struct {
    int penalty;        /* dgh_sipg_penalty */
    int pc;             /* dgh_pc */
    int mode;           /* dgh_solve_mode */
    int maxit;          /* maxIter (mixed: total inner iterations) */
    double rtol;        /* relTol */
    double abstol;      /* tolerance */
    double inner_rtol;  /* mixed: reduction asked of each FP32 correction solve (0: 1e-3) */
    int max_outer;      /* mixed: refinement sweeps (0: 20) */
    int reserved;
    double mass_shift;  /* sigma >= 0: the operator becomes A + sigma M (M the mass matrix with the element
                           Jacobians), the implicit step of dgChorinFoam's viscous term
                           Sp(U) g0/(dt nu) == laplacian(U) + ... (dgChorinFoam.C:99-103); 0: Poisson */
} dgh_sipg_desc;

typedef struct {
    int iters;      /* CG iterations (FP64) or the sum of the FP32 correction iterations (mixed) */
    int outer;      /* refinement sweeps (mixed), 0 otherwise */
    int reason;     /* > 0 converged; < 0 stopped: -3 iteration limit, -4 diverged, -8 indefinite, -9 not finite */
    int reserved;
    double rnorm0;  /* FP64: initial preconditioned residual norm; mixed: initial |b - A u| */
    double rnorm;   /* FP64: final preconditioned residual norm; mixed: final |b - A u| */
    double relres;  /* final |b - A u|_2 / |b|_2, evaluated in FP64 */
    double ms;      /* device time of the solve (HIP events on the context stream) */
} dgh_sipg_stats;

/* operator + preconditioner set-up on a mesh (faces tagged DGH_BC_FIXED_VALUE are Dirichlet, other
 * boundary faces homogeneous Neumann) */
int dgh_sipg_create(dgh_ctx ctx, int mesh, const dgh_sipg_desc* d, int* sipg);
int dgh_sipg_destroy(dgh_ctx ctx, int sipg);
/* info[4] = {N, K, Np, distinct block-Jacobi blocks} */
int dgh_sipg_info(dgh_ctx ctx, int sipg, int* info);
/* b = M f + Dirichlet terms of g; gbc: K x 3 x Nfp nodal boundary values in the owner's face-node order,
 * or NULL for g = 0 (scalar fields) */
int dgh_sipg_rhs(dgh_ctx ctx, int sipg, int f, const double* gbc, int b);
/* dgh_sipg_rhs with normal-derivative data gnb (K x 3 x Nfp nodal values in the owner's face-node order, or
 * NULL) on the boundary faces that are not Dirichlet: the fixedGradient patches of defaultLaplacianScheme.C:210-220,
 * b -= Vf^T W Vf g per face (W the face Gauss weights times the edge Jacobian) */
int dgh_sipg_rhs_neumann(dgh_ctx ctx, int sipg, int f, const double* gbc, const double* gnb, int b);
/* out = A u, in DGH_PREC_FP64 or DGH_PREC_FP32 arithmetic */
int dgh_sipg_apply(dgh_ctx ctx, int sipg, int prec, int u, int out);
/* solves A u = b; u holds the initial guess (KSPSetInitialGuessNonzero, dgMatrix.C:205).
 * Returns DGH_ERR_NOT_CONVERGED (with stats filled) when the iteration stops unconverged. */
int dgh_sipg_solve(dgh_ctx ctx, int sipg, int b, int u, dgh_sipg_stats* stats);

/* ---- operators of the incompressible projection solver ------------------------------------
 * The explicit pieces of dgChorinFoam (tutorials/DG/2D/pearsonVortex/dgChorinFoam/dgChorinFoam.C:63-109) besides
 * its SIPG solves; U and convecU are 2-variable fields (the z components of the reference's vectors are zero).
 * Convection: dg::solveEquation(dgm::Sp(convecU) == dgc::div(U, U)) with the tutorial's "div(U,U) default LF":
 * the dof-field form of defaultConvectionScheme.C:215-303 (the volume flux interpolated from the nodal products
 * U_k U_c) with the Lax-Friedrichs face flux of LFFlux.C:103-211 at the face nodes (one dissipation speed per
 * face: the largest |n.U| of either side), followed by the mass solve: out = M^-1 [<phi, F*> - (grad phi, U U)].
 * Boundary faces take the mesh's fixedValue ghost states when they are set for 2 variables, else the own trace
 * (zeroGradient and fixedGradient patches evaluate to the internal values).  Orders 1..6. */
int dgh_ins_convection(dgh_ctx ctx, int mesh, int U, int out);
/* dgc::grad / dgc::div of a dof field (defaultGrad.C:45-84, defaultDivScheme.C:45-83): element-local strong
 * derivatives at the nodes, out = grad u (u 1 variable, out 2) or div u (u 2 variables, out 1); with Sp on the
 * other side of the equation they are the nodal values the solver uses (dgChorinFoam.C:82, :89) */
int dgh_nodal_grad(dgh_ctx ctx, int mesh, int u, int out);
int dgh_nodal_div(dgh_ctx ctx, int mesh, int u, int out);
/* out = sum_t coef[t] fields[t], 1 <= n <= 8 fields of the same size (the solver's field algebra, e.g.
 * UT = (a0 U^n + a1 U^(n-1) - dt (b0 convecU + b1 convecUold)) / g0, dgChorinFoam.C:76); out may be one of them */
int dgh_field_lincomb(dgh_ctx ctx, int out, int n, const int* fields, const double* coef);
/* variable dvar of dst <- variable svar of src: fields of the same element count and order, possibly on different
 * meshes (e.g. a vector component into the scalar field of a per-component solve) */
int dgh_field_copy_var(dgh_ctx ctx, int dst, int dvar, int src, int svar);

/* ---- ensembles of LDG advection-diffusion cases ----------------------------------------
 * No counterpart in the reference (its dgc::laplacian is an empty stub, defaultLaplacianScheme.C:354-359):
 * many independent cases u_t + div(a u) = nu div(grad u) on periodic split-quad meshes of different
 * sizes and orders, advanced together by LSERK4 (the coefficients of createFields.H:119-138).  The method
 * is LDG (Cockburn & Shu 1998) with alternating fluxes (u^ from the side the switch vector (bx, by) points
 * away from, q^ from the other side, penalty tau max(Js/J) on u jumps), upwind advection, nodal strong
 * form.  Each stage is two passes (q, then the right-hand side with the stage update), each one persistent
 * launch per order present over all cases of that order; dgh_ensemble_run can capture the whole time loop
 * into one HIP graph.  The autotuner picks, per order, the tile shape: how many waves share a 64-element
 * block's rows (1, 2, 4, 8; one element per lane), or 16x16x4 FP64 MFMA tiles on 16-element groups (16). */
typedef struct {
    int N;                 /* order 1..8 */
    int nx, ny;            /* nx x ny quads split into 2 nx ny triangles, periodic in x and y (nx, ny >= 3) */
    double x0, y0, Lx, Ly; /* the box [x0, x0 + Lx] x [y0, y0 + Ly] */
    double ax, ay;         /* velocity */
    double nu;             /* diffusivity (>= 0) */
    double dt;             /* time step of this case */
} dgh_ens_case;
typedef struct {
    double tau;            /* LDG penalty factor (>= 0) */
    double bx, by;         /* switch vector of the alternating fluxes */
    int autotune;          /* 1: time the row splits of every order at creation and keep the fastest */
    int reserved;
} dgh_ens_desc;
int dgh_ensemble_create(dgh_ctx ctx, const dgh_ens_desc* d, int ncases, const dgh_ens_case* cases, int* ens);
int dgh_ensemble_destroy(dgh_ctx ctx, int ens);
/* state of case i, host order [K][Np] (element e = 2 (j nx + i) + t of the split quads, reference node order) */
int dgh_ensemble_set(dgh_ctx ctx, int ens, int i, const double* u);
/* what: 0 the state, 1 the right-hand side of the last dgh_ensemble_rhs, 2 q = (qx, qy), 2 x K x Np */
int dgh_ensemble_get(dgh_ctx ctx, int ens, int i, int what, double* out);
/* right-hand side of every case (kept for dgh_ensemble_get) */
int dgh_ensemble_rhs(dgh_ctx ctx, int ens);
/* nsteps LSERK4 steps of every case; use_graph: the loop runs as one captured HIP graph (built on first use
 * for this step count and kept); *ms = device time (HIP events on the context stream), may be NULL */
int dgh_ensemble_run(dgh_ctx ctx, int ens, int nsteps, int use_graph, float* ms);
/* split[9]: tile shape of each order (1, 2, 4, 8: waves per 64-element block; 16: MFMA tiles; 0: order
 * absent); tune_us[9 x 5]: autotuner time per stage of each order's cases for 1, 2, 4, 8 waves and MFMA tiles
 * (0: not measured or not available); dofs: sum of K Np; nitems: work items.  Any pointer may be NULL. */
int dgh_ensemble_info(dgh_ctx ctx, int ens, int* split, double* tune_us, long long* dofs, int* nitems);

/* ---- diagnostics ---------------------------------------------------------------------- */
/* caps the persistent grid of the stage kernel at blocks (rounded down to a multiple of 8, at least 8; 0: as many
 * as the device keeps resident), so that every wavefront walks several element groups even on small meshes;
 * results do not depend on it */
int dgh_set_launch_cap(dgh_ctx ctx, int blocks);
/* runs one 16x16x4 MFMA in f64 and f32 and returns the max abs error vs the host product */
int dgh_selftest_mfma(dgh_ctx ctx, double* err64, double* err32);
/* one K16 step of bf16x3 split products (v_mfma_f32_16x16x16_bf16, host-split A, device-split B) and of the
 * hi parts alone, 16 x 16 x 16: max abs error vs the FP64 product, relative to the largest entry */
int dgh_selftest_mfma_bf16x3(dgh_ctx ctx, double* err3, double* err1);
/* raw accumulator registers of 16x16x4 MFMAs with A = lane id, B = 1 and A = 1, B = lane id,
 * f64 then f32, 4 blocks of 64 lanes x 4 registers (1024 values) */
int dgh_selftest_mfma_layout(dgh_ctx ctx, double* raw);

#ifdef __cplusplus
}
#endif
#endif
