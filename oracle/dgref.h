/*
 * dgref — FP64 CPU restatement of HopeFOAM-0.1's 2D triangle DG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP implementation in hopefoam_amd/ and the CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the checker.  Nothing in the product links or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to HopeFOAM-0.1/).  Layouts: nodal fields are [var][K][Np]
 * (the reference's element-major [K][Np] order, one block per variable).
 */
#ifndef DGREF_H
#define DGREF_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1D / polynomial building blocks (src/DG/element/polynomials/Legendre/Legendre.C) */
double or_jacobi_p(double x, double alpha, double beta, int n);
double or_grad_jacobi_p(double x, double alpha, double beta, int n);
int or_jacobi_gq(double alpha, double beta, int n, double* x, double* w);   /* n+1 points */
int or_jacobi_gl(double alpha, double beta, int n, double* x);              /* n+1 points */

/* ---- reference triangle (triangleBaseFunction.C, gaussTriangleIntegration.C)
 * All operators are formed in long double and rounded once to FP64. */
int or_np(int N);
int or_nq(int N);                    /* cubature points of order 3(N+1) */
int or_nodes2d(int N, double* r, double* s);
int or_vandermonde2d(int N, int m, const double* r, const double* s, double* V);        /* m x Np */
int or_grad_vandermonde2d(int N, int m, const double* r, const double* s, double* Vr, double* Vs);
int or_invert(int n, const double* A, double* Ainv);                                     /* Gauss-Jordan */
int or_cubature(int order, double* r, double* s, double* w);                             /* returns npts */
int or_fmask(int N, int* fmask);                                                         /* 3 x Nfp */
int or_mass_inverse(int N, double* invM);                                                /* Np x Np */

/* all operators of one order; arrays sized by the caller from or_np/or_nq:
 * V,invV,Dr,Ds,M [Np*Np]; Vq,Drq,Dsq [Nq*Np]; cw [Nq]; Vf [Nfq*Nfp]; gw [Nfq] (Nfq = N+2) */
int or_refelem_ops(int N, double* r, double* s, double* V, double* invV, double* Dr, double* Ds,
                   double* Vq, double* Drq, double* Dsq, double* cw, double* Vf, double* gw, double* M);

/* ---- mesh (dgPolyMesh.C, physicalCellElement.C, physicalElementData.C) */
typedef struct or_mesh or_mesh;
/* vxy: K x 3 x 2 element vertex coordinates (anticlockwise, already unwrapped);
 * vid: K x 3 vertex ids used for connectivity (periodic images share ids);
 * bctag: K x 3 tag for faces without neighbour (1 fixedValue, 2 zeroGradient, 3 reflective wall) */
or_mesh* or_mesh_create(int N, int K, const double* vxy, const int* vid, const int* bctag);
void or_mesh_destroy(or_mesh* m);
int or_mesh_info(const or_mesh* m, int* out7);    /* N, Np, Nfp, Nq, Nfq, K, nbfaces */
int or_mesh_conn(const or_mesh* m, int* EToE, int* EToF);
int or_mesh_nodes(const or_mesh* m, double* x, double* y);   /* K x Np physical nodes */
int or_mesh_geo(const or_mesh* m, double* geo);              /* K x 14: rx ry sx sy J, (nx ny sJ) x 3 */

/* ---- explicit RHS (defaultConvectionScheme.C:47-129, LFFlux.C:45-100, RoeFlux.C:46-191,
 *      defaultGrad.C:86-166, defaultDivScheme.C:84-166, dgMesh.C:129-172) */
/* scalar advection with constant velocity, LF (= upwind) flux.  bcq: K x 3 x Nfp nodal ghost
 * values for fixedValue faces (may be NULL).  rhs = M^-1 (weak divergence). */
int or_rhs_advect(const or_mesh* m, const double* q, double ax, double ay, const double* bcq, double* rhs);
int or_rhs_advect_flux(const or_mesh* m, const double* q, double ax, double ay, const double* bcq, int average,
                       double* rhs);
/* 2D Euler, q = [rho, rhou, rhov, E] x K x Np.  flux: 0 = Roe (RoeFlux.C), 1 = local Lax-Friedrichs
 * (Rusanov, |u.n|+c pointwise).  bcq: 4 x K x 3 x Nfp nodal ghost states (fixedValue) or NULL. */
int or_rhs_euler(const or_mesh* m, const double* q, double gamma, int flux, const double* bcq, double* rhs);

/* time integration.  eq: 0 advection (ax, ay), 1 Euler (gamma, flux).  q updated in place. */
int or_lserk4(const or_mesh* m, int eq, double* q, int nsteps, double dt,
              double a0, double a1, int flux, const double* bcq);
int or_ssprk2(const or_mesh* m, int eq, double* q, int nsteps, double dt,
              double a0, double a1, int flux, const double* bcq);
void or_lserk4_coeffs(double* a, double* b, double* c);

/* reference isentropic-vortex driver (dgEulerFoam.C:64-131, setNonUniformInlet.H, setBoundaryValues.H,
 * eulerError.H): SSP-RK2, Roe, fixedValue exact BCs at t_n, returns mean-abs rho and |rhoU| errors. */
int or_vortex_ref_run(const or_mesh* m, double dt, int nsteps, double* err_rho, double* err_rhou);
void or_vortex_exact(int n, const double* x, const double* y, double t, double x0, double y0,
                     double beta, double gamma, double* rho, double* rhou, double* rhov, double* E);

/* ---- limiter (TVB minmod, Cockburn-Shu 1998 for triangles; not in the reference) */
/* pmean [Np] mean weights, pmid [3 x Np] P1-part values at the edge midpoints, phi [Np x 3] P1 edge functions */
int or_tvb_ops(int N, double* pmean, double* pmid, double* phi);
/* coef [K x 3 x 2] neighbour-pair coefficients, jsel [K x 3] second neighbour face, h [K] longest edge */
int or_tvb_geometry(const or_mesh* m, double* coef, int* jsel, double* h);
/* limits q ([nvar][K][Np]) in place; bcq: nvar x K x 3 x Nfp fixedValue data or NULL; troubled [K] or NULL.
 * Returns the number of troubled cells. */
int or_tvb_limit(const or_mesh* m, double* q, int nvar, double M, double nu, const double* bcq, int* troubled);
/* Zhang-Shu positivity scaling of a 4-variable Euler state q [4][K][Np] in place, checked at the nodes,
 * cubature points and face Gauss points (or_positivity_points: (Np + Nq + 3 Nfq) x Np rows); scaled [K]
 * flags or NULL; returns the number of cells scaled. */
int or_positivity_points(const or_mesh* m, double* V);
int or_positivity(const or_mesh* m, double* q, double gamma, double eps, int* scaled);

/* ---- the reference's Triangle limiter (Trianglelimite.C:61-852): weighted Green-Gauss P1 reconstruction of
 * every cell (sel[e] != 0 only, when sel is given) in primitive variables; q = [rho, rhou, rhov, E] x K x Np in
 * place; tol the density floor (1e-2 in the reference); bcq 4 x K x 3 x Nfp fixedValue data or NULL;
 * flags[K]: bit 0 rebuilt, bit 1 density floor used, bit 2 pressure below tol at a node (may be NULL) */
int or_triangle_limit(const or_mesh* m, double* q, double gamma, double tol, const double* bcq, const int* sel,
                      int* flags);
int or_triangle_limit_pf(const or_mesh* m, double* q, double gamma, double tol, const double* bcq, const int* sel,
                         const int* pfirst, int* flags);

/* ---- LDG advection-diffusion (not in the reference; Cockburn-Shu 1998 alternating fluxes, nodal strong form):
 * u_t + div(a u) = nu div(q), q = grad u, periodic meshes only.  lift [Np x 3 Nfp]; q [2][K][Np] may be NULL;
 * tau: penalty on u jumps (tau max(fscale) per face); (bx, by): the switch direction of the alternating fluxes */
int or_ldg_lift(const or_mesh* m, double* lift);
int or_ldg_rhs(const or_mesh* m, const double* u, double ax, double ay, double nu, double tau, double bx, double by,
               double* rhs, double* q);
int or_ldg_lserk4(const or_mesh* m, double* u, int nsteps, double dt, double ax, double ay, double nu, double tau,
                  double bx, double by);

/* ---- SIPG (defaultLaplacianScheme.C:62-288, dgMatrix.C:163-446, dgScalarMatrix.C:67-188) */
/* penalty_mode 0: reference gtau = fscale*20*(N+1)^2; 1: tau = (N+1)^2 / h with h = 1/fscale.
 * Assembles the reference's (negative) SIPG matrix A as CSR (row-major, K*Np rows, 4*Np nnz/row max)
 * and the right-hand side b = M f + Dirichlet terms, where f and g are nodal arrays (K x Np / K x 3 x Nfp);
 * g may be NULL (zero). */
int or_sipg_nnz(const or_mesh* m);
int or_sipg_assemble(const or_mesh* m, int penalty_mode, const double* f, const double* g,
                     int* rowptr, int* col, double* val, double* b);
/* PETSc-semantics PCG (KSPCG, PCJACOBI, KSP_NORM_PRECONDITIONED, KSPConvergedDefault) on CSR.
 * x holds the initial guess on entry.  Returns iterations; *rnorm = final preconditioned residual norm;
 * hist (maxit + 1 values, or NULL) receives the preconditioned residual norm of every iterate. */
int or_cg_jacobi(int n, const int* rowptr, const int* col, const double* val, const double* b,
                 double* x, double rtol, double abstol, int maxit, double* rnorm0, double* rnorm, double* hist);
int or_csr_matvec(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
