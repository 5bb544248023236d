// Reference-triangle operators, host side (FP64).
//
// Same element as HopeFOAM-0.1's triangleBaseFunction (warp & blend nodes,
// orthonormal Dubiner basis, triangleBaseFunction.C:46-276) with the reference's
// over-integration rule: triangle cubature of order 3(N+1) and a Gauss-Legendre
// face rule with N+2 points (gaussIntegration.C:66-68, gaussTriangleIntegration.C:45-97).
#pragma once
#include <vector>

namespace dg {

struct RefElem {
    int N = 0, Np = 0, Nfp = 0, Nq = 0, Nfq = 0;
    std::vector<double> r, s;            // Np nodes
    std::vector<double> V, invV, Dr, Ds; // Np x Np
    std::vector<double> M, invM;         // cubature mass matrix (unit Jacobian) and inverse
    std::vector<double> cr, cs, cw;      // Nq cubature points / weights
    std::vector<double> Vq, Drq, Dsq;    // Nq x Np
    std::vector<double> gr, gw;          // Nfq face Gauss points / weights
    std::vector<double> Vf;              // Nfq x Nfp  (face nodes -> face Gauss points)
    std::vector<int> fmask;              // 3 x Nfp   (face f runs from vertex f to f+1)
    // folded operators used by the kernels
    std::vector<double> Pr, Ps;          // Np x Nq : M^-1 Drq^T diag(cw), M^-1 Dsq^T diag(cw)
    std::vector<double> Lift;            // Np x (3*Nfq): M^-1 E_f Vf^T diag(gw)
    std::vector<double> LiftN;           // Np x (3*Nfp): M^-1 E (face mass on the face nodes), strong form
    // TVB limiter operators
    std::vector<double> tvb_p4;          // 4 x Np: cell mean, then the P1 part at the edge midpoints
    std::vector<double> tvb_phi;         // Np x 3: P1 functions with phi_i(m_k) = delta_ik at the nodes
    std::vector<double> tvb_wbc;         // Nfp: edge average of the face-node interpolant
};

// Builds all operators of order N (1..8). Returns false for unsupported N.
bool build_refelem(int N, RefElem& re);

// building blocks, exposed for tests through the C ABI
double jacobi_p(double x, double alpha, double beta, int n);
void gauss_jacobi(double alpha, double beta, int n, std::vector<double>& x, std::vector<double>& w);
void gauss_lobatto(int n, std::vector<double>& x);
bool invert(int n, const std::vector<double>& A, std::vector<double>& Ainv);

}  // namespace dg
