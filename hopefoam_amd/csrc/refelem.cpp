// Reference-triangle operators (host).  See refelem.hpp.
//
// Basis and nodes are the reference's: orthonormal Jacobi polynomials
// (Legendre.C:174-205), the Dubiner basis sqrt(2) P_i(a) P_j^{(2i+1,0)}(b) (1-b)^i
// (Legendre.C:246-275) and its gradient (Legendre.C:371-438), warp & blend nodes
// with the alpha-optimised blend (triangleBaseFunction.C:122-186).  Gauss points
// are found here by Newton iteration with deflation on P_{n+1}^{(alpha,beta)}
// (the reference solves the Golub-Welsch eigenproblem; both give the same rule).
//
// Everything is computed in long double and rounded once to double at the end, so the
// operators are the correctly rounded values of the exact ones to within an ulp.  This
// matters: the over-integrated Euler RHS cancels free-stream fluxes of size O(1) against
// each other and then applies M^-1 (entries ~400 at N = 6), so last-bit differences in
// the operators show up at 1e-12 in the RHS.
#include "refelem.hpp"

#include <cmath>
#include <cstdlib>

#include "dg_common.hpp"

namespace dg {

typedef long double real;
typedef std::vector<real> rvec;

static const real kPi = 3.14159265358979323846264338327950288L;

#include "cubature_tri_table.inc"

static real gamma_int(int x) {
    real a = 1.0L;
    for (int i = 2; i < x; ++i) a *= i;
    return a;
}

static real jp(real x, real alpha, real beta, int n) {
    const real g0 = std::pow(2.0L, alpha + beta + 1) / (alpha + beta + 1) * gamma_int(int(alpha) + 1) *
                    gamma_int(int(beta) + 1) / gamma_int(int(alpha + beta) + 1);
    real p_prev = 1.0L / std::sqrt(g0);
    if (n == 0) return p_prev;
    const real g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0;
    real p = ((alpha + beta + 2) / 2 * x + (alpha - beta) / 2) / std::sqrt(g1);
    real a_old = 2 / (2 + alpha + beta) * std::sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3));
    for (int i = 1; i < n; ++i) {
        const real h = 2.0L * i + alpha + beta;
        const real a_new = 2 / (h + 2) *
            std::sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta) / (h + 1) / (h + 3));
        const real b_new = -(alpha * alpha - beta * beta) / h / (h + 2);
        const real p_next = (-a_old * p_prev + (x - b_new) * p) / a_new;
        p_prev = p;
        p = p_next;
        a_old = a_new;
    }
    return p;
}

static real gjp(real x, real alpha, real beta, int n) {
    return n == 0 ? 0.0L : std::sqrt(real(n) * (n + alpha + beta + 1)) * jp(x, alpha + 1, beta + 1, n - 1);
}

double jacobi_p(double x, double alpha, double beta, int n) { return double(jp(x, alpha, beta, n)); }

// Roots of P_{n+1}^{(alpha,beta)} by Newton with deflation; for alpha = beta = 0 also the
// Gauss-Legendre weights w = 2 / ((1 - x^2) P'_{m}(x)^2) with P the classical Legendre polynomial.
static void gauss_jacobi_ld(real alpha, real beta, int n, rvec& x, rvec& w) {
    const int m = n + 1;
    x.assign(m, 0.0L);
    w.assign(m, 0.0L);
    if (m == 1) {
        x[0] = -(alpha - beta) / (alpha + beta + 2);
        w[0] = 2.0L;
        return;
    }
    for (int k = 0; k < m; ++k) {
        real r = -std::cos((2.0L * k + 1.0L) * kPi / (2.0L * m));
        if (k > 0) r = 0.5L * (r + x[k - 1]);
        for (int it = 0; it < 200; ++it) {
            real s = 0.0L;
            for (int j = 0; j < k; ++j) s += 1.0L / (r - x[j]);
            const real f = jp(r, alpha, beta, m);
            const real df = gjp(r, alpha, beta, m);
            const real delta = -f / (df - s * f);
            r += delta;
            if (std::fabs(delta) < 1e-21L) break;
        }
        x[k] = r;
    }
    if (alpha == beta) {   // symmetric rule: enforce x_k = -x_{m-1-k} exactly
        for (int k = 0; k < m / 2; ++k) {
            const real xs = 0.5L * (x[m - 1 - k] - x[k]);
            x[k] = -xs;
            x[m - 1 - k] = xs;
        }
        if (m % 2) x[m / 2] = 0.0L;
    }
    if (alpha == 0 && beta == 0) {
        const real norm = std::sqrt((2.0L * m + 1.0L) / 2.0L);   // orthonormal = norm * classical
        for (int k = 0; k < m; ++k) {
            const real dp = gjp(x[k], 0, 0, m) / norm;
            w[k] = 2.0L / ((1.0L - x[k] * x[k]) * dp * dp);
        }
        for (int k = 0; k < m / 2; ++k) w[k] = w[m - 1 - k] = 0.5L * (w[k] + w[m - 1 - k]);
    }
}

static void gauss_lobatto_ld(int n, rvec& x) {
    x.assign(n + 1, 0.0L);
    x[0] = -1.0L;
    x[n] = 1.0L;
    if (n < 2) return;
    rvec xi, wi;
    gauss_jacobi_ld(1.0L, 1.0L, n - 2, xi, wi);
    for (int i = 1; i < n; ++i) x[i] = xi[i - 1];
}

static bool invert_ld(int n, const rvec& A, rvec& Ainv) {
    // LU with partial pivoting, then solve for the identity's columns
    rvec lu(A);
    std::vector<int> piv(n);
    for (int i = 0; i < n; ++i) piv[i] = i;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(lu[r * n + c]) > std::fabs(lu[p * n + c])) p = r;
        if (lu[p * n + c] == 0.0L) return false;
        if (p != c) {
            for (int k = 0; k < n; ++k) std::swap(lu[c * n + k], lu[p * n + k]);
            std::swap(piv[c], piv[p]);
        }
        for (int r = c + 1; r < n; ++r) {
            const real f = lu[r * n + c] / lu[c * n + c];
            lu[r * n + c] = f;
            for (int k = c + 1; k < n; ++k) lu[r * n + k] -= f * lu[c * n + k];
        }
    }
    Ainv.assign(size_t(n) * n, 0.0L);
    rvec col(n);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) col[i] = (piv[i] == j) ? 1.0L : 0.0L;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) col[i] -= lu[i * n + k] * col[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k) col[i] -= lu[i * n + k] * col[k];
            col[i] /= lu[i * n + i];
        }
        for (int i = 0; i < n; ++i) Ainv[i * n + j] = col[i];
    }
    return true;
}

static std::vector<double> rounded(const rvec& v) { return std::vector<double>(v.begin(), v.end()); }

void gauss_jacobi(double alpha, double beta, int n, std::vector<double>& x, std::vector<double>& w) {
    rvec xl, wl;
    gauss_jacobi_ld(alpha, beta, n, xl, wl);
    x = rounded(xl);
    w = rounded(wl);
}

void gauss_lobatto(int n, std::vector<double>& x) {
    rvec xl;
    gauss_lobatto_ld(n, xl);
    x = rounded(xl);
}

bool invert(int n, const std::vector<double>& A, std::vector<double>& Ainv) {
    rvec Al(A.begin(), A.end()), Il;
    if (!invert_ld(n, Al, Il)) return false;
    Ainv = rounded(Il);
    return true;
}

static void warp_factor(int N, const rvec& rout, rvec& warp) {
    rvec lgl;
    gauss_lobatto_ld(N, lgl);
    rvec req(N + 1), VeqT(size_t(N + 1) * (N + 1)), inv;
    for (int i = 0; i <= N; ++i) req[i] = 2.0L / N * i - 1;
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) VeqT[j * (N + 1) + i] = jp(req[i], 0, 0, j);
    invert_ld(N + 1, VeqT, inv);
    warp.assign(rout.size(), 0.0L);
    for (size_t k = 0; k < rout.size(); ++k) {
        real w = 0.0L;
        for (int i = 0; i <= N; ++i) {
            real l = 0.0L;   // Lagrange polynomial i on the equidistant nodes, at rout[k]
            for (int j = 0; j <= N; ++j) l += inv[i * (N + 1) + j] * jp(rout[k], 0, 0, j);
            w += l * (lgl[i] - req[i]);
        }
        const real zf = std::fabs(rout[k]) < 1.0L - 1.0e-10L ? 1.0L : 0.0L;
        const real sf = 1.0L - (zf * rout[k]) * (zf * rout[k]);
        warp[k] = w / sf + w * (zf - 1.0L);
    }
}

static void nodes2d(int N, rvec& r, rvec& s) {
    // the reference's optimised blend parameters are the 4-digit decimals below
    static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                      1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
    const int Np = (N + 1) * (N + 2) / 2;
    const real alpha = N < 16 ? real(alpopt[N - 1]) : 5.0L / 3.0L;
    rvec L1(Np), L2(Np), L3(Np), x(Np), y(Np), d1(Np), d2(Np), d3(Np);
    int sk = 0;
    for (int n = 1; n <= N + 1; ++n)
        for (int m = 1; m <= N + 2 - n; ++m, ++sk) {
            L1[sk] = real(n - 1) / N;
            L3[sk] = real(m - 1) / N;
        }
    const real sq3 = std::sqrt(3.0L);
    for (int i = 0; i < Np; ++i) {
        L2[i] = 1.0L - L1[i] - L3[i];
        x[i] = L3[i] - L2[i];
        y[i] = (2 * L1[i] - L2[i] - L3[i]) / sq3;
        d1[i] = L3[i] - L2[i];
        d2[i] = L1[i] - L3[i];
        d3[i] = L2[i] - L1[i];
    }
    rvec w1, w2, w3;
    warp_factor(N, d1, w1);
    warp_factor(N, d2, w2);
    warp_factor(N, d3, w3);
    const real c1 = std::cos(2 * kPi / 3), c2 = std::cos(4 * kPi / 3);
    const real s1 = std::sin(2 * kPi / 3), s2 = std::sin(4 * kPi / 3);
    r.resize(Np);
    s.resize(Np);
    for (int i = 0; i < Np; ++i) {
        const real a = 4 * L2[i] * L3[i] * w1[i] * (1 + (alpha * L1[i]) * (alpha * L1[i]));
        const real b = 4 * L1[i] * L3[i] * w2[i] * (1 + (alpha * L2[i]) * (alpha * L2[i]));
        const real c = 4 * L1[i] * L2[i] * w3[i] * (1 + (alpha * L3[i]) * (alpha * L3[i]));
        const real xx = x[i] + a + c1 * b + c2 * c;
        const real yy = y[i] + s1 * b + s2 * c;
        const real l1 = (sq3 * yy + 1.0L) / 3.0L;
        const real l2 = (-3.0L * xx - sq3 * yy + 2.0L) / 6.0L;
        const real l3 = (3.0L * xx - sq3 * yy + 2.0L) / 6.0L;
        r[i] = l3 - l2 - l1;
        s[i] = l1 - l2 - l3;
    }
}

static void dubiner(int N, const rvec& r, const rvec& s, rvec* V, rvec* Vr, rvec* Vs) {
    const int Np = (N + 1) * (N + 2) / 2;
    const int m = int(r.size());
    if (V) V->assign(size_t(m) * Np, 0.0L);
    if (Vr) Vr->assign(size_t(m) * Np, 0.0L);
    if (Vs) Vs->assign(size_t(m) * Np, 0.0L);
    for (int n = 0; n < m; ++n) {
        const real a = s[n] != 1.0L ? 2 * (1 + r[n]) / (1 - s[n]) - 1 : -1.0L;
        const real b = s[n];
        int sk = 0;
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N - i; ++j, ++sk) {
                const real fa = jp(a, 0, 0, i), gb = jp(b, 2 * i + 1, 0, j);
                if (V) (*V)[n * Np + sk] = std::sqrt(2.0L) * fa * gb * std::pow(1 - b, i);
                if (!Vr) continue;
                const real dfa = gjp(a, 0, 0, i), dgb = gjp(b, 2 * i + 1, 0, j);
                const real c = std::pow(2.0L, i + 0.5L);
                const real hb = 0.5L * (1.0L - b);
                real vr, vs;
                if (i > 0) {
                    vr = dfa * gb * c * std::pow(hb, i - 1);
                    const real t = dgb * std::pow(hb, i) - 0.5L * i * gb * std::pow(hb, i - 1);
                    vs = (dfa * gb * 0.5L * (1 + a) * std::pow(hb, i - 1) + fa * t) * c;
                } else {
                    vr = dfa * gb * c;
                    vs = (dfa * gb * 0.5L * (1 + a) + fa * dgb) * c;
                }
                (*Vr)[n * Np + sk] = vr;
                (*Vs)[n * Np + sk] = vs;
            }
    }
}

static void matmul(int m, int k, int n, const real* A, const real* B, real* C) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            real acc = 0.0L;
            for (int l = 0; l < k; ++l) acc += A[i * k + l] * B[l * n + j];
            C[i * n + j] = acc;
        }
}

bool build_refelem(int N, RefElem& re) {
    if (N < 1 || N > 8) return false;
    const int order = 3 * (N + 1);
    if (order > dg_cub_max_order) return false;
    re.N = N;
    re.Np = (N + 1) * (N + 2) / 2;
    re.Nfp = N + 1;
    re.Nfq = N + 2;
    re.Nq = dg_cub_npts[order];
    const int Np = re.Np, Nq = re.Nq, Nfp = re.Nfp, Nfq = re.Nfq;
    rvec r, s, V, Vr, Vs, invV;
    nodes2d(N, r, s);
    dubiner(N, r, s, &V, &Vr, &Vs);
    if (!invert_ld(Np, V, invV)) return false;
    rvec Dr(size_t(Np) * Np), Ds(size_t(Np) * Np);
    matmul(Np, Np, Np, Vr.data(), invV.data(), Dr.data());
    matmul(Np, Np, Np, Vs.data(), invV.data(), Ds.data());
    // cubature of order 3(N+1): the table values are the rule (15 significant digits)
    rvec cr(dg_cub_r[order], dg_cub_r[order] + Nq), cs(dg_cub_s[order], dg_cub_s[order] + Nq);
    rvec cw(dg_cub_w[order], dg_cub_w[order] + Nq);
    rvec Vc, Vcr, Vcs;
    dubiner(N, cr, cs, &Vc, &Vcr, &Vcs);
    rvec Vq(size_t(Nq) * Np), Drq(size_t(Nq) * Np), Dsq(size_t(Nq) * Np);
    matmul(Nq, Np, Np, Vc.data(), invV.data(), Vq.data());
    matmul(Nq, Np, Np, Vcr.data(), invV.data(), Drq.data());
    matmul(Nq, Np, Np, Vcs.data(), invV.data(), Dsq.data());
    // face rule and face-node -> face-Gauss interpolation
    rvec gr, gw, gl;
    gauss_jacobi_ld(0, 0, N + 1, gr, gw);
    gauss_lobatto_ld(N, gl);
    rvec V1(size_t(Nfp) * Nfp), iV1, Vg(size_t(Nfq) * Nfp), Vf(size_t(Nfq) * Nfp);
    for (int i = 0; i < Nfp; ++i)
        for (int j = 0; j < Nfp; ++j) V1[i * Nfp + j] = jp(gl[i], 0, 0, j);
    for (int i = 0; i < Nfq; ++i)
        for (int j = 0; j < Nfp; ++j) Vg[i * Nfp + j] = jp(gr[i], 0, 0, j);
    invert_ld(Nfp, V1, iV1);
    matmul(Nfq, Nfp, Nfp, Vg.data(), iV1.data(), Vf.data());
    // face maps (triangleBaseFunction.C:75-94)
    re.fmask.assign(3 * Nfp, 0);
    for (int i = 0; i < Nfp; ++i) re.fmask[i] = i;
    re.fmask[Nfp] = N;
    re.fmask[2 * Nfp] = Np - 1;
    for (int i = 1; i < Nfp; ++i) {
        re.fmask[Nfp + i] = re.fmask[Nfp + i - 1] + (Nfp - i);
        re.fmask[2 * Nfp + i] = re.fmask[2 * Nfp + i - 1] - i - 1;
    }
    // the stage kernels take face nodes from dg::face_node at compile time: it must be this table
    for (int f = 0; f < 3; ++f)
        for (int k = 0; k < Nfp; ++k)
            if (re.fmask[f * Nfp + k] != dg::face_node_n(N, f, k)) return false;
    // cubature mass matrix (unit Jacobian) and inverse
    rvec M(size_t(Np) * Np), invM;
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            real acc = 0.0L;
            for (int q = 0; q < Nq; ++q) acc += Vq[q * Np + i] * cw[q] * Vq[q * Np + j];
            M[i * Np + j] = acc;
        }
    if (!invert_ld(Np, M, invM)) return false;
    // folded operators
    rvec DrT(size_t(Np) * Nq), DsT(size_t(Np) * Nq);
    for (int q = 0; q < Nq; ++q)
        for (int j = 0; j < Np; ++j) {
            DrT[j * Nq + q] = Drq[q * Np + j] * cw[q];
            DsT[j * Nq + q] = Dsq[q * Np + j] * cw[q];
        }
    rvec Pr(size_t(Np) * Nq), Ps(size_t(Np) * Nq);
    matmul(Np, Np, Nq, invM.data(), DrT.data(), Pr.data());
    matmul(Np, Np, Nq, invM.data(), DsT.data(), Ps.data());
    rvec E(size_t(Np) * 3 * Nfq, 0.0L), Lift(size_t(Np) * 3 * Nfq);
    for (int f = 0; f < 3; ++f)
        for (int g = 0; g < Nfq; ++g)
            for (int k = 0; k < Nfp; ++k)
                E[re.fmask[f * Nfp + k] * (3 * Nfq) + f * Nfq + g] += Vf[g * Nfp + k] * gw[g];
    matmul(Np, Np, 3 * Nfq, invM.data(), E.data(), Lift.data());
    // nodal lift of the strong form: M^-1 times the face mass matrix on each face's nodes
    // (Vf^T diag(gw) Vf, exact for the degree-2N products)
    rvec En(size_t(Np) * 3 * Nfp, 0.0L), LiftN(size_t(Np) * 3 * Nfp);
    for (int f = 0; f < 3; ++f)
        for (int a = 0; a < Nfp; ++a)
            for (int b = 0; b < Nfp; ++b) {
                real m1 = 0.0L;
                for (int g = 0; g < Nfq; ++g) m1 += Vf[g * Nfp + a] * gw[g] * Vf[g * Nfp + b];
                En[re.fmask[f * Nfp + a] * (3 * Nfp) + f * Nfp + b] += m1;
            }
    matmul(Np, Np, 3 * Nfp, invM.data(), En.data(), LiftN.data());

    // TVB limiter: mean = (V^-1 u)_0 psi_0; the P1 part keeps the modes (0,0), (0,1), (1,0) of the
    // (i, j) ordering, i.e. indices 0, 1, N+1; edge midpoints of edges 0, 1, 2 in (r, s)
    {
        const rvec mr = {0.0L, 0.0L, -1.0L}, ms = {-1.0L, 0.0L, 0.0L};
        rvec Vm;
        dubiner(N, mr, ms, &Vm, nullptr, nullptr);
        const int p1[3] = {0, 1, N + 1};
        re.tvb_p4.assign(size_t(4) * Np, 0.0);
        for (int j = 0; j < Np; ++j) {
            re.tvb_p4[j] = double(Vm[0] * invV[j]);
            for (int i = 0; i < 3; ++i) {
                real a = 0.0L;
                for (int k = 0; k < 3; ++k) a += Vm[i * Np + p1[k]] * invV[size_t(p1[k]) * Np + j];
                re.tvb_p4[size_t(1 + i) * Np + j] = double(a);
            }
        }
        re.tvb_phi.assign(size_t(Np) * 3, 0.0);
        for (int n = 0; n < Np; ++n) {
            const real l0 = -(r[n] + s[n]) * 0.5L, l1 = (r[n] + 1.0L) * 0.5L, l2 = (s[n] + 1.0L) * 0.5L;
            re.tvb_phi[n * 3 + 0] = double(1.0L - 2.0L * l2);
            re.tvb_phi[n * 3 + 1] = double(1.0L - 2.0L * l0);
            re.tvb_phi[n * 3 + 2] = double(1.0L - 2.0L * l1);
        }
        re.tvb_wbc.assign(Nfp, 0.0);
        for (int k = 0; k < Nfp; ++k) {
            real a = 0.0L;
            for (int g = 0; g < Nfq; ++g) a += 0.5L * gw[g] * Vf[g * Nfp + k];
            re.tvb_wbc[k] = double(a);
        }
    }

    re.r = rounded(r); re.s = rounded(s); re.V = rounded(V); re.invV = rounded(invV);
    re.Dr = rounded(Dr); re.Ds = rounded(Ds); re.M = rounded(M); re.invM = rounded(invM);
    re.cr = rounded(cr); re.cs = rounded(cs); re.cw = rounded(cw);
    re.Vq = rounded(Vq); re.Drq = rounded(Drq); re.Dsq = rounded(Dsq);
    re.gr = rounded(gr); re.gw = rounded(gw); re.Vf = rounded(Vf);
    re.Pr = rounded(Pr); re.Ps = rounded(Ps); re.Lift = rounded(Lift); re.LiftN = rounded(LiftN);
    return true;
}

}  // namespace dg
