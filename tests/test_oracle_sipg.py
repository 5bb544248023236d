"""Pins of the oracle's SIPG assembly and PETSc-order CG (CPU only).

What fixes them without running the reference:
  * consistency: SIPG with exact quadrature reproduces any solution that lies in the DG space, so a
    degree-N polynomial u with f = lap(u) and Dirichlet data g = u must come back exactly -- every volume,
    face, penalty and boundary term and its sign enters this;
  * symmetry and negative definiteness of the reference's (negative) form; constants in the kernel on a
    periodic mesh;
  * O(h^(N+1)) convergence on the manufactured solution sin(2 pi x) sin(3 pi y);
  * the CG iterates against scipy's independent CG on the same SPD system S = -A with the Jacobi
    preconditioner, and the stopping rule (preconditioned norm, relative to the initial one) recomputed
    from those iterates;
  * the device's one-reduction form of the recurrence (restated here) against the oracle's PETSc-order CG:
    same iteration counts, residual histories and solutions.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from meshgen import split_quads
from oracle_lib import OracleMesh, cg_jacobi, refelem_ops


def dirichlet_square(n, N, penalty, ufun, lapfun, tags=None):
    vxy, vid = split_quads(n, n, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    bct = np.ones(vid.shape, np.int32) if tags is None else tags(vxy)
    m = OracleMesh(N, vxy, vid, bct)
    x, y = m.nodes()
    fm = refelem_ops(N)["fmask"]
    g = np.stack([ufun(x[:, fm[f]], y[:, fm[f]]) for f in range(3)], axis=1)
    A, b = m.sipg_system(penalty, lapfun(x, y), g)
    return m, A, b, ufun(x, y)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("penalty", [0, 1])
def test_polynomial_solution_is_reproduced(N, penalty):
    c = np.array([0.3, -1.2, 0.7, 0.25, -0.4, 0.9, 0.15, -0.35, 0.6, -0.2])
    mons = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
    use = [(k, (a, b)) for k, (a, b) in enumerate(mons) if a + b <= N]

    def u(x, y):
        return sum(c[k] * x ** a * y ** b for k, (a, b) in use)

    def lap(x, y):
        out = np.zeros_like(x)
        for k, (a, b) in use:
            if a >= 2:
                out += c[k] * a * (a - 1) * x ** (a - 2) * y ** b
            if b >= 2:
                out += c[k] * b * (b - 1) * x ** a * y ** (b - 2)
        return out

    m, A, b, ue = dirichlet_square(3, N, penalty, u, lap)
    sol = spla.spsolve(A.tocsc(), b)
    assert np.abs(sol - ue.ravel()).max() < 1e-10 * max(1.0, np.abs(ue).max())


@pytest.mark.parametrize("penalty", [0, 1])
def test_symmetric_negative_definite(penalty):
    m, A, b, _ = dirichlet_square(4, 3, penalty, lambda x, y: 0 * x, lambda x, y: 0 * x)
    D = A.toarray()
    assert np.abs(D - D.T).max() <= 1e-13 * np.abs(D).max()
    ev = np.linalg.eigvalsh(0.5 * (D + D.T))
    assert ev.max() < 0


def test_constants_in_kernel_on_periodic_mesh():
    vxy, vid = split_quads(4, 4, 0.0, 1.0, 0.0, 1.0, periodic=(True, True))
    m = OracleMesh(3, vxy, vid)
    A, _ = m.sipg_system(0, np.zeros((m.K, m.Np)))
    one = np.ones(A.shape[0])
    assert np.abs(A @ one).max() <= 1e-12 * abs(A).max()
    # a non-constant field is not in the kernel
    x, y = m.nodes()
    assert np.abs(A @ np.sin(2 * np.pi * x).ravel()).max() > 1e-3


def test_neumann_faces_have_no_face_term():
    # tags 2 (zero gradient) on every boundary: pure Neumann problem, constants in the kernel again
    vxy, vid = split_quads(3, 3, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    m = OracleMesh(2, vxy, vid, np.full(vid.shape, 2, np.int32))
    A, _ = m.sipg_system(1, np.zeros((m.K, m.Np)))
    assert np.abs(A @ np.ones(A.shape[0])).max() <= 1e-12 * abs(A).max()


def test_manufactured_solution_rate():
    def u(x, y):
        return np.sin(2 * np.pi * x) * np.sin(3 * np.pi * y)

    def lap(x, y):
        return -13 * np.pi ** 2 * u(x, y)

    errs = []
    for n in (8, 16):
        m, A, b, ue = dirichlet_square(n, 2, 0, u, lap)
        errs.append(np.abs(spla.spsolve(A.tocsc(), b) - ue.ravel()).max())
    assert np.log2(errs[0] / errs[1]) > 2.6    # O(h^3) for N = 2


def small_system(N=2, n=4, penalty=0):
    def u(x, y):
        return np.sin(2 * np.pi * x) * np.sin(3 * np.pi * y)
    return dirichlet_square(n, N, penalty, u, lambda x, y: -13 * np.pi ** 2 * u(x, y))


@pytest.mark.parametrize("k", [1, 5, 17])
def test_cg_iterates_match_scipy(k):
    m, A, b, _ = small_system()
    S = (-A).tocsr()
    dinv = 1.0 / S.diagonal()
    x0 = np.cos(np.arange(A.shape[0]) * 0.37)
    xo, its, _, _ = cg_jacobi(A, b, x0, rtol=1e-300, abstol=0.0, maxit=k)
    assert its == k
    xs, _ = spla.cg(S, -b, x0=x0.copy(), rtol=1e-300, atol=0.0, maxiter=k,
                    M=spla.LinearOperator(S.shape, matvec=lambda v: dinv * v))
    assert np.linalg.norm(xo - xs) <= 1e-10 * np.linalg.norm(xs)


def test_cg_stopping_rule():
    m, A, b, _ = small_system()
    S = (-A).tocsr()
    dinv = 1.0 / S.diagonal()
    x0 = np.zeros(A.shape[0])
    rtol = 1e-6
    x, its, rn0, rn = cg_jacobi(A, b, x0, rtol=rtol)
    # preconditioned residual norms of scipy's iterates on the same system
    hist = []
    spla.cg(S, -b, x0=x0.copy(), rtol=1e-300, atol=0.0, maxiter=its + 3,
            M=spla.LinearOperator(S.shape, matvec=lambda v: dinv * v),
            callback=lambda xk: hist.append(np.linalg.norm(dinv * (-b - S @ xk))))
    z0 = np.linalg.norm(dinv * (-b - S @ x0))
    assert rn0 == pytest.approx(z0, rel=1e-12)
    first = next(i + 1 for i, v in enumerate(hist) if v <= rtol * z0)
    assert abs(first - its) <= 1
    assert rn <= rtol * rn0
    # the recorded history is the same quantity, and the rule stops at its first value <= ttol
    _, its_h, _, _, h = cg_jacobi(A, b, x0, rtol=rtol, history=True)
    assert its_h == its and h[0] == rn0 and h[-1] == rn
    assert np.all(h[:-1] > rtol * rn0) and h[-1] <= rtol * rn0
    np.testing.assert_allclose(h[1:11], hist[:10], rtol=1e-8)
    # nonzero initial guess (KSPSetInitialGuessNonzero): starting from the solution needs no iteration
    _, its2, _, _ = cg_jacobi(A, b, spla.spsolve(A.tocsc(), b), rtol=rtol, abstol=1e-9)
    assert its2 == 0


def single_reduction_pcg(S, dinv, rhs, x, rtol, maxit, drop_correction=False):
    """The recurrence of the device PCG (sipg.hip: cg_finish, cg_update_kernel), restated: PCG in the
    Chronopoulos & Gear form, one reduction (u.u, r.u, u.S u) per iteration, KSPCG's test on |u| = |B r|.
    drop_correction: p.S p taken as u.S u (a plausible slip), which the pin must catch."""
    r = rhs - S @ x
    u = dinv * r
    w = S @ u
    its, gamma, alpha, p, ap, hist = 0, r @ u, None, None, None, []
    dp = np.sqrt(u @ u)
    ttol = rtol * dp
    hist.append(dp)
    if dp <= ttol:
        return x, 0, hist
    beta = 0.0
    pw = u @ w
    while True:
        alpha = gamma / pw
        p = u.copy() if its == 0 else u + beta * p
        ap = w.copy() if its == 0 else w + beta * ap
        x = x + alpha * p
        r = r - alpha * ap
        u = dinv * r
        w = S @ u
        gnew, dp = r @ u, np.sqrt(u @ u)
        its += 1
        hist.append(dp)
        if dp <= ttol or its >= maxit:
            return x, its, hist
        beta = gnew / gamma
        pw = (u @ w) if drop_correction else (u @ w) - beta * gnew / alpha
        gamma = gnew


@pytest.mark.parametrize("N,n,penalty", [(3, 4, 1), (5, 8, 1)])
def test_single_reduction_recurrence_matches_petsc_order_cg(N, n, penalty):
    """The device's one-reduction PCG takes the PETSc-order oracle CG's iterations and iterates: the same
    iteration counts at every tolerance, preconditioned residual histories within 1e-8 of the initial norm
    and solutions to 1e-10; with the p.S p correction dropped it takes about twice the iterations."""
    def u(x, y):
        return np.sin(2 * np.pi * x) * np.sin(3 * np.pi * y)
    m, A, b, _ = dirichlet_square(n, N, penalty, u, lambda x, y: -13 * np.pi ** 2 * u(x, y))
    S = (-A).tocsr()
    dinv = 1.0 / S.diagonal()
    x0 = np.zeros(A.shape[0])
    for rtol in (1e-4, 1e-8, 1e-12):
        xo, its, rn0, _, ho = cg_jacobi(A, b, x0, rtol=rtol, history=True)
        xs, its_s, hs = single_reduction_pcg(S, dinv, -b, x0.copy(), rtol, 100000)
        assert its_s == its
        np.testing.assert_allclose(hs, ho, rtol=0, atol=1e-8 * rn0)
        assert np.linalg.norm(xs - xo) <= 1e-10 * np.linalg.norm(xo)
    _, its_bad, _ = single_reduction_pcg(S, dinv, -b, x0.copy(), 1e-8, 100000, drop_correction=True)
    _, its_ref, _, _ = cg_jacobi(A, b, x0, rtol=1e-8)
    assert its_bad > 1.5 * its_ref


@pytest.mark.parametrize("sigma", [0.5, 40.0])
def test_mass_shift_row_sums(sigma):
    """Helmholtz form A - sigma blockdiag(J_e M): with zero-gradient sides the Laplacian of a constant
    vanishes, so each element's rows of A_sigma 1 sum to -sigma times its area, and all of them to -sigma
    times the area of the unit square."""
    vxy, vid = split_quads(3, 4, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    m = OracleMesh(3, vxy, vid, np.full(vid.shape, 2, np.int32))
    A, _ = m.sipg_system(1, np.zeros((m.K, m.Np)), mass_shift=sigma)
    r = (A @ np.ones(A.shape[0])).reshape(m.K, m.Np)
    v = vxy
    area = 0.5 * np.abs((v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) -
                        (v[:, 2, 0] - v[:, 0, 0]) * (v[:, 1, 1] - v[:, 0, 1]))
    np.testing.assert_allclose(r.sum(axis=1), -sigma * area, rtol=1e-12, atol=1e-12 * sigma)
    assert r.sum() == pytest.approx(-sigma, rel=1e-12)


@pytest.mark.parametrize("N", [1, 2, 3])
def test_helmholtz_polynomial_solution_is_reproduced(N):
    """-lap u + sigma u = g with Dirichlet data: a degree-N polynomial lies in the DG space and comes back
    exactly from the shifted system (f = lap u - sigma u in the reference's sign)."""
    sigma = 7.5
    c = np.array([0.3, -1.2, 0.7, 0.25, -0.4, 0.9, 0.15, -0.35, 0.6, -0.2])
    mons = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
    use = [(k, (a, b)) for k, (a, b) in enumerate(mons) if a + b <= N]

    def u(x, y):
        return sum(c[k] * x ** a * y ** b for k, (a, b) in use)

    def lap(x, y):
        out = np.zeros_like(x)
        for k, (a, b) in use:
            if a >= 2:
                out += c[k] * a * (a - 1) * x ** (a - 2) * y ** b
            if b >= 2:
                out += c[k] * b * (b - 1) * x ** a * y ** (b - 2)
        return out

    vxy, vid = split_quads(3, 3, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    m = OracleMesh(N, vxy, vid, np.ones(vid.shape, np.int32))
    x, y = m.nodes()
    fm = refelem_ops(N)["fmask"]
    g = np.stack([u(x[:, fm[f]], y[:, fm[f]]) for f in range(3)], axis=1)
    A, b = m.sipg_system(0, lap(x, y) - sigma * u(x, y), g, mass_shift=sigma)
    sol = spla.spsolve(A.tocsc(), b)
    ue = u(x, y).ravel()
    assert np.abs(sol - ue).max() < 1e-10 * max(1.0, np.abs(ue).max())
    # the unshifted system with the same data does not reproduce it
    A0, _ = m.sipg_system(0, lap(x, y) - sigma * u(x, y), g)
    assert np.abs(spla.spsolve(A0.tocsc(), b) - ue).max() > 1e-3
