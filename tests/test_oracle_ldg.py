"""Pins of the LDG advection-diffusion oracle (oracle/ldg.c; CPU only).

LDG is not in the reference (dgc::laplacian is an empty stub there), so the oracle restates the published
method (Cockburn & Shu 1998, alternating fluxes; nodal strong form of Hesthaven & Warburton 7.1) and is
pinned by what the mathematics fixes:
  * nu = 0 gives the upwind advection operator, which must equal the reference's cubature weak form
    (or_rhs_advect, itself pinned by the reference's known answers) to round-off;
  * the lift operator integrates: 1^T M LIFT_f = the face weights (sum of each face's mass rows);
  * with a = 0 the operator is symmetric and negative semi-definite in the mass inner product (LDG is a
    symmetric, adjoint-consistent primal form), with constants in its kernel;
  * the discrete total mass is conserved on periodic meshes; advection + diffusion is dissipative;
  * manufactured solution: the advected, decaying Fourier mode converges at order >= N + 1/2 under mesh
    refinement (observed N + 1 on these meshes).
"""
import numpy as np
import pytest

from meshgen import split_quads
from oracle_lib import OracleMesh

BETA = (1.0, 1.0 / np.pi)


def mesh(N, n, L=1.0, jitter=0.0, seed=0):
    vxy, vid = split_quads(n, n, 0.0, L, 0.0, L, periodic=(True, True))
    if jitter:
        # move interior vertex ids consistently (periodic images share ids): perturb by id
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, (vid.max() + 1, 2)) * (L / n)
        vxy = vxy + d[vid]
    return OracleMesh(N, vxy, vid)


def mass_inner(om, a, b):
    from oracle_lib import refelem_ops
    Mr = refelem_ops(om.N)["M"]
    J = om.geo()[:, 4]
    return float(np.sum(J[:, None] * (a @ Mr) * b))


@pytest.mark.parametrize("N", [1, 2, 4, 6, 8])
def test_no_diffusion_is_the_reference_advection(N):
    om = mesh(N, 5, jitter=0.15)
    x, y = om.nodes()
    u = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.3 * np.cos(3 * x * y)
    for a in ((1.0, 0.5), (-0.7, 1.3)):
        r = om.ldg_rhs(u, a[0], a[1], 0.0, tau=2.0)
        want = om.rhs_advect(u, a[0], a[1])
        assert np.abs(r - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("N", [1, 3, 6])
def test_lift_integrates_face_mass(N):
    from oracle_lib import refelem_ops
    om = mesh(N, 3)
    lift = om.ldg_lift()
    M = refelem_ops(N)["M"]
    w = np.ones(om.Np) @ M @ lift            # sum over nodes of M LIFT = face weights per face node
    Nfp = om.Nfp
    from numpy.polynomial.legendre import leggauss
    xg, wg = leggauss(Nfp + 1)
    # the face-node quadrature weights of the exact face mass: integral of each face Lagrange function
    from oracle_lib import lib
    gl = np.zeros(Nfp)
    lib().or_jacobi_gl(0.0, 0.0, N, gl)
    V = np.polynomial.legendre.legvander(gl, N)
    c = np.linalg.solve(V, np.eye(Nfp))     # Lagrange coefficients (Legendre basis)
    Vg = np.polynomial.legendre.legvander(xg, N)
    wk = (wg @ Vg) @ c
    for f in range(3):
        np.testing.assert_allclose(w[f * Nfp:(f + 1) * Nfp], wk, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("N", [1, 2, 3, 5])
def test_diffusion_symmetric_negative_semidefinite(N):
    om = mesh(N, 4, jitter=0.2, seed=N)
    rng = np.random.default_rng(N)
    u = rng.standard_normal((om.K, om.Np))
    v = rng.standard_normal((om.K, om.Np))
    for tau in (0.0, 1.0):
        Lu = om.ldg_rhs(u, 0.0, 0.0, 1.0, tau=tau, beta=BETA)
        Lv = om.ldg_rhs(v, 0.0, 0.0, 1.0, tau=tau, beta=BETA)
        a, b = mass_inner(om, v, Lu), mass_inner(om, Lv, u)
        scale = np.sqrt(abs(mass_inner(om, u, Lu) * mass_inner(om, v, Lv)))
        assert abs(a - b) <= 1e-11 * scale
        assert mass_inner(om, u, Lu) < 0 and mass_inner(om, v, Lv) < 0
        np.testing.assert_allclose(om.ldg_rhs(np.full_like(u, 2.5), 0.0, 0.0, 1.0, tau=tau), 0.0,
                                   atol=1e-13 * np.abs(Lu).max())
    # a switch direction flipped is a different, equally valid scheme: still symmetric
    Lu = om.ldg_rhs(u, 0.0, 0.0, 1.0, tau=0.5, beta=(-0.3, 1.0))
    Lv = om.ldg_rhs(v, 0.0, 0.0, 1.0, tau=0.5, beta=(-0.3, 1.0))
    assert abs(mass_inner(om, v, Lu) - mass_inner(om, Lv, u)) <= 1e-11 * abs(mass_inner(om, u, Lu))


@pytest.mark.parametrize("N", [2, 4])
def test_conservation_and_dissipation(N):
    om = mesh(N, 5, jitter=0.15)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((om.K, om.Np))
    r = om.ldg_rhs(u, 0.8, -0.6, 0.02, tau=1.0)
    assert abs(mass_inner(om, np.ones_like(u), r)) <= 1e-11 * np.abs(r).max()
    assert mass_inner(om, u, r) < 0


def exact(x, y, t, a, nu):
    k2 = 2 * (2 * np.pi) ** 2
    return np.exp(-nu * k2 * t) * np.sin(2 * np.pi * (x - a[0] * t)) * np.cos(2 * np.pi * (y - a[1] * t))


def l2_error(om, u, ue):
    return np.sqrt(mass_inner(om, u - ue, u - ue))


@pytest.mark.parametrize("N", [1, 2, 3])
def test_manufactured_convergence(N):
    a, nu, T = (1.0, 0.5), 0.01, 0.05
    errs = []
    for n in (6, 12):
        om = mesh(N, n)
        x, y = om.nodes()
        h = 1.0 / n
        dt = 0.25 * min(h / (np.hypot(*a) * (N + 1) ** 2), h * h / (nu * (N + 1) ** 4))
        steps = int(np.ceil(T / dt))
        u = om.ldg_lserk4(exact(x, y, 0.0, a, nu), steps, T / steps, a[0], a[1], nu, tau=1.0)
        errs.append(l2_error(om, u, exact(x, y, T, a, nu)))
    rate = np.log2(errs[0] / errs[1])
    print("N=%d errors %s rate %.2f" % (N, errs, rate))
    assert rate >= N + 0.5
