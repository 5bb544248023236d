"""Pins of the CPU oracle (oracle/) against the reference's published known answers,
library routines and closed forms.  These run without a GPU."""
import numpy as np
import pytest
from scipy.special import roots_jacobi, roots_legendre

from meshgen import load_fixture_mesh, split_quads
from oracle_lib import OracleMesh, lib, refelem_ops


# ---------------------------------------------------------------- known answers
def test_vortex_user_guide_known_answer():
    """UG 1.7-1.8 (HopeFOAM-0.1_User_Guide.pdf pp.15-16): vortex1024.msh, N=4, t=2,
    rhoError 8.807979526797244e-06, rhoUError 1.865574862711117e-05 (eulerError.H:37-38).
    Reproduced with dt = 0.004 (UG Table 1.1)."""
    vxy, tri = load_fixture_mesh("vortex1024")
    m = OracleMesh(4, vxy, tri, np.ones_like(tri))
    er, eu = m.vortex_ref_run(0.004, 500)
    assert abs(er / 8.807979526797244e-06 - 1) < 1e-9
    assert abs(eu / 1.865574862711117e-05 - 1) < 1e-9


# SLIDES slide 18 (4 significant digits), dt from UG Table 1.1
SLIDE18 = [
    ("vortex0256", 1, 0.04, 1.101e-02, 2.361e-02),
    ("vortex0256", 2, 0.02, 3.481e-03, 6.384e-03),
    ("vortex0256", 3, 0.008, 6.523e-04, 1.610e-03),
    ("vortex1024", 1, 0.02, 3.344e-03, 7.182e-03),
    ("vortex1024", 2, 0.01, 3.621e-04, 8.060e-04),
    ("vortex0256", 4, 0.008, 2.352e-04, 5.187e-04),
]


@pytest.mark.parametrize("name,N,dt,erho,erhou", SLIDE18)
def test_vortex_slide18_table(name, N, dt, erho, erhou):
    vxy, tri = load_fixture_mesh(name)
    m = OracleMesh(N, vxy, tri, np.ones_like(tri))
    er, eu = m.vortex_ref_run(dt, int(round(2.0 / dt)))
    assert abs(er / erho - 1) < 6e-4
    assert abs(eu / erhou - 1) < 6e-4


# ---------------------------------------------------------------- reference element
@pytest.mark.parametrize("n", [0, 1, 3, 6, 9])
def test_gauss_legendre_matches_library(n):
    x = np.zeros(n + 1)
    w = np.zeros(n + 1)
    lib().or_jacobi_gq(0.0, 0.0, n, x, w)
    xr, wr = roots_legendre(n + 1)
    assert np.allclose(x, xr, atol=1e-14, rtol=0)
    assert np.allclose(w, wr, atol=1e-14, rtol=0)


@pytest.mark.parametrize("n", [2, 4, 8])
def test_gauss_lobatto_matches_library(n):
    x = np.zeros(n + 1)
    lib().or_jacobi_gl(0.0, 0.0, n, x)
    xi, _ = roots_jacobi(n - 1, 1, 1)
    assert np.allclose(x, np.concatenate([[-1], xi, [1]]), atol=1e-14, rtol=0)


def test_jacobi_orthonormal():
    x, w = roots_legendre(30)
    for a, b in [(0, 0), (1, 0), (3, 0), (1, 1)]:
        wt = (1 - x) ** a * (1 + x) ** b * w
        P = np.array([[lib().or_jacobi_p(xi, a, b, n) for xi in x] for n in range(6)])
        G = (P * wt) @ P.T
        assert np.allclose(G, np.eye(6), atol=1e-13)
    # derivative against a centred difference of the polynomial
    for n in range(1, 6):
        h = 1e-6
        for xi in (-0.7, 0.1, 0.55):
            fd = (lib().or_jacobi_p(xi + h, 2, 0, n) - lib().or_jacobi_p(xi - h, 2, 0, n)) / (2 * h)
            assert abs(lib().or_grad_jacobi_p(xi, 2, 0, n) - fd) < 1e-7


@pytest.mark.parametrize("order", [6, 9, 12, 15, 18, 21, 24, 27])
def test_cubature_exactness(order):
    n = 400
    r = np.zeros(n); s = np.zeros(n); w = np.zeros(n)
    npts = lib().or_cubature(order, r, s, w)
    r, s, w = r[:npts], s[:npts], w[:npts]
    # exact integral over the reference triangle of ((1+r)/2)^a ((1+s)/2)^b = 4 a! b! / (a+b+2)!
    from math import factorial
    for a in range(order + 1):
        for b in range(order + 1 - a):
            exact = 4.0 * factorial(a) * factorial(b) / factorial(a + b + 2)
            got = np.sum(w * ((1 + r) / 2) ** a * ((1 + s) / 2) ** b)
            assert abs(got - exact) < 1e-13 * exact + 1e-15, (order, a, b, got, exact)


@pytest.mark.parametrize("N", [1, 2, 4, 6, 8])
def test_refelem_invariants(N):
    o = refelem_ops(N)
    r, s = o["r"], o["s"]
    Np = (N + 1) * (N + 2) // 2
    assert r.size == Np
    one = np.ones(Np)
    assert np.abs(o["Dr"] @ one).max() < 1e-12
    assert np.abs(o["Dr"] @ r - 1).max() < 1e-12 and np.abs(o["Ds"] @ s - 1).max() < 1e-12
    assert np.abs(o["Dr"] @ s).max() < 1e-12 and np.abs(o["Ds"] @ r).max() < 1e-12
    # derivative of a degree-N monomial is exact
    p = r ** (N - 1) * s if N > 1 else r
    dp = (N - 1) * r ** (N - 2) * s if N > 1 else one
    assert np.abs(o["Dr"] @ p - dp).max() < 1e-10
    # exact mass matrix V^-T V^-1 equals the cubature mass matrix
    Mex = o["invV"].T @ o["invV"]
    assert np.abs(o["M"] - Mex).max() < 1e-13
    # vertices are nodes; face 0 nodes are the GL nodes on s = -1
    fm = o["fmask"]
    gl = np.zeros(N + 1)
    lib().or_jacobi_gl(0.0, 0.0, N, gl)
    assert np.allclose(r[fm[0]], gl, atol=1e-13) and np.allclose(s[fm[0]], -1, atol=1e-13)
    assert np.allclose(s[fm[1]], -r[fm[1]], atol=1e-13)       # hypotenuse r + s = 0
    assert np.allclose(r[fm[2]], -1, atol=1e-13)
    assert np.allclose(s[fm[2]], gl[::-1], atol=1e-13)         # face 2 runs v2 -> v0
    # face Gauss interpolation reproduces degree-N polynomials along the edge
    gx = np.zeros(N + 2); gw = np.zeros(N + 2)
    lib().or_jacobi_gq(0.0, 0.0, N + 1, gx, gw)
    assert np.allclose(o["Vf"] @ gl ** N, gx ** N, atol=1e-13)
    assert np.allclose(o["gw"], gw)
    # cubature interpolation reproduces the polynomial at the cubature points
    npts = o["Vq"].shape[0]
    cr = np.zeros(400); cs = np.zeros(400); cw = np.zeros(400)
    lib().or_cubature(3 * (N + 1), cr, cs, cw)
    cr, cs = cr[:npts], cs[:npts]
    f = lambda a, b: a ** N + 2 * a * b ** (N - 1) - b
    assert np.allclose(o["Vq"] @ f(r, s), f(cr, cs), atol=1e-11)


# ---------------------------------------------------------------- mesh and RHS invariants
def test_periodic_connectivity():
    vxy, vid = split_quads(6, 5, 0.0, 1.0, 0.0, 1.0)
    m = OracleMesh(3, vxy, vid)
    E, F = m.conn()
    assert (E >= 0).all() and m.nbfaces == 0
    for e in range(m.K):
        for f in range(3):
            assert E[E[e, f], F[e, f]] == e and F[E[e, f], F[e, f]] == f
    g = m.geo()
    assert np.allclose(g[:, 4], (1 / 6) * (1 / 5) / 4)     # J = area / 2


@pytest.mark.parametrize("N", [1, 3, 5])
def test_advection_rhs_exact_for_polynomials(N):
    """Weak form with exact traces (fixedValue boundary = exact polynomial) gives
    rhs = -a.grad(q) exactly for polynomial q of degree <= N."""
    vxy, vid = split_quads(3, 4, -1.0, 1.0, -1.0, 0.5, periodic=(False, False))
    bct = np.ones((vid.shape[0], 3), np.int32)
    m = OracleMesh(N, vxy, vid, bct)
    x, y = m.nodes()
    fq = lambda x, y: x ** N - 3 * x * y ** (N - 1) + 0.5 * y
    gq = lambda x, y: (N * x ** (N - 1) - 3 * y ** (N - 1), -3 * (N - 1) * x * y ** max(N - 2, 0) + 0.5)
    q = fq(x, y)
    fm = refelem_ops(N)["fmask"]
    bcq = np.zeros((m.K, 3, N + 1))
    for f in range(3):
        bcq[:, f, :] = q[:, fm[f]]
    ax, ay = 1.0, 0.5
    rhs = m.rhs_advect(q, ax, ay, bcq)
    gx, gy = gq(x, y)
    assert np.abs(rhs + ax * gx + ay * gy).max() < 1e-10 * max(1, np.abs(gx).max())


@pytest.mark.parametrize("N", [1, 3, 5])
def test_average_flux_derivatives_exact_for_polynomials(N):
    """The reference's "average" (central) flux with a = (1, 0) and (0, 1) and exact boundary traces: the
    right-hand sides are minus the x and y derivatives of any polynomial of degree <= N, exactly."""
    vxy, vid = split_quads(3, 4, -1.0, 1.0, -1.0, 0.5, periodic=(False, False))
    bct = np.ones((vid.shape[0], 3), np.int32)
    m = OracleMesh(N, vxy, vid, bct)
    x, y = m.nodes()
    q = x ** N - 3 * x * y ** (N - 1) + 0.5 * y
    gx = N * x ** (N - 1) - 3 * y ** (N - 1)
    gy = -3 * (N - 1) * x * y ** max(N - 2, 0) + 0.5
    fm = refelem_ops(N)["fmask"]
    bcq = np.ascontiguousarray(np.stack([q[:, fm[f]] for f in range(3)], axis=1))
    rx = m.rhs_advect(q, 1.0, 0.0, bcq, average=True)
    ry = m.rhs_advect(q, 0.0, 1.0, bcq, average=True)
    scale = max(1.0, np.abs(gx).max(), np.abs(gy).max())
    assert np.abs(rx + gx).max() < 1e-10 * scale and np.abs(ry + gy).max() < 1e-10 * scale


@pytest.mark.parametrize("N", [2, 4])
def test_average_flux_is_skew_adjoint(N):
    """On a periodic mesh the central-flux DG divergence of a u is skew-adjoint in the mass inner product:
    u^T M rhs(u) = 0 for every u (no dissipation), while the upwind flux dissipates (u^T M rhs(u) < 0)."""
    vxy, vid = split_quads(5, 4, 0.0, 1.0, 0.0, 1.0, periodic=(True, True))
    m = OracleMesh(N, vxy, vid)
    u = np.random.default_rng(N).standard_normal((m.K, m.Np))
    J = m.geo()[:, 4]
    Mref = refelem_ops(N)["M"]
    def energy_rate(r):
        return float(np.einsum("ei,ij,ej->", u, Mref, r * J[:, None]))
    ra = m.rhs_advect(u, 0.8, -0.35, average=True)
    ru = m.rhs_advect(u, 0.8, -0.35)
    scale = float(np.sqrt(np.einsum("ei,ij,ej->", u, Mref, u * J[:, None]) *
                          np.einsum("ei,ij,ej->", ra, Mref, ra * J[:, None])))
    assert abs(energy_rate(ra)) < 1e-12 * scale
    assert energy_rate(ru) < -1e-3 * scale


def test_euler_rhs_free_stream_and_conservation():
    vxy, vid = split_quads(5, 4, 0.0, 2.0, 0.0, 1.0)
    for flux in (0, 1):
        m = OracleMesh(4, vxy, vid)
        q = np.zeros((4, m.K, m.Np))
        q[0] = 1.2; q[1] = 0.3; q[2] = -0.2; q[3] = 2.5
        # free-stream preservation holds to the accuracy of the cubature table (15 significant
        # digits in gaussTriangleIntegrationDataTable.C), amplified by M^-1: ~1e-12 here
        assert np.abs(m.rhs_euler(q, 1.4, flux)).max() < 2e-11
        x, y = m.nodes()
        q[0] = 1 + 0.2 * np.sin(np.pi * x) * np.cos(2 * np.pi * y)
        q[1] = q[0] * 0.4
        q[3] = 2.5 + 0.1 * np.cos(np.pi * x)
        rhs = m.rhs_euler(q, 1.4, flux)
        # sum over elements of 1^T M_e rhs_e = 0 on a periodic mesh
        o = refelem_ops(4)
        J = m.geo()[:, 4]
        tot = np.einsum("i,ij,vej,e->v", np.ones(m.Np), o["M"], rhs, J)
        assert np.abs(tot).max() < 1e-12


def test_lserk4_translation_convergence():
    """Periodic advection of a smooth profile: error falls with h at high order."""
    errs = []
    for n in (4, 8):
        vxy, vid = split_quads(n, n, -1.0, 1.0, -1.0, 1.0)
        m = OracleMesh(3, vxy, vid)
        x, y = m.nodes()
        q0 = np.sin(np.pi * x) * np.sin(np.pi * y)
        T, steps = 0.25, 50 * n // 4
        q = m.lserk4(0, q0, steps, T / steps, 1.0, 0.5)
        ex = np.sin(np.pi * (x - T)) * np.sin(np.pi * (y - 0.5 * T))
        errs.append(np.abs(q - ex).max())
    assert errs[1] < errs[0] / 10


def test_oracle_refuses_non_contiguous_arrays():
    """The oracle's C functions read raw row-major buffers: a strided view is refused, not misread."""
    vxy, vid = split_quads(3, 3, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    m = OracleMesh(2, vxy, vid, np.ones((vid.shape[0], 3), np.int32))
    x, y = m.nodes()
    q = x + y
    fm = refelem_ops(2)["fmask"]
    bcq = np.stack([q[:, fm[f]] for f in range(3)], axis=1)   # not C-contiguous
    assert not bcq.flags["C_CONTIGUOUS"]
    with pytest.raises(ValueError, match="C-contiguous"):
        m.rhs_advect(q, 1.0, 0.0, bcq)
