"""Pins of the projection-solver oracle (oracle/chorin_ref.py), CPU only.

The reference publishes no numbers for dgChorinFoam (the User Guide gives time-step tables only), so its pieces
are pinned by what the mathematics fixes:
  * the dof-field LF convection divides exactly a quadratic flux (linear U, N >= 2) given exact boundary data, and
    conserves momentum on a periodic mesh (every face flux enters its two elements with opposite signs);
  * the nodal gradient and divergence differentiate degree-N polynomials exactly;
  * the SIPG right-hand side with Neumann (fixedGradient) data reproduces degree-N polynomial Poisson solutions
    exactly with Dirichlet data on some boundary faces and the exact normal derivative on the others;
  * the whole scheme converges to the exact Taylor-Green vortex at the rate of the discretisation on the
    tutorial's three meshes.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import chorin_ref as cr  # noqa: E402

from hopefoam_amd import caseio  # noqa: E402
from meshgen import split_quads  # noqa: E402
from oracle_lib import OracleMesh, refelem_ops  # noqa: E402

CASE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "pearsonVortex")


def _mesh(N, nx=4, ny=3, periodic=(False, False), tag=1):
    vxy, vid = split_quads(nx, ny, -0.5, 0.5, -0.4, 0.6, periodic=periodic)
    om = OracleMesh(N, vxy, vid, np.full(vid.shape, tag, np.int32))
    return om, refelem_ops(N)


@pytest.mark.parametrize("N", [2, 3, 5])
def test_convection_exact_for_quadratic_flux(N):
    """U linear: U U is quadratic, so its nodal interpolant is exact, the traces are continuous (the LF jump
    terms vanish) and the weak divergence with exact boundary data is the exact divergence; after the mass
    solve convecU holds div(U U) = U (div U) + (U . grad) U at the nodes."""
    om, ops = _mesh(N)
    x, y = om.nodes()
    a = (0.3, -0.7, 0.4)
    b = (-0.2, 0.5, 0.9)
    U = np.stack([a[0] + a[1] * x + a[2] * y, b[0] + b[1] * x + b[2] * y])
    fm = ops["fmask"]
    bcU = np.stack([np.stack([U[c][:, fm[f]] for f in range(3)], axis=1) for c in range(2)])
    E, F = om.conn()
    got = cr.convection(ops, om.geo(), E, F, U, bcU, E < 0)
    div = a[1] + b[2]
    want = np.stack([U[0] * div + U[0] * a[1] + U[1] * a[2], U[1] * div + U[0] * b[1] + U[1] * b[2]])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11)


def test_convection_conserves_momentum_periodic():
    """On a periodic mesh the integral of div(U U) is zero for any U (each face flux leaves one element and
    enters the other): sum_e 1^T J_e M convecU_e = 0, with the LF dissipation active (discontinuous U)."""
    N = 3
    om, ops = _mesh(N, 5, 4, periodic=(True, True))
    rng = np.random.default_rng(3)
    U = rng.standard_normal((2, om.K, om.Np))
    E, F = om.conn()
    got = cr.convection(ops, om.geo(), E, F, U)
    J = om.geo()[:, 4]
    tot = np.einsum("cej,jk,e->c", got, ops["M"], J)
    scale = np.abs(np.einsum("cej,jk,e->cek", got, ops["M"], J)).sum()
    assert np.all(np.abs(tot) < 1e-13 * scale), tot


@pytest.mark.parametrize("N", [1, 3, 6])
def test_nodal_derivatives_exact(N):
    om, ops = _mesh(N)
    x, y = om.nodes()
    u = x ** N - 2 * x * y ** (N - 1) + 0.5 * y + 1.0
    ux = N * x ** (N - 1) - 2 * y ** (N - 1)
    uy = (-2 * (N - 1) * x * y ** (N - 2) if N > 1 else 0.0 * x) + 0.5
    g = cr.nodal_grad(ops, om.geo(), u)
    np.testing.assert_allclose(g[0], ux, atol=1e-11)
    np.testing.assert_allclose(g[1], uy, atol=1e-11)
    V = np.stack([u, x * y ** (N - 1)])
    dv = (N - 1) * x * y ** (N - 2) if N > 1 else 0.0 * x
    np.testing.assert_allclose(cr.nodal_div(ops, om.geo(), V), ux + dv, atol=1e-11)


@pytest.mark.parametrize("N", [2, 4])
def test_neumann_data_reproduce_polynomials(N):
    """-laplacian u = f with u of degree N: Dirichlet data on the faces with x = -0.5 or y = 0.6 (tag 1), the exact
    normal derivative on the other boundary faces (fixedGradient, b -= Vf^T W Vf g): the solve reproduces u at
    the nodes."""
    vxy, vid = split_quads(4, 3, -0.5, 0.5, -0.4, 0.6, periodic=(False, False))
    tag = np.full(vid.shape, 2, np.int32)
    for e in range(len(vid)):
        for f in range(3):
            m = 0.5 * (vxy[e, f] + vxy[e, (f + 1) % 3])
            if abs(m[0] + 0.5) < 1e-12 or abs(m[1] - 0.6) < 1e-12:
                tag[e, f] = 1
    om = OracleMesh(N, vxy, vid, tag)
    ops = refelem_ops(N)
    x, y = om.nodes()
    fm = ops["fmask"]
    u = x ** N + 2 * x * y - y ** 2 + 0.3
    lap = N * (N - 1) * x ** (N - 2) - 2.0
    gD = np.stack([u[:, fm[f]] for f in range(3)], axis=1)
    geo = om.geo()
    xb = np.stack([x[:, fm[f]] for f in range(3)], axis=1)
    yb = np.stack([y[:, fm[f]] for f in range(3)], axis=1)
    ux = N * xb ** (N - 1) + 2 * yb
    uy = 2 * xb - 2 * yb
    gN = ux * geo[:, 5::3][:, :, None] + uy * geo[:, 6::3][:, :, None]
    E, _ = om.conn()
    A, b = om.sipg_system(0, lap, gD)
    neu = (E < 0) & (tag != 1)
    assert neu.any() and ((E < 0) & (tag == 1)).any()
    b = b - cr.face_mass_term(ops, geo, E, gN, neu).reshape(-1)
    import scipy.sparse.linalg as spla
    sol = spla.spsolve(A.tocsc(), b).reshape(u.shape)
    np.testing.assert_allclose(sol, u, atol=1e-9)
    # and without the Neumann term the solution is wrong
    sol0 = spla.spsolve(A.tocsc(), om.sipg_system(0, lap, gD)[1]).reshape(u.shape)
    assert np.abs(sol0 - u).max() > 1e-3


def chorin_oracle(meshname, N=2, nsteps=100):
    case = caseio.Case(CASE, mesh=meshname)
    m = case.mesh2d()
    om_p = OracleMesh(N, m.vertices(), m.tris, case.bc_tags(["p"]))
    om_U = OracleMesh(N, m.vertices(), m.tris, case.bc_tags(["U"]))
    tp = caseio.read_dict(os.path.join(CASE, "constant", "transportProperties"))
    ref = cr.ChorinRef(om_p, om_U, refelem_ops(N), float(tp["nu"][-1]), float(case.control["deltaT"]))
    for _ in range(nsteps):
        ref.step()
    return ref


def test_scheme_converges_to_taylor_green_vortex():
    """The tutorial's case (N = 2, dt = 0.001, 100 steps to t = 0.1, nu = 0.01) on its meshes h, h/2, h/4: the
    pressure and velocity errors against the exact solution fall at about the order N + 1 = 3 of the
    discretisation (measured 2.74 / 2.80 for p, 1.61 / 2.99 for U; the coarsest mesh has 14 triangles)."""
    errs = [chorin_oracle(m).errors() for m in ("vortexh.msh", "vortex0.5h.msh", "vortex0.25h.msh")]
    ep = [e[0] for e in errs]
    eu = [e[2] for e in errs]
    print("mean |p - p_exact|", ep, "mean |U - U_exact|", eu)
    assert ep[0] > ep[1] > ep[2] and eu[0] > eu[1] > eu[2]
    assert np.log2(ep[1] / ep[2]) > 2.5 and np.log2(eu[1] / eu[2]) > 2.5
    assert ep[2] < 4e-3 and eu[2] < 3e-3


@pytest.mark.parametrize("N", [1, 2])
def test_convection_face_fluxes_closed_form(N):
    """Two triangles of the unit square with different linear velocities (a jump on the diagonal) and boundary
    data equal to their own traces: the integral of convecU over each triangle is the net outward numerical flux
    (the volume term of a weak divergence integrates to zero against the constant), sum_f |e_f| sum_k w_k F*_f(k)
    with the face-node rule w (trapezoid for N = 1, Simpson for N = 2: exact for the interpolated face flux),
    F* = (n.U) U on the outer faces and on the diagonal the LF flux 1/2 (vO UO + vN UN) + 1/2 maxV (UO - UN) with
    one maxV per face: the largest |n.U| over the face nodes of both sides.  Pins the sign and the face-wise
    dissipation speed of the LF flux, which the smooth-data tests cannot see."""
    vxy = np.array([[[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]], [[0.0, 0.0], [1.0, 1.0], [0.0, 1.0]]])
    vid = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    om = OracleMesh(N, vxy, vid, np.ones((2, 3), np.int32))
    ops = refelem_ops(N)
    x, y = om.nodes()
    lin = [lambda x, y: np.stack([0.6 + 0.5 * x - 0.2 * y, -0.3 + 0.4 * x + 0.9 * y]),
           lambda x, y: np.stack([-0.8 + 0.1 * x + 0.7 * y, 0.5 - 0.6 * x + 0.2 * y])]
    U = np.stack([lin[e](x[e], y[e]) for e in range(2)], axis=1)
    fm = ops["fmask"]
    bcU = np.stack([np.stack([U[c][:, fm[f]] for f in range(3)], axis=1) for c in range(2)])
    E, F = om.conn()
    got = cr.convection(ops, om.geo(), E, F, U, bcU, E < 0)
    J = om.geo()[:, 4]
    w = {1: [0.5, 0.5], 2: [1 / 6, 4 / 6, 1 / 6]}[N]
    for e in range(2):
        integral = J[e] * (ops["M"] @ got[:, e].T).sum(axis=0)
        want = np.zeros(2)
        for f in range(3):
            a, b = vxy[e, f], vxy[e, (f + 1) % 3]
            L = np.hypot(*(b - a))
            n = np.array([b[1] - a[1], a[0] - b[0]]) / L          # outward for anticlockwise vertices
            pts = [a + (b - a) * s for s in np.linspace(0.0, 1.0, N + 1)]
            Uo = np.array([lin[e](*p) for p in pts])                # [k][2]
            if E[e, f] >= 0:
                Un = np.array([lin[E[e, f]](*p) for p in pts])
                vo, vn = Uo @ n, Un @ n
                mv = max(np.abs(vo).max(), np.abs(vn).max())
                Fs = 0.5 * (vo[:, None] * Uo + vn[:, None] * Un) + 0.5 * mv * (Uo - Un)
            else:
                Fs = (Uo @ n)[:, None] * Uo
            want += L * np.einsum("k,kc->c", np.array(w), Fs)
        np.testing.assert_allclose(integral, want, rtol=1e-12, atol=1e-13)


def _case_ref(meshname, N, dt):
    case = caseio.Case(CASE, mesh=meshname)
    m = case.mesh2d()
    om_p = OracleMesh(N, m.vertices(), m.tris, case.bc_tags(["p"]))
    om_U = OracleMesh(N, m.vertices(), m.tris, case.bc_tags(["U"]))
    return cr.ChorinRef(om_p, om_U, refelem_ops(N), 0.01, dt)


def test_pressure_neumann_data_vanish_for_the_exact_vortex():
    """pEqnCorrect's nbp = n . (convecU + nu curl curl U + du/dt) is -dp/dn; on the tutorial's inlet (x = +-1/2,
    y = +-1/2 sides) the Taylor-Green pressure has dp/dn = 0, and nu curl curl u cancels du/dt exactly.  With the
    exact velocity at N = 5 the discrete nbp is at the discretisation level (1.5e-3); a flipped curl, a flipped
    du/dt term or a missing viscous term leaves 8 pi^2 nu |n.u| ~ 0.4 - 0.8."""
    ref = _case_ref("vortex0.25h.msh", 5, 0.001)
    t = 0.05
    U = cr.exact_u(ref.x, ref.y, t, 0.01)
    fm = ref.ops["fmask"]
    bcU = np.stack([np.stack([U[c][:, fm[f]] for f in range(3)], axis=1) for c in range(2)])
    cu = cr.convection(ref.ops, ref.geo, ref.EToE, ref.EToF, U, bcU, ref.inlet)
    nbp = ref.pressure_neumann(U, cu, np.exp(-4 * np.pi ** 2 * 0.01 * t))
    assert ref.inlet.any()
    assert np.abs(nbp[ref.inlet]).max() < 5e-3
    assert np.abs(nbp[~ref.inlet]).max() == 0.0


def test_scheme_second_order_in_time():
    """BDF2 / second-order extrapolation after the first step: at N = 5 on the finest mesh (spatial error far
    below the temporal one) the errors at t = 0.1 fall at about dt^2 for dt = 0.02, 0.01, 0.005 (measured p 2.17 /
    1.83, U 2.09 / 2.04).  Without the extrapolated pressure Neumann data (b0 nbp + b1 nbp_old) the pressure error
    grows as dt shrinks."""
    errs = []
    for dt in (0.02, 0.01, 0.005):
        ref = _case_ref("vortex0.25h.msh", 5, dt)
        for _ in range(int(round(0.1 / dt))):
            ref.step()
        errs.append(ref.errors())
    e = np.array(errs)
    rp = np.log2(e[:-1, 0] / e[1:, 0])
    ru = np.log2(e[:-1, 2] / e[1:, 2])
    print("temporal rates p", rp, "U", ru)
    assert np.all(rp > 1.6) and np.all(ru > 1.8)
