"""The projection solver dgChorinFoam on the GPU (tutorials/DG/2D/pearsonVortex): its operators against the oracle
(oracle/chorin_ref.py), a few whole steps against the oracle's direct-solve run, and the tutorial run on its three
meshes against the exact Taylor-Green vortex."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import chorin_ref as cr  # noqa: E402

from meshgen import split_quads  # noqa: E402
from oracle_lib import OracleMesh, refelem_ops  # noqa: E402

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")
caseio = pytest.importorskip("hopefoam_amd.caseio")
chorin = pytest.importorskip("hopefoam_amd.chorin")

CASE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "pearsonVortex")


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("N,periodic", [(2, True), (4, True), (3, False), (6, False)])
def test_convection_matches_oracle(ctx, N, periodic):
    """div(U U) with the LF flux on a discontinuous random velocity: a periodic mesh, and an open one whose left
    and bottom sides are fixedValue (the ghost states) and the others copy the trace."""
    vxy, vid = split_quads(6, 5, -0.5, 0.5, -0.5, 0.5, periodic=(periodic, periodic))
    tag = np.full(vid.shape, dgh.DGH_BC_ZERO_GRADIENT, np.int32)
    if not periodic:
        for e in range(len(vid)):
            for f in range(3):
                m = 0.5 * (vxy[e, f] + vxy[e, (f + 1) % 3])
                if abs(m[0] + 0.5) < 1e-12 or abs(m[1] + 0.5) < 1e-12:
                    tag[e, f] = dgh.DGH_BC_FIXED_VALUE
    om = OracleMesh(N, vxy, vid, tag)
    gm = ctx.mesh(N, vxy, vid, tag)
    ops = refelem_ops(N)
    rng = np.random.default_rng(N)
    U = rng.standard_normal((2, om.K, om.Np))
    E, F = om.conn()
    fixed = (E < 0) & (tag == dgh.DGH_BC_FIXED_VALUE)
    bcU = rng.standard_normal((2, om.K, 3, om.Nfp)) * fixed[None, :, :, None]
    want = cr.convection(ops, om.geo(), E, F, U, bcU, fixed)
    if fixed.any():
        gm.set_bc_values(np.ascontiguousarray(bcU))
    fu, fo = gm.field(2, U), gm.field(2)
    gm.ins_convection(fu, fo)
    err = rel(fo.download(), want)
    print("N=%d periodic=%s convection rel L2 %.3e" % (N, periodic, err), flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("N", [1, 3, 5, 8])
def test_nodal_derivatives_match_oracle(ctx, N):
    vxy, vid = split_quads(5, 4, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    om = OracleMesh(N, vxy, vid)
    gm = ctx.mesh(N, vxy, vid, np.full(vid.shape, 2, np.int32))
    ops = refelem_ops(N)
    rng = np.random.default_rng(10 + N)
    u = rng.standard_normal((om.K, om.Np))
    V = rng.standard_normal((2, om.K, om.Np))
    fu, fg = gm.field(1, u), gm.field(2)
    gm.nodal_grad(fu, fg)
    fv, fd = gm.field(2, V), gm.field(1)
    gm.nodal_div(fv, fd)
    assert rel(fg.download(), cr.nodal_grad(ops, om.geo(), u)) <= 1e-13
    assert rel(fd.download(), cr.nodal_div(ops, om.geo(), V)) <= 1e-13


@pytest.mark.parametrize("N", [2, 5])
def test_sipg_rhs_neumann_matches_oracle(ctx, N):
    """b = M f + Dirichlet terms - Vf^T W Vf g on the Neumann faces, against the oracle's assembled right-hand side
    and its face mass term."""
    vxy, vid = split_quads(5, 4, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    tag = np.full(vid.shape, dgh.DGH_BC_ZERO_GRADIENT, np.int32)
    tag[::3] = dgh.DGH_BC_FIXED_VALUE
    om = OracleMesh(N, vxy, vid, tag)
    gm = ctx.mesh(N, vxy, vid, tag)
    ops = refelem_ops(N)
    rng = np.random.default_rng(N)
    f = rng.standard_normal((om.K, om.Np))
    gD = rng.standard_normal((om.K, 3, om.Nfp))
    gN = rng.standard_normal((om.K, 3, om.Nfp))
    E, _ = om.conn()
    neu = (E < 0) & (tag != dgh.DGH_BC_FIXED_VALUE)
    assert neu.any()
    _, b = om.sipg_system(0, f, gD)
    want = b - cr.face_mass_term(ops, om.geo(), E, gN, neu).reshape(-1)
    S = gm.sipg()
    ff, fb = gm.field(1, f), gm.field(1)
    S.rhs(ff, fb, gbc=gD, gnb=gN)
    assert rel(fb.download().ravel(), want) <= 1e-13
    S.rhs(ff, fb, gbc=gD)
    assert rel(fb.download().ravel(), b) <= 1e-13


def test_field_lincomb_and_copy_var(ctx):
    vxy, vid = split_quads(3, 3, 0.0, 1.0, 0.0, 1.0)
    gm = ctx.mesh(3, vxy, vid)
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal((2, 2, gm.K, gm.Np))
    fa, fb, fo = gm.field(2, a), gm.field(2, b), gm.field(2)
    fo.lincomb([(2.0, fa), (-0.5, fb)])
    np.testing.assert_allclose(fo.download(), 2 * a - 0.5 * b, rtol=0, atol=1e-15)
    fa.lincomb([(3.0, fa)])
    np.testing.assert_allclose(fa.download(), 3 * a, rtol=0, atol=1e-15)
    s = gm.field(1)
    s.copy_var(0, fb, 1)
    np.testing.assert_array_equal(s.download(), b[1])
    fo.copy_var(0, s, 0)
    np.testing.assert_array_equal(fo.download()[0], b[1])


def _oracle(meshname, N, nsteps):
    case = caseio.Case(CASE, mesh=meshname)
    m = case.mesh2d()
    om_p = OracleMesh(N, m.vertices(), m.tris, case.bc_tags(["p"]))
    om_U = OracleMesh(N, m.vertices(), m.tris, case.bc_tags(["U"]))
    ref = cr.ChorinRef(om_p, om_U, refelem_ops(N), 0.01, float(case.control["deltaT"]))
    for _ in range(nsteps):
        ref.step()
    return ref


@pytest.mark.parametrize("meshname", ["vortexh.msh", "vortex0.5h.msh"])
def test_steps_match_oracle(ctx, meshname):
    """Five steps (the first-order start and four BDF2 steps) of the device solver against the oracle's
    direct-solve run: U and p to 1e-10 relative."""
    sol = chorin.ChorinSolver(ctx, CASE, mesh=meshname)
    sol.run(5)
    ref = _oracle(meshname, sol.N, 5)
    U, p = sol.fields()
    eu, ep = rel(U, ref.U), rel(p, ref.p)
    its = [[s["iters"] for s in st] for st in sol.stats]
    print("%s: 5 steps, U rel %.3e, p rel %.3e; PCG iterations per step %s" % (meshname, eu, ep, its), flush=True)
    assert eu <= 1e-10 and ep <= 1e-10


def test_tutorial_run_converges(ctx):
    """The tutorial as its files set it (N = 2, dt = 0.001, end time 0.1, nu = 0.01) on vortexh, vortex0.5h and
    vortex0.25h: the pressure and velocity errors against the exact solution (chorinerror.H's mean |p - p_exact|
    and the same for U) equal the oracle's to 1e-8 relative and fall at the rate of the discretisation."""
    errs = []
    for name in ("vortexh.msh", "vortex0.5h.msh", "vortex0.25h.msh"):
        sol = chorin.ChorinSolver(ctx, CASE, mesh=name)
        sol.run()
        assert abs(sol.t - 0.1) < 1e-12
        e = sol.errors()
        ref = _oracle(name, sol.N, 100).errors()
        print("%s: p mean %.4e (oracle %.4e) max %.4e; U mean %.4e (oracle %.4e)" %
              (name, e[0], ref[0], e[1], e[2], ref[2]), flush=True)
        assert abs(e[0] / ref[0] - 1) < 1e-8 and abs(e[2] / ref[2] - 1) < 1e-8
        errs.append(e)
    e = np.array(errs)
    rp, ru = np.log2(e[:-1, 0] / e[1:, 0]), np.log2(e[:-1, 2] / e[1:, 2])
    print("rates p %s U %s" % (rp, ru), flush=True)
    assert rp[-1] > 2.5 and ru[-1] > 2.5
