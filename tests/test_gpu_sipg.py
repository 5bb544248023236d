"""GPU parity of the matrix-free SIPG operator and its PCG solves against the oracle.

The oracle assembles the reference's SIPG matrix (defaultLaplacianScheme.C:62-288) and runs the
PETSc-order Jacobi CG; the HIP path never forms the matrix.  Bars: operator and right-hand side
relative L2 <= 1e-12 (FP64), CG iteration count within +-1 of the oracle's, and solutions within
the solver tolerance of the oracle's / the exact discrete solution.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from meshgen import split_quads
from oracle_lib import OracleMesh, cg_jacobi, refelem_ops

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def u_sin(x, y):
    return np.sin(2 * np.pi * x) * np.sin(3 * np.pi * y)


def tags_mixed(vxy):
    """Dirichlet (1) on y = 0 / y = 1, zero-gradient (2) on x = 0 / x = 1."""
    t = np.ones(vxy.shape[:2], np.int32)
    for f in range(3):
        mx = 0.5 * (vxy[:, f, 0] + vxy[:, (f + 1) % 3, 0])
        t[:, f] = np.where((np.abs(mx) < 1e-12) | (np.abs(mx - 1) < 1e-12), 2, 1)
    return t


def setup(ctx, N, nx, ny, penalty, tags="dirichlet", ufun=u_sin, lapfun=None, mass_shift=0.0, **kw):
    vxy, vid = split_quads(nx, ny, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    bct = np.ones(vid.shape, np.int32) if tags == "dirichlet" else tags_mixed(vxy)
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    fm = refelem_ops(N)["fmask"]
    g = np.stack([ufun(x[:, fm[f]], y[:, fm[f]]) for f in range(3)], axis=1)
    f = (lapfun or (lambda X, Y: -13 * np.pi ** 2 * u_sin(X, Y)))(x, y)
    A, b = om.sipg_system(penalty, f, g, mass_shift=mass_shift)
    gm = ctx.mesh(N, vxy, vid, bct)
    sp = gm.sipg(penalty=penalty, mass_shift=mass_shift, **kw)
    return dict(om=om, A=A, b=b, f=f, g=g, gm=gm, sp=sp, ue=ufun(x, y))


@pytest.mark.parametrize("N", [1, 3, 5, 8])
@pytest.mark.parametrize("penalty", [0, 1])
def test_apply_matches_oracle(ctx, N, penalty):
    s = setup(ctx, N, 5, 4, penalty, tags="mixed")
    gm, A = s["gm"], s["A"]
    rng = np.random.default_rng(10 * N + penalty)
    u = rng.standard_normal((gm.K, gm.Np))
    fu, fo = gm.field(1, u), gm.field(1)
    s["sp"].apply(fu, fo)
    want = A @ u.ravel()
    e64 = rel(fo.download(), want)
    print("N=%d penalty=%d apply rel L2 %.3e" % (N, penalty, e64))
    assert e64 <= 1e-12
    s["sp"].apply(fu, fo, prec=dgh.DGH_PREC_FP32)
    e32 = rel(fo.download(), want)
    print("  fp32 %.3e" % e32)
    assert e32 <= 2e-5


@pytest.mark.parametrize("N", [2, 5])
def test_rhs_matches_oracle(ctx, N):
    s = setup(ctx, N, 6, 6, 0, tags="mixed")
    gm = s["gm"]
    ff, fb = gm.field(1, s["f"]), gm.field(1)
    s["sp"].rhs(ff, fb, s["g"])
    assert rel(fb.download(), s["b"]) <= 1e-12
    s["sp"].rhs(ff, fb, None)
    _, b0 = s["om"].sipg_system(0, s["f"], None)
    assert rel(fb.download(), b0) <= 1e-12


def roundoff_envelope(A, b, x0, rtol, seeds=16):
    """Oracle iteration counts on copies of A whose entries carry 2e-16 relative noise (the size of the
    differences between the assembled and the matrix-free operator): the spread that round-off alone
    gives the iteration count on this system."""
    its = []
    for seed in range(seeds):
        rng = np.random.default_rng(seed)
        Ap = A.copy()
        Ap.data = Ap.data * (1 + 2e-16 * rng.standard_normal(Ap.data.size))
        its.append(cg_jacobi(Ap.tocsr(), b, x0, rtol=rtol)[1])
    return min(its), max(its)


def check_iterations(A, b, x0, rtol, got, slack=0.25):
    """Iteration-count parity with the oracle.  When round-off cannot move the oracle's own count (its
    preconditioned residual crosses the tolerance cleanly) the GPU count must be within +-1 of it.  When
    the residual hovers near the tolerance, round-off alone moves the count by several iterations; the
    GPU must then stop where the oracle's residual is within `slack` of the tolerance: between the first
    iterate at or below (1 + slack) ttol and the first at or below (1 - slack) ttol."""
    _, its_o, rn0, _, h = cg_jacobi(A, b, x0, rtol=rtol * (1 - slack), maxit=100000, history=True)
    ttol = rtol * rn0
    its_o = next(k for k, v in enumerate(h) if v <= ttol)
    lo_env, hi_env = roundoff_envelope(A, b, x0, rtol)
    if lo_env == hi_env == its_o:
        ok = abs(got - its_o) <= 1
        how = "+-1 of oracle %d" % its_o
    else:
        lo = next(k for k, v in enumerate(h) if v <= (1 + slack) * ttol)
        hi = len(h) - 1
        ok = lo - 1 <= got <= hi + 1
        how = "oracle %d, round-off envelope %d..%d, hover window %d..%d" % (its_o, lo_env, hi_env, lo, hi)
    print("  iterations gpu %d: %s" % (got, how))
    return ok, its_o


@pytest.mark.parametrize("N,n,penalty", [(2, 8, 0), (5, 8, 0), (5, 16, 1), (3, 12, 1), (4, 10, 0), (6, 6, 1)])
def test_cg_jacobi_iterations_match_oracle(ctx, N, n, penalty):
    s = setup(ctx, N, n, n, penalty, rtol=1e-10)
    gm, A, b = s["gm"], s["A"], s["b"]
    x0 = np.zeros(A.shape[0])
    xo, its_o, rn0_o, rn_o = cg_jacobi(A, b, x0, rtol=1e-10)
    fb, fu = gm.field(1, b), gm.field(1, x0)
    st = s["sp"].solve(fb, fu)
    xg = fu.download().ravel()
    print("N=%d n=%d penalty=%d iterations gpu %d oracle %d, rnorm0 %.6e/%.6e, relres %.2e, %.2f ms" % (
        N, n, penalty, st["iters"], its_o, st["rnorm0"], rn0_o, st["relres"], st["ms"]))
    ok, _ = check_iterations(A, b, x0, 1e-10, st["iters"])
    assert ok
    assert st["rnorm0"] == pytest.approx(rn0_o, rel=1e-12)
    assert st["rnorm"] <= 1e-10 * st["rnorm0"]
    assert st["reason"] > 0
    xd = spla.spsolve(A.tocsc(), b)
    assert rel(xg, xd) <= 1e-7
    assert rel(xg, xo) <= 1e-7


def test_cg_initial_guess_and_iteration_limit(ctx):
    s = setup(ctx, 3, 6, 6, 0, rtol=1e-8, abstol=1e-9, maxit=7)
    gm, A, b = s["gm"], s["A"], s["b"]
    fb = gm.field(1, b)
    fu = gm.field(1, spla.spsolve(A.tocsc(), b))
    st = s["sp"].solve(fb, fu, check=False)
    assert st["iters"] == 0 and st["reason"] > 0
    fu.upload(np.zeros(A.shape[0]))
    st = s["sp"].solve(fb, fu, check=False)
    _, its_o, _, _ = cg_jacobi(A, b, np.zeros(A.shape[0]), rtol=1e-8, abstol=1e-9, maxit=7)
    assert st["iters"] == its_o == 7 and st["reason"] == -3
    with pytest.raises(dgh.DghError):
        s["sp"].solve(fb, fu)


@pytest.mark.parametrize("pc", [0, 1])
def test_solver_reuse_replays_captured_iterations(ctx, pc):
    """The 32-iteration chunks are one captured HIP graph per solver and precision, replayed for every chunk of
    every later solve.  Three solves in a row on one solver -- different right-hand sides and fields, with other
    GPU work (an operator application) between them -- each give the oracle CG's iteration count and the same
    bits as a first solve on a fresh solver."""
    s = setup(ctx, 4, 10, 10, 1, pc=pc, rtol=1e-10)
    gm, A, b = s["gm"], s["A"], s["b"]
    rng = np.random.default_rng(3)
    rhs = [b, b + 0.1 * rng.standard_normal(b.shape), 2.0 * b]
    other = gm.field(1, rng.standard_normal(b.shape))
    scratch = gm.field(1)
    got = []
    for r in rhs:
        fb, fu = gm.field(1, r), gm.field(1)
        st = s["sp"].solve(fb, fu)
        s["sp"].apply(other, scratch)
        got.append((st["iters"], fu.download().ravel()))
        if pc == 0:
            ok, _ = check_iterations(A, r, np.zeros(A.shape[0]), 1e-10, st["iters"])
            assert ok
        assert rel(got[-1][1], spla.spsolve(A.tocsc(), r)) <= 1e-7
    for r, (its, x) in zip(rhs, got):
        fresh = gm.sipg(penalty=1, pc=pc, rtol=1e-10)
        fb, fu = gm.field(1, r), gm.field(1)
        st = fresh.solve(fb, fu)
        assert st["iters"] == its
        np.testing.assert_array_equal(fu.download().ravel(), x)
        fresh.destroy()


def test_block_jacobi(ctx):
    s = setup(ctx, 5, 16, 16, 1, pc=dgh.DGH_PC_BLOCK_JACOBI, rtol=1e-10)
    gm, A, b = s["gm"], s["A"], s["b"]
    fb, fu = gm.field(1, b), gm.field(1)
    st = s["sp"].solve(fb, fu)
    _, its_pj, _, _ = cg_jacobi(A, b, np.zeros(A.shape[0]), rtol=1e-10)
    print("block Jacobi: %d iterations (point Jacobi %d), %d distinct blocks" % (st["iters"], its_pj,
                                                                                 s["sp"].nblocks))
    assert st["reason"] > 0 and st["iters"] < its_pj
    assert s["sp"].nblocks < gm.K
    assert rel(fu.download().ravel(), spla.spsolve(A.tocsc(), b)) <= 1e-7


@pytest.mark.parametrize("pc", [0, 1])
def test_mixed_precision_refinement(ctx, pc):
    s = setup(ctx, 5, 16, 16, 1, pc=pc, mode=dgh.DGH_SOLVE_MIXED, rtol=1e-10, maxit=100000)
    gm, A, b = s["gm"], s["A"], s["b"]
    fb, fu = gm.field(1, b), gm.field(1)
    st = s["sp"].solve(fb, fu)
    xg = fu.download().ravel()
    true_rel = np.linalg.norm(b - A @ xg) / np.linalg.norm(b)
    print("mixed pc=%d: %d inner iterations, %d sweeps, relres %.2e (oracle-matrix %.2e)" % (
        pc, st["iters"], st["outer"], st["relres"], true_rel))
    assert st["reason"] > 0 and st["relres"] <= 1e-10
    assert true_rel <= 2e-10
    assert rel(xg, spla.spsolve(A.tocsc(), b)) <= 1e-8


def test_polynomial_exact_on_gpu(ctx):
    c = [0.4, -0.8, 0.3, 0.5, -0.25, 0.7, 0.2, -0.1, 0.45, -0.3]

    def u(x, y):
        return (c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y + c[6] * x ** 3 +
                c[7] * x * x * y + c[8] * x * y * y + c[9] * y ** 3)

    def lap(x, y):
        return 2 * c[3] + 2 * c[5] + 6 * c[6] * x + 2 * c[7] * y + 2 * c[8] * x + 6 * c[9] * y

    s = setup(ctx, 3, 5, 5, 0, ufun=u, lapfun=lap, rtol=1e-13, abstol=0.0)
    gm = s["gm"]
    ff, fb, fu = gm.field(1, s["f"]), gm.field(1), gm.field(1)
    s["sp"].rhs(ff, fb, s["g"])
    s["sp"].solve(fb, fu)
    assert np.abs(fu.download() - s["ue"]).max() <= 1e-9


def test_ragged_K(ctx):
    # K = 2 * 7 * 5 = 70: the last group of 16 elements is partly padding
    s = setup(ctx, 4, 7, 5, 1, tags="mixed", rtol=1e-10)
    gm, A, b = s["gm"], s["A"], s["b"]
    assert gm.K % 16 != 0
    fb, fu = gm.field(1, b), gm.field(1)
    st = s["sp"].solve(fb, fu)
    _, its_o, rn0_o, _ = cg_jacobi(A, b, np.zeros(A.shape[0]), rtol=1e-10)
    print("ragged K=%d" % gm.K)
    ok, _ = check_iterations(A, b, np.zeros(A.shape[0]), 1e-10, st["iters"])
    assert ok
    assert st["rnorm0"] == pytest.approx(rn0_o, rel=1e-12)


@pytest.mark.parametrize("mode,pc", [(0, 0), (1, 1)])
def test_poisson_128x128_N5(ctx, mode, pc):
    """The 128 x 128 split-quad Poisson problem at N = 5 (K = 32768), penalty (N+1)^2/h: sizes the
    oracle cannot solve quickly, so the checks are the FP64 residual and the exact solution."""
    gm = ctx.mesh_structured(5, 128, 128, 0.0, 1.0, 0.0, 1.0, periodic=(False, False),
                             bc_tag=dgh.DGH_BC_FIXED_VALUE)
    x, y = gm.nodes()
    ue = u_sin(x, y)
    fm = refelem_ops(5)["fmask"]
    g = np.stack([ue[:, fm[f]] for f in range(3)], axis=1)
    sp = gm.sipg(penalty=dgh.DGH_SIPG_PENALTY_H, pc=pc, mode=mode, rtol=1e-10, maxit=200000)
    ff, fb, fu = gm.field(1, -13 * np.pi ** 2 * ue), gm.field(1), gm.field(1)
    sp.rhs(ff, fb, g)
    st = sp.solve(fb, fu)
    err = np.abs(fu.download() - ue).max()
    print("128x128 N=5 mode=%d pc=%d: %d iterations (%d sweeps) %.1f ms, relres %.2e, max error %.2e" % (
        mode, pc, st["iters"], st["outer"], st["ms"], st["relres"], err))
    assert st["reason"] > 0
    if mode == 1:
        assert st["relres"] <= 1e-10
    assert err <= 1e-7


@pytest.mark.gpu
def test_groups_beyond_one_pass(ctx):
    """A mesh with more element groups (4608, 192 x 192 split quads at N = 2) than the persistent grids have
    wavefronts (4096 on a 256-CU device), so some wavefronts take a second group: the operator stays symmetric
    (u.Av = v.Au in FP64 and FP32), and the point- and block-Jacobi solves and the mixed refinement converge to
    the same solution, near the exact one."""
    N, n = 2, 192
    gm = ctx.mesh_structured(N, n, n, 0.0, 1.0, 0.0, 1.0, periodic=(False, False), bc_tag=dgh.DGH_BC_FIXED_VALUE)
    assert (gm.K + 15) // 16 > 4096
    x, y = gm.nodes()
    ue = u_sin(x, y)
    rng = np.random.default_rng(5)
    uu, vv = rng.standard_normal(ue.shape), rng.standard_normal(ue.shape)
    sp = gm.sipg(penalty=dgh.DGH_SIPG_PENALTY_H, pc=0, rtol=1e-10, maxit=200000)
    fu, fv, fo = gm.field(1, uu), gm.field(1, vv), gm.field(1)
    for prec, tol in ((dgh.DGH_PREC_FP64, 1e-12), (dgh.DGH_PREC_FP32, 1e-5)):
        sp.apply(fu, fo, prec)
        au = fo.download()
        sp.apply(fv, fo, prec)
        av = fo.download()
        a, b = float(np.sum(vv * au)), float(np.sum(uu * av))
        print("prec %d: v.Au %.12e u.Av %.12e" % (prec, a, b))
        assert abs(a - b) <= tol * np.sqrt(np.sum(au * au) * np.sum(vv * vv))
    fm = refelem_ops(N)["fmask"]
    g = np.stack([ue[:, fm[f]] for f in range(3)], axis=1)
    ff, fb = gm.field(1, -13 * np.pi ** 2 * ue), gm.field(1)
    sols = []
    for mode, pc in ((dgh.DGH_SOLVE_FP64, dgh.DGH_PC_JACOBI), (dgh.DGH_SOLVE_FP64, dgh.DGH_PC_BLOCK_JACOBI),
                     (dgh.DGH_SOLVE_MIXED, dgh.DGH_PC_BLOCK_JACOBI)):
        s2 = gm.sipg(penalty=dgh.DGH_SIPG_PENALTY_H, pc=pc, mode=mode, rtol=1e-10, maxit=200000)
        s2.rhs(ff, fb, g)
        fx = gm.field(1)
        st = s2.solve(fb, fx)
        u = fx.download()
        err = np.abs(u - ue).max()
        print("mode %d pc %d: %d iterations (%d sweeps), relres %.2e, max error %.2e" % (
            mode, pc, st["iters"], st["outer"], st["relres"], err))
        assert st["reason"] > 0 and err <= 1e-5   # discretisation error 2.1e-6
        sols.append(u)
        s2.destroy()
    for u in sols[1:]:
        assert np.abs(u - sols[0]).max() <= 1e-6


@pytest.mark.parametrize("N", [1, 5, 8])
def test_apply_with_mass_shift_matches_oracle(ctx, N):
    """Helmholtz operator A + sigma M (the implicit viscous step of the Chorin solver): the apply against the
    oracle's shifted assembled matrix."""
    s = setup(ctx, N, 5, 4, 1, tags="mixed", mass_shift=3.0)
    gm, A = s["gm"], s["A"]
    u = np.random.default_rng(N).standard_normal((gm.K, gm.Np))
    fu, fo = gm.field(1, u), gm.field(1)
    s["sp"].apply(fu, fo)
    e64 = rel(fo.download(), A @ u.ravel())
    s["sp"].apply(fu, fo, prec=dgh.DGH_PREC_FP32)
    e32 = rel(fo.download(), A @ u.ravel())
    print("N=%d mass shift apply rel L2 %.3e, fp32 %.3e" % (N, e64, e32))
    assert e64 <= 1e-12 and e32 <= 2e-5


@pytest.mark.parametrize("sigma", [10.0, 1000.0])
def test_cg_helmholtz_matches_oracle(ctx, sigma):
    """PCG on -lap u + sigma u = g (u = sin(2 pi x) sin(3 pi y), Dirichlet): iteration count of the oracle's
    PETSc-order CG on the shifted assembled system within the round-off envelope, the same solution, and the
    manufactured solution to discretisation error.  A larger shift conditions the system better."""
    lapf = lambda X, Y: -(13 * np.pi ** 2 + sigma) * u_sin(X, Y)   # noqa: E731
    s = setup(ctx, 5, 8, 8, 0, lapfun=lapf, mass_shift=sigma, rtol=1e-10)
    gm, A, b = s["gm"], s["A"], s["b"]
    x0 = np.zeros(A.shape[0])
    xo, its_o, rn0_o, _ = cg_jacobi(A, b, x0, rtol=1e-10)
    fb, fu = gm.field(1, b), gm.field(1, x0)
    st = s["sp"].solve(fb, fu)
    xg = fu.download().ravel()
    err = np.abs(xg - s["ue"].ravel()).max()
    print("sigma=%g: iterations gpu %d oracle %d, max error vs exact %.2e" % (sigma, st["iters"], its_o, err))
    ok, _ = check_iterations(A, b, x0, 1e-10, st["iters"])
    assert ok
    assert st["rnorm0"] == pytest.approx(rn0_o, rel=1e-12)
    assert rel(xg, spla.spsolve(A.tocsc(), b)) <= 1e-7
    assert err <= 1e-4
    # block Jacobi and the mixed solve reach the same solution
    for kw in ({"pc": dgh.DGH_PC_BLOCK_JACOBI}, {"pc": dgh.DGH_PC_BLOCK_JACOBI, "mode": dgh.DGH_SOLVE_MIXED}):
        sp2 = gm.sipg(penalty=0, mass_shift=sigma, rtol=1e-10, **kw)
        fu2 = gm.field(1, x0)
        st2 = sp2.solve(fb, fu2)
        print("  %s: %d iterations, relres %.2e" % (kw, st2["iters"], st2["relres"]))
        assert rel(fu2.download().ravel(), xg) <= 1e-7
        sp2.destroy()
