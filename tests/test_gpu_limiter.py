"""GPU parity of the TVB minmod limiter against the oracle (same troubled cells, limited values to 1e-13)."""
import numpy as np
import pytest

from meshgen import split_quads
from oracle_lib import OracleMesh, refelem_ops

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")
cases = pytest.importorskip("hopefoam_amd.cases")


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def compare(ctx, N, vxy, vid, bct, q, M, nu=1.5, bcq=None):
    om = OracleMesh(N, vxy, vid, bct)
    want, tr_o = om.tvb_limit(q, M=M, nu=nu, bcq=bcq)
    gm = ctx.mesh(N, vxy, vid, bct)
    if bcq is not None:
        gm.set_bc_values(bcq)
    nvar = q.size // (gm.K * gm.Np)
    f = gm.field(nvar, q)
    tr_g, n = gm.limit(f, M=M, nu=nu)
    got = f.download()
    diff = np.flatnonzero(tr_g != tr_o)
    print("N=%d K=%d nvar=%d: troubled gpu %d oracle %d, %d differ" % (N, gm.K, nvar, n, tr_o.sum(), diff.size))
    assert n == tr_g.sum()
    assert diff.size == 0
    scale = max(1.0, np.abs(q).max())
    err = np.abs(got - want).max() / scale
    print("  max |gpu - oracle| / scale = %.3e" % err)
    assert err <= 1e-13
    return tr_g


@pytest.mark.parametrize("N", [1, 3, 6, 8])
def test_scalar_periodic(ctx, N):
    vxy, vid = split_quads(12, 10, 0.0, 1.0, 0.0, 1.0, periodic=(True, True))
    om = OracleMesh(N, vxy, vid)
    x, y = om.nodes()
    rng = np.random.default_rng(N)
    q = np.where(x + 0.37 * y < 0.61, 1.5, 0.2) + 0.05 * np.sin(7 * x) * np.cos(5 * y) + 1e-3 * rng.standard_normal(x.shape)
    tr = compare(ctx, N, vxy, vid, None, q, M=1.0)
    assert tr.any() and not tr.all()


@pytest.mark.parametrize("tagname", ["zero_gradient", "wall", "fixed"])
def test_euler_riemann_boundaries(ctx, tagname):
    N = 3
    vxy, vid = split_quads(24, 6, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    tag = {"zero_gradient": 2, "wall": 3, "fixed": 1}[tagname]
    bct = np.full(vid.shape, tag, np.int32)
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    q = cases.riemann4(x, y, seed=11)
    bcq = None
    if tag == 1:
        fm = refelem_ops(N)["fmask"]
        bcq = np.ascontiguousarray(np.stack([np.stack([q[v][:, fm[f]] for f in range(3)], axis=1)
                                             for v in range(4)]) * 1.3)
    tr = compare(ctx, N, vxy, vid, bct, q, M=50.0, bcq=bcq)
    assert tr.any()


def test_riemann_512x128_N3(ctx):
    """The limiter configuration's mesh (K = 131072, N = 3) on seeded four-quadrant data."""
    N = 3
    vxy, vid = split_quads(512, 128, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, 2, np.int32)
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    q = cases.riemann4(x, y, seed=1000)
    compare(ctx, N, vxy, vid, bct, q, M=50.0)


@pytest.mark.parametrize("N", [2, 3, 6])
def test_positivity_matches_oracle(ctx, N):
    vxy, vid = split_quads(24, 8, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, 2, np.int32)
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    q, _ = om.tvb_limit(cases.riemann4(x, y, seed=11), M=50.0)
    want, sc = om.positivity(q)
    gm = ctx.mesh(N, vxy, vid, bct)
    f = gm.field(4, q)
    # M = 1e30: the TVB step touches nothing, only the positivity pass acts
    gm.limit(f, M=1e30, positivity=True, flags=False)
    got = f.download()
    print("N=%d: %d cells scaled by the oracle, max |gpu - oracle| %.2e" % (N, sc.sum(), np.abs(got - want).max()))
    assert sc.any()
    changed = np.any(np.abs(got - q) > 0, axis=(0, 2))
    assert np.array_equal(changed, sc)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(q).max()
    rho, mx, my, E = got
    Vp = om.positivity_points()
    r_pts = rho @ Vp.T
    p_pts = 0.4 * ((E @ Vp.T) - 0.5 * ((mx @ Vp.T) ** 2 + (my @ Vp.T) ** 2) / r_pts)
    assert r_pts.min() > 0 and p_pts.min() > 0


def test_limited_euler_steps(ctx):
    """Limiter (TVB + positivity) after every LSERK4 stage of a Roe Euler run from Riemann data: GPU (stage
    kernel + limiter) against the same sequence on the GPU with the limiter replaced by the oracle's."""
    N = 2
    vxy, vid = split_quads(32, 8, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, 2, np.int32)
    om = OracleMesh(N, vxy, vid, bct)
    gm = ctx.mesh(N, vxy, vid, bct)
    x, y = om.nodes()

    def oracle_limit(q):
        q, _ = om.tvb_limit(q, M=50.0)
        return om.positivity(q)[0]

    # discontinuous data must be limited before the first RHS (unlimited, the cubature points of cut
    # cells see negative pressure)
    q0 = oracle_limit(cases.riemann4(x, y, seed=5))
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    qa, qb, res = gm.field(4, q0), gm.field(4), gm.field(4)
    qc, qd, res2 = gm.field(4, q0), gm.field(4), gm.field(4)
    dt = 2e-4
    for stage in range(5):
        gm.lserk4_stage(d, qa, qb, res, stage, dt)
        gm.limit(qb, M=50.0, flags=False, positivity=True)
        qa, qb = qb, qa
        gm.lserk4_stage(d, qc, qd, res2, stage, dt)
        qd.upload(oracle_limit(qd.download()))
        qc, qd = qd, qc
        a, b = qa.download(), qc.download()
        assert np.isfinite(b).all()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("kind,prec", [("tvb", "fp64"), ("mlp", "fp64"), ("tvb", "bf16x3"), ("mlp", "bf16x3"),
                                       ("all", "bf16x3"), ("tvb", "fp32")])
def test_run_limited_matches_stage_calls(ctx, kind, prec):
    """dgh_lserk4_run_limited (the limiter after every stage inside the library's step loop) gives bitwise the
    state of the same steps taken as dgh_lserk4_stage + dgh_limit calls; order 6, where the learned indicator
    reads 64 inputs (an untrained network: its decisions only have to be the same in both sequences).  With bf16x3
    stages the loop takes the limiter's moments from the stage kernel instead of the moments pass."""
    N = 6
    vxy, vid = split_quads(16, 4, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, 2, np.int32)
    gm = ctx.mesh(N, vxy, vid, bct)
    x, y = gm.nodes()
    net = ctx.mlp(N, seed=11) if kind == "mlp" else None
    q = gm.field(4, cases.riemann4(x, y, seed=3))
    gm.limit(q, M=50.0, flags=False, positivity=True)
    q0 = q.download()
    pc = {"fp64": dgh.DGH_PREC_FP64, "fp32": dgh.DGH_PREC_FP32, "bf16x3": dgh.DGH_PREC_BF16X3}[prec]
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, pc, 0, 0.0, 0.0, 1.4)
    dt, nsteps = 1e-4, 3
    M = 1.0   # M h^2 small against the Riemann jumps on this coarse mesh: the TVB test flags cells
    lk = dgh.DGH_LIM_ALL if kind == "all" else dgh.DGH_LIM_TVB
    ld = gm.limiter_desc(M=M, mlp=net, positivity=True, kind=lk)
    fa = gm.field(4, q0)
    gm.lserk4_run(d, fa, nsteps, dt, limiter=ld)
    qa, qb, res = gm.field(4, q0), gm.field(4), gm.field(4)
    ntr = 0
    for _ in range(nsteps):
        for s in range(5):
            gm.lserk4_stage(d, qa, qb, res, s, dt)
            ntr += gm.limit(qb, M=M, flags=False, mlp=net, positivity=True, kind=lk)[1]
            qa, qb = qb, qa
    a, b = fa.download(), qa.download()
    print("%s %s: %d troubled cells over %d stages, max |run - stages| = %.3e" % (kind, prec, ntr, 5 * nsteps,
                                                                              np.abs(a - b).max()))
    assert ntr > 0
    assert np.array_equal(a, b)


def test_run_limited_rejects_bad_limiters(ctx):
    """dgh_lserk4_run_limited checks its limiter on the host: a detect-only descriptor is refused before any
    stage runs, and an unknown kind is reported at the first stage with q unchanged."""
    vxy, vid = split_quads(8, 4, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    gm = ctx.mesh(3, vxy, vid, np.full(vid.shape, 2, np.int32))
    x, y = gm.nodes()
    q0 = cases.riemann4(x, y, seed=4)
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    f = gm.field(4, q0)
    bad = gm.limiter_desc(M=50.0)
    bad.detect_only = 1
    with pytest.raises(dgh.DghError, match="detect-only"):
        gm.lserk4_run(d, f, 2, 1e-4, limiter=bad)
    bad = gm.limiter_desc(M=50.0)
    bad.kind = 99
    with pytest.raises(dgh.DghError, match="kind"):
        gm.lserk4_run(d, f, 2, 1e-4, limiter=bad)
    np.testing.assert_array_equal(f.download(), q0)
    gm.lserk4_run(d, f, 2, 1e-4, limiter=gm.limiter_desc(M=50.0, positivity=True))   # the context still works
    assert np.all(np.isfinite(f.download()))
