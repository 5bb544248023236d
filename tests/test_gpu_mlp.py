"""GPU parity of the learned troubled-cell indicator against the numpy oracle (oracle/mlp_ref.py):
input features, FP32 forward, gradients of the training loss, BF16 MFMA inference, the fused MLP limiter,
and a small end-to-end training on TVB labels."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

from meshgen import split_quads  # noqa: E402
from oracle_lib import OracleMesh, refelem_ops, tvb_ops  # noqa: E402

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")
cases = pytest.importorskip("hopefoam_amd.cases")
indicator = pytest.importorskip("hopefoam_amd.indicator")


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def random_net(ctx, N=3, seed=3, scale=0.3):
    net = ctx.mlp(N, seed=seed)
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(net.nparams) * scale).astype(np.float32)
    net.set_params(p)
    return net, p


def riemann_state(N, nx=16, ny=8, seed=2):
    vxy, vid = split_quads(nx, ny, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, 2, np.int32)
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    return vxy, vid, bct, om, cases.riemann4(x, y, seed)


@pytest.mark.parametrize("N", [1, 3, 6])
def test_features_match_oracle(ctx, N):
    vxy, vid, bct, om, q = riemann_state(N)
    q = q * (1.0 + 0.05 * np.sin(3.0 * om.nodes()[0]))[None]
    gm = ctx.mesh(N, vxy, vid, bct)
    X = gm.mlp_features(gm.field(4, q), 50.0)
    pmean, pmid, _ = tvb_ops(N)
    mean = q @ pmean
    E, _ = om.conn()
    _, _, h = om.tvb_geometry()
    nb = np.where(E[None] >= 0, mean[:, np.maximum(E, 0)], mean[:, :, None])   # zero-gradient ghost = own mean
    want = mlp_ref.features(N, q, mean, nb, refelem_ops(N)["invV"], q @ pmid.T, h, 50.0)
    err = np.abs(X - want).max() / max(1.0, np.abs(want).max())
    print("N=%d features max rel err %.2e" % (N, err))
    assert X.shape == want.shape and err <= 1e-6


@pytest.mark.parametrize("N", [3, 6])
def test_forward_fp32_and_bf16(ctx, N):
    """32 inputs (N = 3, two 16-deep K steps) and 64 (N = 6, four)"""
    net, p = random_net(ctx, N)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((333, net.nin)).astype(np.float32)
    z32 = net.forward(X, dgh.DGH_MLP_FP32)
    want = mlp_ref.forward(p, X)
    e32 = np.abs(z32 - want).max() / max(1.0, np.abs(want).max())
    zb = net.forward(X, dgh.DGH_MLP_BF16)
    wantb = mlp_ref.forward_bf16(p, X)
    eb = np.abs(zb - wantb).max() / max(1.0, np.abs(wantb).max())
    print("forward fp32 rel err %.2e, bf16 vs bf16 oracle %.2e, bf16 vs fp32 %.2e" % (
        e32, eb, np.abs(zb - want).max() / np.abs(want).max()))
    assert e32 <= 1e-5
    assert eb <= 1e-2


@pytest.mark.parametrize("N,n,pw", [(3, 64, 1.0), (3, 200, 3.0), (3, 1000, 0.5), (6, 200, 2.0)])
def test_grad_matches_oracle(ctx, N, n, pw):
    net, p = random_net(ctx, N, seed=n)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, net.nin)).astype(np.float32)
    y = (rng.random(n) < 0.3).astype(np.float32)
    loss, g = net.grad(X, y, pw)
    L, gw = mlp_ref.loss_grad(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64), pw)
    err = np.abs(g - gw).max() / np.abs(gw).max()
    print("n=%d pw=%.1f: loss %.6f / %.6f, grad max rel err %.2e" % (n, pw, loss, L, err))
    assert loss == pytest.approx(L, rel=1e-5)
    assert err <= 1e-4


@pytest.mark.parametrize("N", [3, 6])
def test_fused_limiter_decisions(ctx, N):
    """The limiter's fused BF16 inference decides exactly as the standalone BF16 forward on the features;
    cells both it and the TVB test flag are rewritten exactly as the oracle's TVB limiter does."""
    vxy, vid, bct, om, q = riemann_state(N, 24, 8, seed=9)
    gm = ctx.mesh(N, vxy, vid, bct)
    net, p = random_net(ctx, N, scale=0.2)
    X = gm.mlp_features(gm.field(4, q), 50.0)
    z = net.forward(X, dgh.DGH_MLP_BF16)
    thr = float(np.median(1 / (1 + np.exp(-z.astype(np.float64)))))
    zthr = np.log(thr / (1 - thr))
    f = gm.field(4, q)
    flags, n = gm.limit(f, M=50.0, mlp=net, threshold=thr)
    clear = np.abs(z - zthr) > 1e-4
    assert np.array_equal(flags[clear], (z >= zthr)[clear]) and n == flags.sum()
    out = f.download()
    want, tvb = om.tvb_limit(q, M=50.0)
    both = flags & tvb
    assert both.any()
    np.testing.assert_allclose(out[:, both], want[:, both], rtol=0, atol=1e-13 * np.abs(q).max())
    np.testing.assert_array_equal(out[:, ~flags], q[:, ~flags])
    # detect only: flags, q untouched
    g = gm.field(4, q)
    flags2, _ = gm.limit(g, M=50.0, mlp=net, threshold=thr, detect_only=True)
    assert np.array_equal(flags2, flags)
    np.testing.assert_array_equal(g.download(), q)


def test_limited_euler_steps_with_mlp(ctx):
    """A short Riemann run limited by a trained-on-the-fly network stays finite with positive density."""
    X, y = indicator.collect(ctx, seeds=[0, 1, 2], N=3, nx=48, ny=12, steps=6, every=2, keep_clean=0.2)
    net, losses = indicator.train(ctx, X, y, N=3, epochs=6, batch=256, lr=3e-3)
    print("samples %d (troubled %d), epoch losses %s" % (y.size, int(y.sum()), np.round(losses, 4)))
    assert losses[-1] < losses[0]
    zb = net.forward(X, dgh.DGH_MLP_BF16)
    z32 = net.forward(X, dgh.DGH_MLP_FP32)
    agree = np.mean((zb >= 0) == (z32 >= 0))
    acc = np.mean((z32 >= 0) == (y > 0.5))
    print("training accuracy %.4f, bf16/fp32 decision agreement %.4f" % (acc, agree))
    assert acc >= 0.9 and agree >= 0.99
    mesh = indicator.riemann_mesh(ctx, 3, 48, 12)
    run = indicator.RiemannRun(mesh, 7, indicator.mlp_limiter(mesh, 50.0, net, 0.5), 3, 48, 12)
    for _ in range(10):
        run.step()
    q = run.q.download()
    assert np.all(np.isfinite(q)) and q[0].min() > 0


def test_indicator_gain_over_no_detection(ctx):
    """What the trained network gains.  Against the TVB indicator (the reference decision) on held-out Riemann runs it
    misses few troubled cells and limits a small fraction; on the smooth isentropic vortex, where limiting
    every cell (the same solver without a detection step) clips the extrema, the network-indicated solver
    keeps the accuracy of the unlimited scheme."""
    X, y = indicator.collect(ctx, seeds=range(8), N=3, nx=96, ny=24, steps=30, every=6, keep_clean=0.05,
                             smooth_seeds=range(10000, 10040))
    net, _ = indicator.train(ctx, X, y, N=3, epochs=8, pos_weight=2.0)
    r = indicator.evaluate(ctx, net, seeds=[1000, 1001], N=3, nx=96, ny=24, steps=30)
    r.update(indicator.vortex_errors(ctx, net, N=3, n=24, T=0.5))
    print({k: round(v, 7) for k, v in r.items()})
    assert r["false_negative_rate"] <= 0.05
    assert r["limited_fraction_mlp"] <= 0.25
    assert r["vortex_l2_rho_mlp"] <= 1.5 * r["vortex_l2_rho_none"]
    assert r["vortex_l2_rho_all"] >= 5.0 * r["vortex_l2_rho_mlp"]


WEIGHTS_N6 = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hopefoam_amd", "data",
                          "indicator_N6.npz")


def test_saved_network_roundtrip(ctx, tmp_path):
    """save_net / load_net keep the order and every parameter bit; the reloaded network computes the same."""
    net, p = random_net(ctx, N=6, seed=5)
    path = str(tmp_path / "net.npz")
    indicator.save_net(net, path)
    back = indicator.load_net(ctx, path)
    assert back.N == 6 and back.nin == net.nin
    np.testing.assert_array_equal(back.params(), p)
    X = np.random.default_rng(5).standard_normal((128, net.nin)).astype(np.float32)
    np.testing.assert_array_equal(back.forward(X, dgh.DGH_MLP_BF16), net.forward(X, dgh.DGH_MLP_BF16))


def test_committed_network_leaves_smooth_vortex_alone(ctx):
    """The order-6 network the mixed+learned benchmark loads flags no cell of the smooth isentropic vortex, so
    the limited run equals the unlimited one bit for bit, while limiting every cell changes the solution."""
    net = indicator.load_net(ctx, WEIGHTS_N6)
    assert net.N == 6
    mesh = ctx.mesh_structured(6, 16, 16, 0.0, 10.0, 0.0, 10.0, periodic=(True, True))
    x, y = mesh.nodes()
    q0 = cases.vortex_periodic(x, y, 0.0)
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_LF, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    dt, nsteps = cases.stable_dt(6, 10.0 / 16, 10.0 / 16, indicator.max_speed(q0), 0.4), 20
    runs = {}
    for name, lim in (("none", None), ("mlp", mesh.limiter_desc(M=50.0, mlp=net)),
                      ("all", mesh.limiter_desc(M=50.0, kind=dgh.DGH_LIM_ALL))):
        f = mesh.field(4, q0)
        mesh.lserk4_run(d, f, nsteps, dt, limiter=lim)
        runs[name] = f
    flags, n = mesh.limit(runs["mlp"], M=50.0, detect_only=True, mlp=net)
    a, b, c = (runs[k].download() for k in ("none", "mlp", "all"))
    print("flagged %d of %d; |all - none| = %.3e" % (n, mesh.K, np.abs(c - a).max()))
    assert n == 0
    np.testing.assert_array_equal(b, a)
    assert np.abs(c - a).max() > 1e-4


def test_committed_network_agrees_with_tvb_on_riemann_data(ctx):
    """On a held-out four-quadrant Riemann run at the training resolution (256 x 64 x 2, N = 6), the committed
    network flags the cells the TVB test flags (the labels it was trained on): few misses, few extra cells."""
    net = indicator.load_net(ctx, WEIGHTS_N6)
    N, nx, ny = 6, 256, 64
    mesh = indicator.riemann_mesh(ctx, N, nx, ny)
    cnt = {"fn": 0, "fp": 0, "pos": 0, "neg": 0}

    def compare(f, stage):
        ft, _ = mesh.limit(f, M=50.0, detect_only=True)
        fm, _ = mesh.limit(f, M=50.0, detect_only=True, mlp=net)
        cnt["fn"] += int(np.sum(ft & ~fm))
        cnt["fp"] += int(np.sum(~ft & fm))
        cnt["pos"] += int(ft.sum())
        cnt["neg"] += int((~ft).sum())

    run = indicator.RiemannRun(mesh, 1005, indicator.tvb_limiter(mesh, 50.0), N, nx, ny)
    for _ in range(6):
        run.step(compare)
    fnr, fpr = cnt["fn"] / max(cnt["pos"], 1), cnt["fp"] / max(cnt["neg"], 1)
    print("TVB-troubled %d, false negatives %.4f, false positives %.4f" % (cnt["pos"], fnr, fpr))
    assert cnt["pos"] > 100
    assert fnr <= 0.02 and fpr <= 0.03


@pytest.mark.parametrize("N,pw", [(3, 1.0), (6, 2.5)])
def test_adam_steps_match_oracle(ctx, N, pw):
    """The training kernel's Adam update (mlp_reduce_adam_kernel) against the oracle's bias-corrected Adam step
    (Kingma & Ba 2015) over 10 consecutive steps: one sample per step, so the step's gradient is exactly the one
    dgh_mlp_grad returns for the current parameters; the oracle applies it to the GPU's parameters and moments.
    A wrong beta, bias-correction t, eps or sign moves parameters by up to lr and fails."""
    net, p = random_net(ctx, N, seed=11 + N)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((1, net.nin)).astype(np.float32)
    y = np.ones(1, np.float32)
    lr = 3e-3
    m = np.zeros(net.nparams, np.float32)
    v = np.zeros(net.nparams, np.float32)
    cur = net.params()
    for t in range(1, 11):
        _, g = net.grad(X, y, pw)
        net.train(X, y, epochs=1, batch=1, lr=lr, pos_weight=pw, seed=t)
        got = net.params()
        want, m, v = mlp_ref.adam_step(cur, m, v, g, t, lr)
        err = np.abs(got - want).max()
        moved = np.abs(want - cur).max()
        assert moved > 0.5 * lr                     # the step is not trivially small
        assert err <= 1e-3 * lr + 1e-7 * np.abs(want).max(), (t, err)
        cur = got
    print("N=%d: 10 Adam steps, max |p_gpu - p_oracle| %.2e (lr %.0e)" % (N, err, lr))
