"""Learned troubled-cell indicator workflow, all on the GPU.

Data: seeded four-quadrant 2D Riemann runs (Euler, Roe flux, LSERK4 with the TVB minmod limiter and the
positivity scaling after every stage) on the split-quad mesh of [0,2] x [0,0.5] with zero-gradient
boundaries.  At snapshot steps the MLP inputs of every cell (dgh_mlp_features) and the TVB troubled-cell
flags (dgh_limit, detect only) are taken from each stage's state before it is limited; all troubled
cells and a fraction of the others are kept.  The Riemann states are piecewise constant away from the
waves, so every untroubled cell they offer is constant; seeded smooth periodic states on coarser
meshes (labelled by the same TVB test) supply untroubled cells with smooth variation.

Training: dgh_mlp_train (FP32 forward/backward kernels, Adam).  Evaluation on held-out seeds: the
false-negative rate and false-positive rate of the network against the TVB flags on the TVB-limited
trajectory, the fraction of cells each indicator limits, and the relative L2 deviation of the
MLP-limited solution from the TVB-limited one -- next to the same solver without a detection step
(every cell limited), which is what the indicator saves.
"""
import math

import numpy as np

from . import cases, dgh

X0, X1, Y0, Y1 = 0.0, 2.0, 0.0, 0.5


def riemann_mesh(ctx, N, nx, ny):
    return ctx.mesh_structured(N, nx, ny, X0, X1, Y0, Y1, periodic=(False, False), bc_tag=dgh.DGH_BC_ZERO_GRADIENT)


def max_speed(q, gamma=cases.GAMMA):
    rho, ru, rv, E = q
    u, v = ru / rho, rv / rho
    p = (gamma - 1) * (E - 0.5 * rho * (u * u + v * v))
    return float(np.max(np.hypot(u, v) + np.sqrt(gamma * np.maximum(p, 1e-12) / rho)))


class RiemannRun:
    """One seeded run; `limit(field)` (returning the number of limited cells) is applied to the initial
    state and after every RK stage."""

    def __init__(self, mesh, seed, limit, N, nx, ny, cfl=0.3):
        self.mesh = mesh
        x, y = mesh.nodes()
        q0 = cases.riemann4(x, y, seed, X0, X1, Y0, Y1)
        self.q = mesh.field(4, q0)
        self.tmp = mesh.field(4)
        self.res = mesh.field(4)
        self.limit = limit
        self.limit(self.q)
        self.limited = 0      # cells limited after RK stages
        self.checked = 0
        hx, hy = (X1 - X0) / nx, (Y1 - Y0) / ny
        # the Riemann fans do not exceed 1.5 x the initial largest wave speed at these states
        self.dt = cases.stable_dt(N, hx, hy, 1.5 * max_speed(q0), cfl)
        self.desc = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, cases.GAMMA)

    def step(self, before_limit=None):
        """one LSERK4 step; before_limit(field, stage) sees every stage's state before it is limited"""
        for s in range(5):
            self.mesh.lserk4_stage(self.desc, self.q, self.tmp, self.res, s, self.dt)
            if before_limit is not None:
                before_limit(self.tmp, s)
            self.limited += self.limit(self.tmp)
            self.checked += self.mesh.K
            self.q, self.tmp = self.tmp, self.q


def tvb_limiter(mesh, M):
    return lambda f: mesh.limit(f, M=M, flags=False, positivity=True)[1]


def mlp_limiter(mesh, M, mlp, threshold):
    return lambda f: mesh.limit(f, M=M, flags=False, mlp=mlp, threshold=threshold, positivity=True)[1]


def all_limiter(mesh, M):
    return lambda f: mesh.limit(f, M=M, flags=False, kind=dgh.DGH_LIM_ALL, positivity=True)[1]


def run_limited(mesh, q0, dt, nsteps, limit, flux=dgh.DGH_FLUX_ROE):
    """LSERK4 Euler run from q0 with `limit` (or None) after every stage; returns the final state"""
    desc = dgh.RhsDesc(dgh.DGH_EQ_EULER, flux, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, cases.GAMMA)
    q, tmp, res = mesh.field(4, q0), mesh.field(4), mesh.field(4)
    if limit is not None:
        limit(q)
    for _ in range(nsteps):
        for s in range(5):
            mesh.lserk4_stage(desc, q, tmp, res, s, dt)
            if limit is not None:
                limit(tmp)
            q, tmp = tmp, q
    return q.download()


def vortex_errors(ctx, net, N=3, n=32, T=1.0, M=50.0, threshold=0.5):
    """L2 error against the exact isentropic vortex (periodic [0,10]^2, n x n x 2) of the unlimited
    scheme, the MLP-indicated limiter, the TVB-indicated limiter and the limiter on every cell"""
    mesh = ctx.mesh_structured(N, n, n, 0.0, 10.0, 0.0, 10.0, periodic=(True, True))
    x, y = mesh.nodes()
    q0 = cases.vortex_periodic(x, y, 0.0)
    h = 10.0 / n
    nsteps = int(math.ceil(T / cases.stable_dt(N, h, h, max_speed(q0), 0.3)))
    dt = T / nsteps
    exact = cases.vortex_periodic(x, y, T)
    out = {}
    for name, lim in (("none", None), ("mlp", mlp_limiter(mesh, M, net, threshold)), ("tvb", tvb_limiter(mesh, M)),
                      ("all", all_limiter(mesh, M))):
        q = run_limited(mesh, q0, dt, nsteps, lim)
        out["vortex_l2_rho_" + name] = float(np.sqrt(np.mean((q[0] - exact[0]) ** 2)))
    return out


def collect_smooth(ctx, seeds, N=3, M=50.0, keep_clean=1.0, rng_seed=1):
    """TVB-labelled samples from seeded smooth periodic states on [0,10]^2 with 12..32 quads a side"""
    rng = np.random.default_rng(rng_seed)
    sizes = (12, 16, 24, 32)
    meshes = {n: ctx.mesh_structured(N, n, n, 0.0, 10.0, 0.0, 10.0, periodic=(True, True)) for n in sizes}
    fields = {n: meshes[n].field(4) for n in sizes}
    nodes = {n: meshes[n].nodes() for n in sizes}
    Xs, ys = [], []
    for seed in seeds:
        n = sizes[seed % 4]
        mesh, f = meshes[n], fields[n]
        f.upload(cases.smooth_random(*nodes[n], seed))
        X = mesh.mlp_features(f, M)
        flags, _ = mesh.limit(f, M=M, detect_only=True)
        keep = flags | (rng.random(flags.size) < keep_clean)
        Xs.append(X[keep])
        ys.append(flags[keep].astype(np.float32))
    return np.concatenate(Xs), np.concatenate(ys)


def collect(ctx, seeds, N=3, nx=128, ny=32, steps=40, every=8, M=50.0, keep_clean=0.05, rng_seed=0, smooth_seeds=()):
    """TVB-labelled samples (X [n][2 Np + 6] float32, y [n] float32) from the given Riemann seeds (and smooth
    seeds)."""
    rng = np.random.default_rng(rng_seed)
    mesh = riemann_mesh(ctx, N, nx, ny)
    Xs, ys = [], []
    def sample(f, stage):
        X = mesh.mlp_features(f, M)
        flags, _ = mesh.limit(f, M=M, detect_only=True)
        keep = flags | (rng.random(flags.size) < keep_clean)
        Xs.append(X[keep])
        ys.append(flags[keep].astype(np.float32))

    for seed in seeds:
        run = RiemannRun(mesh, seed, tvb_limiter(mesh, M), N, nx, ny)
        for st in range(steps):
            run.step(sample if st % every == 0 else None)
    if len(smooth_seeds):
        Xsm, ysm = collect_smooth(ctx, smooth_seeds, N, M)
        Xs.append(Xsm)
        ys.append(ysm)
    return np.concatenate(Xs), np.concatenate(ys)


def save_net(net, path):
    """network parameters (float32) and order to an .npz file"""
    np.savez(path, N=np.int32(net.N), params=net.params())


def load_net(ctx, path):
    """a network written by save_net, on ctx's device"""
    with np.load(path, allow_pickle=False) as z:
        net = ctx.mlp(int(z["N"]))
        net.set_params(z["params"])
    return net


def train(ctx, X, y, N=3, epochs=8, batch=1024, lr=2e-3, pos_weight=1.0, seed=0):
    net = ctx.mlp(N, seed=seed)
    losses = net.train(X, y, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
    return net, losses


def evaluate(ctx, net, seeds, N=3, nx=128, ny=32, steps=40, M=50.0, threshold=0.5):
    """Indicator quality on the TVB-limited trajectories and the deviation of the MLP-limited runs."""
    mesh = riemann_mesh(ctx, N, nx, ny)
    cnt = {"fn": 0, "fp": 0, "pos": 0, "neg": 0}
    lim_tvb = lim_mlp = cells = 0
    devs, devs_all = [], []

    def compare_with_tvb(f, stage):
        ft, _ = mesh.limit(f, M=M, detect_only=True)
        fm, _ = mesh.limit(f, M=M, detect_only=True, mlp=net, threshold=threshold)
        cnt["fn"] += int(np.sum(ft & ~fm))
        cnt["fp"] += int(np.sum(~ft & fm))
        cnt["pos"] += int(ft.sum())
        cnt["neg"] += int((~ft).sum())

    for seed in seeds:
        a = RiemannRun(mesh, seed, tvb_limiter(mesh, M), N, nx, ny)
        b = RiemannRun(mesh, seed, mlp_limiter(mesh, M, net, threshold), N, nx, ny)
        c = RiemannRun(mesh, seed, all_limiter(mesh, M), N, nx, ny)
        for _ in range(steps):
            a.step(compare_with_tvb)
            b.step()
            c.step()
        lim_tvb += a.limited
        lim_mlp += b.limited
        cells += a.checked
        qa, qb, qc = a.q.download(), b.q.download(), c.q.download()
        devs.append(float(np.linalg.norm(qb[0] - qa[0]) / np.linalg.norm(qa[0])))
        devs_all.append(float(np.linalg.norm(qc[0] - qa[0]) / np.linalg.norm(qa[0])))
        if not np.all(np.isfinite(qb)):
            devs[-1] = math.inf
    return {
        "false_negative_rate": cnt["fn"] / max(cnt["pos"], 1),
        "false_positive_rate": cnt["fp"] / max(cnt["neg"], 1),
        "limited_fraction_tvb": lim_tvb / max(cells, 1),
        "limited_fraction_mlp": lim_mlp / max(cells, 1),
        "l2_rho_deviation_mlp_vs_tvb": float(np.max(devs)),
        "l2_rho_deviation_mean": float(np.mean(devs)),
        "l2_rho_deviation_limit_all_vs_tvb": float(np.max(devs_all)),
        "l2_rho_deviation_limit_all_mean": float(np.mean(devs_all)),
    }


def limiter_cost(ctx, net, N=3, nx=512, ny=128, M=50.0, threshold=0.5, reps=50, seed=1000):
    """device time per limiter call (TVB detection vs fused MLP detection, both with positivity), from
    wall-clock over reps calls between synchronisations"""
    import time
    mesh = riemann_mesh(ctx, N, nx, ny)
    x, y = mesh.nodes()
    q = mesh.field(4, cases.riemann4(x, y, seed, X0, X1, Y0, Y1))
    out = {}
    for name, kw in (("tvb", {}), ("mlp", {"mlp": net, "threshold": threshold}), ("all", {"kind": dgh.DGH_LIM_ALL})):
        f = mesh.field(4)
        f.copy_from(q)
        mesh.limit(f, M=M, flags=False, positivity=True, **kw)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            f.copy_from(q)
            mesh.limit(f, M=M, flags=False, positivity=True, **kw)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(reps):
            f.copy_from(q)
        ctx.sync()
        t2 = time.perf_counter()
        out["limiter_us_" + name] = 1e6 * ((t1 - t0) - (t2 - t1)) / reps
    return out


def main(argv=None):
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description="train and evaluate the learned troubled-cell indicator on the GPU")
    ap.add_argument("--N", type=int, default=3)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--train-seeds", type=int, default=200)
    ap.add_argument("--eval-seeds", type=int, default=20)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--keep-clean", type=float, default=0.02)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--M", type=float, default=50.0)
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--pos-weight", type=float, default=2.0)
    ap.add_argument("--smooth-seeds", type=int, default=2000)
    ap.add_argument("--save", default=None, help="write the trained network to this .npz file")
    args = ap.parse_args(argv)
    ctx = dgh.Context(0)
    t0 = time.perf_counter()
    X, y = collect(ctx, range(args.train_seeds), args.N, args.nx, args.ny, args.steps, args.every, args.M,
                   args.keep_clean, smooth_seeds=range(10000, 10000 + args.smooth_seeds))
    ok = np.all(np.isfinite(X), axis=1)
    X, y = X[ok], y[ok]
    t1 = time.perf_counter()
    print("# collected %d samples (%d troubled) in %.1f s" % (y.size, int(y.sum()), t1 - t0), flush=True)
    net, losses = train(ctx, X, y, args.N, epochs=args.epochs, pos_weight=args.pos_weight)
    t2 = time.perf_counter()
    print("# trained %d epochs in %.1f s, losses %s" % (args.epochs, t2 - t1, np.round(losses, 4).tolist()),
          flush=True)
    res = evaluate(ctx, net, range(1000, 1000 + args.eval_seeds), args.N, args.nx, args.ny, args.steps, args.M,
                   args.threshold)
    t3 = time.perf_counter()
    if args.save:
        save_net(net, args.save)
    res.update(limiter_cost(ctx, net, args.N, args.nx, args.ny, args.M, args.threshold))
    res.update(vortex_errors(ctx, net, args.N, M=args.M, threshold=args.threshold))
    res.update({"samples": int(y.size), "troubled_samples": int(y.sum()), "epoch_losses": losses.tolist(),
                "train_seeds": args.train_seeds, "smooth_seeds": args.smooth_seeds, "eval_seeds": args.eval_seeds,
                "steps": args.steps,
                "mesh": "%dx%dx2 on [0,2]x[0,0.5], N=%d" % (args.nx, args.ny, args.N), "M": args.M,
                "threshold": args.threshold, "seconds": {"collect": t1 - t0, "train": t2 - t1,
                                                         "evaluate": t3 - t2}})
    print(json.dumps(res))
    ctx.close()


if __name__ == "__main__":
    main()
