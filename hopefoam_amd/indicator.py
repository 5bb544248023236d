"""Learned troubled-cell indicator workflow, all on the GPU.

Data: seeded four-quadrant 2D Riemann runs (Euler, Roe flux, LSERK4 with the TVB minmod limiter and the
positivity scaling after every stage) on the split-quad mesh of [0,2] x [0,0.5] with zero-gradient boundaries.  At snapshot steps
the MLP inputs of every cell (dgh_mlp_features) and the TVB troubled-cell flags (dgh_limit, detect only)
are taken from each stage's state before it is limited; all troubled cells and a fraction of the others
are kept.

Training: dgh_mlp_train (FP32 forward/backward kernels, Adam).  Evaluation on held-out seeds: the
false-negative rate and false-positive rate of the network against the TVB flags on the TVB-limited
trajectory, the fraction of cells each indicator limits, and the relative L2 deviation of the
MLP-limited solution from the TVB-limited one.
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


def collect(ctx, seeds, N=3, nx=128, ny=32, steps=40, every=8, M=50.0, keep_clean=0.05, rng_seed=0):
    """TVB-labelled samples (X [n][2 Np + 6] float32, y [n] float32) from the given seeds."""
    rng = np.random.default_rng(rng_seed)
    mesh = riemann_mesh(ctx, N, nx, ny)
    Xs, ys = [], []
    def sample(f, stage):
        X = mesh.mlp_features(f)
        flags, _ = mesh.limit(f, M=M, detect_only=True)
        keep = flags | (rng.random(flags.size) < keep_clean)
        Xs.append(X[keep])
        ys.append(flags[keep].astype(np.float32))

    for seed in seeds:
        run = RiemannRun(mesh, seed, tvb_limiter(mesh, M), N, nx, ny)
        for st in range(steps):
            run.step(sample if st % every == 0 else None)
    return np.concatenate(Xs), np.concatenate(ys)


def train(ctx, X, y, N=3, epochs=8, batch=1024, lr=2e-3, pos_weight=1.0, seed=0):
    net = ctx.mlp(N, seed=seed)
    losses = net.train(X, y, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
    return net, losses


def evaluate(ctx, net, seeds, N=3, nx=128, ny=32, steps=40, M=50.0, threshold=0.5):
    """Indicator quality on the TVB-limited trajectories and the deviation of the MLP-limited runs."""
    mesh = riemann_mesh(ctx, N, nx, ny)
    cnt = {"fn": 0, "fp": 0, "pos": 0, "neg": 0}
    lim_tvb = lim_mlp = cells = 0
    devs = []

    def judge(f, stage):
        ft, _ = mesh.limit(f, M=M, detect_only=True)
        fm, _ = mesh.limit(f, M=M, detect_only=True, mlp=net, threshold=threshold)
        cnt["fn"] += int(np.sum(ft & ~fm))
        cnt["fp"] += int(np.sum(~ft & fm))
        cnt["pos"] += int(ft.sum())
        cnt["neg"] += int((~ft).sum())

    for seed in seeds:
        a = RiemannRun(mesh, seed, tvb_limiter(mesh, M), N, nx, ny)
        b = RiemannRun(mesh, seed, mlp_limiter(mesh, M, net, threshold), N, nx, ny)
        for _ in range(steps):
            a.step(judge)
            b.step()
        lim_tvb += a.limited
        lim_mlp += b.limited
        cells += a.checked
        qa, qb = a.q.download(), b.q.download()
        devs.append(float(np.linalg.norm(qb[0] - qa[0]) / np.linalg.norm(qa[0])))
        if not np.all(np.isfinite(qb)):
            devs[-1] = math.inf
    return {
        "false_negative_rate": cnt["fn"] / max(cnt["pos"], 1),
        "false_positive_rate": cnt["fp"] / max(cnt["neg"], 1),
        "limited_fraction_tvb": lim_tvb / max(cells, 1),
        "limited_fraction_mlp": lim_mlp / max(cells, 1),
        "l2_rho_deviation_mlp_vs_tvb": float(np.max(devs)),
        "l2_rho_deviation_mean": float(np.mean(devs)),
    }
