"""GPU parity of the LDG advection-diffusion ensembles (dgh_ensemble_*) against the oracle (oracle/ldg.c):
right-hand side and auxiliary q of every case to 1e-12, LSERK4 steps to 1e-11, the captured HIP graph
replayed several times against the eager loop (bit-identical), the autotuner's choices, and conservation
over an uneven ensemble of orders 1-8."""
import numpy as np
import pytest

from meshgen import split_quads
from oracle_lib import OracleMesh

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")

BETA = (1.0, 1.0 / np.pi)
TAU = 1.0


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def make_cases(spec, cfl=0.2):
    cases = []
    for (N, nx, ny, L, a, nu) in spec:
        h = L / max(nx, ny)
        dt = cfl * min(h / (np.hypot(*a) * (N + 1) ** 2), h * h / (max(nu, 1e-12) * (N + 1) ** 4))
        cases.append(dict(N=N, nx=nx, ny=ny, x0=-0.3, y0=0.1, Lx=L, Ly=L * ny / nx, ax=a[0], ay=a[1], nu=nu, dt=dt))
    return cases


def oracle_mesh(c):
    vxy, vid = split_quads(c["nx"], c["ny"], c["x0"], c["x0"] + c["Lx"], c["y0"], c["y0"] + c["Ly"],
                           periodic=(True, True))
    return OracleMesh(c["N"], vxy, vid)


def initial(om, c, seed):
    x, y = om.nodes()
    rng = np.random.default_rng(seed)
    kx, ky = rng.integers(1, 3, 2)
    px, py = rng.uniform(0, 2 * np.pi, 2)
    return (np.sin(2 * np.pi * kx * (x - c["x0"]) / c["Lx"] + px) * np.cos(2 * np.pi * ky * (y - c["y0"]) / c["Ly"] + py)
            + 0.5)


SPEC = [(1, 7, 5, 1.0, (1.0, 0.5), 0.01), (2, 6, 6, 2.0, (-0.7, 1.2), 0.005), (3, 9, 4, 1.0, (0.3, -1.1), 0.0),
        (4, 5, 5, 1.5, (1.5, 0.2), 0.002), (5, 4, 6, 1.0, (-1.0, -1.0), 0.01), (6, 4, 4, 1.0, (0.6, 0.9), 0.003),
        (7, 3, 5, 1.0, (1.1, -0.4), 0.008), (8, 3, 3, 1.0, (-0.5, 0.5), 0.001), (4, 20, 13, 3.0, (0.8, 0.8), 0.004)]


@pytest.mark.parametrize("autotune", [False, True])
def test_rhs_and_q_match_oracle(ctx, autotune):
    cases = make_cases(SPEC)
    ens = ctx.ensemble(cases, tau=TAU, beta=BETA, autotune=autotune)
    oms, us = [], []
    for i, c in enumerate(cases):
        om = oracle_mesh(c)
        u = initial(om, c, i)
        ens.set(i, u)
        oms.append(om)
        us.append(u)
    ens.rhs()
    info = ens.info()
    print("split", info["split"], "items", info["nitems"])
    for i, c in enumerate(cases):
        want, q = oms[i].ldg_rhs(us[i], c["ax"], c["ay"], c["nu"], tau=TAU, beta=BETA, with_q=True)
        got = ens.get(i, 1)
        gq = ens.get(i, 2)
        e_r = np.abs(got - want).max() / np.abs(want).max()
        e_q = np.abs(gq - q).max() / np.abs(q).max()
        print("case %d N=%d K=%d: rhs %.2e q %.2e" % (i, c["N"], 2 * c["nx"] * c["ny"], e_r, e_q))
        assert e_r <= 1e-12 and e_q <= 1e-12
        np.testing.assert_array_equal(ens.get(i, 0), us[i])
    ens.close()


def test_lserk4_steps_match_oracle(ctx):
    cases = make_cases(SPEC)
    ens = ctx.ensemble(cases, tau=TAU, beta=BETA, autotune=False)
    oms, us = [], []
    for i, c in enumerate(cases):
        om = oracle_mesh(c)
        u = initial(om, c, 10 + i)
        ens.set(i, u)
        oms.append(om)
        us.append(u)
    steps = 20
    ens.run(steps, graph=False)
    for i, c in enumerate(cases):
        want = oms[i].ldg_lserk4(us[i], steps, c["dt"], c["ax"], c["ay"], c["nu"], tau=TAU, beta=BETA)
        got = ens.get(i)
        err = np.abs(got - want).max() / np.abs(want).max()
        print("case %d N=%d: %d steps max rel err %.2e" % (i, c["N"], steps, err))
        assert err <= 1e-11
    ens.close()


def test_graph_replays_match_eager(ctx):
    cases = make_cases(SPEC[:6])
    a = ctx.ensemble(cases, tau=TAU, beta=BETA, autotune=False)
    b = ctx.ensemble(cases, tau=TAU, beta=BETA, autotune=False)
    for i, c in enumerate(cases):
        u = initial(oracle_mesh(c), c, 20 + i)
        a.set(i, u)
        b.set(i, u)
    for rep in range(3):
        a.run(7, graph=True)          # the same graph replayed (7 steps = 35 stages: the buffer flips each time)
        b.run(7, graph=False)
        for i in range(len(cases)):
            np.testing.assert_array_equal(a.get(i), b.get(i))
        a.rhs()                       # eager work between the replays
    a.close()
    b.close()


def test_autotuner_choices(ctx):
    cases = make_cases([(N, 12, 12, 1.0, (1.0, 0.3), 0.005) for N in range(1, 9)])
    ens = ctx.ensemble(cases, autotune=True)
    info = ens.info()
    print("split", info["split"][1:], "\n", np.round(info["tune_us"][1:], 1))
    for N in range(1, 9):
        P = int(info["split"][N])
        assert P in (1, 2, 4, 8, 16)
        Np = (N + 1) * (N + 2) // 2
        if P != 16:
            assert -(-Np // P) <= 8
        measured = info["tune_us"][N]
        k = [1, 2, 4, 8, 16].index(P)
        assert measured[k] > 0
        assert measured[k] == min(t for t in measured if t > 0)
        if N >= 2:
            assert measured[4] > 0          # the MFMA tiles were measured
    ens.close()


def test_uneven_ensemble_conserves(ctx):
    rng = np.random.default_rng(7)
    spec = []
    for _ in range(48):
        N = int(rng.integers(1, 9))
        n = int(rng.integers(3, 24))
        th = rng.uniform(0, 2 * np.pi)
        sp = rng.uniform(0.5, 2.0)
        spec.append((N, n, n, 1.0, (sp * np.cos(th), sp * np.sin(th)), rng.uniform(1e-3, 1e-2)))
    cases = make_cases(spec)
    ens = ctx.ensemble(cases, autotune=True)
    from oracle_lib import refelem_ops
    totals, energy = [], []
    for i, c in enumerate(cases):
        om = oracle_mesh(c)
        u = initial(om, c, 100 + i)
        ens.set(i, u)
        M = refelem_ops(c["N"])["M"]
        J = om.geo()[:, 4]
        totals.append((J, M, np.sum(J[:, None] * (u @ M))))
        energy.append(np.sum(J[:, None] * (u @ M) * u))
    ens.run(30, graph=True)
    for i, c in enumerate(cases):
        J, M, t0 = totals[i]
        u = ens.get(i)
        assert np.isfinite(u).all()
        assert abs(np.sum(J[:, None] * (u @ M)) - t0) <= 1e-12 * max(1.0, abs(t0))
        assert np.sum(J[:, None] * (u @ M) * u) < energy[i]
    ens.close()
