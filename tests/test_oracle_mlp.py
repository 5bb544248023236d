"""Pins of the numpy MLP oracle (oracle/mlp_ref.py), CPU only: finite-difference gradients, the bf16
rounding rule, and closed-form feature values."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

from meshgen import split_quads  # noqa: E402
from oracle_lib import OracleMesh, refelem_ops, tvb_ops  # noqa: E402


def random_params(nin, seed=0):
    rng = np.random.default_rng(seed)
    n = mlp_ref.H * nin + mlp_ref.H * mlp_ref.H + 3 * mlp_ref.H + 1
    return rng.standard_normal(n) * 0.3


@pytest.mark.parametrize("pw", [1.0, 4.0])
def test_gradient_matches_finite_differences(pw):
    nin = 26
    rng = np.random.default_rng(1)
    p = random_params(nin)
    X = rng.standard_normal((9, nin))
    y = (rng.random(9) < 0.4).astype(float)
    L, g = mlp_ref.loss_grad(p, X, y, pw)
    assert L == pytest.approx(mlp_ref.loss(p, X, y, pw), rel=1e-14)
    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [p.size - 1, nin * 64 + 3, p.size - 40]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mlp_ref.loss(p + e, X, y, pw) - mlp_ref.loss(p - e, X, y, pw)) / 2e-6
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -7 + 2 ** -9, 3.14159265], np.float32)
    got = mlp_ref.bf16(x)
    # ulp of bf16 at 1 is 2^-7: 1 + 2^-8 is a tie -> even (1); 1 + 3 2^-8 -> 1 + 2^-6; 1 + 2^-7 + 2^-9 -> 1 + 2^-7
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0 + 2 ** -7, 3.140625], np.float32)
    np.testing.assert_array_equal(got, want)


def test_features_closed_form():
    N = 3
    vxy, vid = split_quads(4, 4, 0.0, 1.0, 0.0, 1.0, periodic=(True, True))
    m = OracleMesh(N, vxy, vid)
    ops = refelem_ops(N)
    invV = ops["invV"]
    pmean, pmid, _ = tvb_ops(N)
    _, _, h = m.tvb_geometry()
    q = np.stack([np.full((m.K, m.Np), 1.3), np.zeros((m.K, m.Np)), np.zeros((m.K, m.Np)), np.full((m.K, m.Np), 2.0)])
    mean = q @ pmean
    E, _ = m.conn()
    X = mlp_ref.features(N, q, mean, mean[:, E], invV, q @ pmid.T, h, 50.0)
    want = np.zeros(2 * m.Np + 8)
    want[0] = want[m.Np] = 1.0
    want[-2:] = -10.0          # no deviation: the TVB scale is clipped at its floor
    np.testing.assert_allclose(X, np.broadcast_to(want, X.shape), atol=1e-6)
    # a linear field has only the degree-1 modes (1, N+1) besides the mean; its TVB scale is
    # log10(max |midpoint deviation| / (M h^2))
    x, y = m.nodes()
    q[0] = 2.0 + 0.3 * x - 0.2 * y
    X = mlp_ref.features(N, q, q @ pmean, (q @ pmean)[:, E], invV, q @ pmid.T, h, 50.0)
    other = [k for k in range(m.Np) if k not in (0, 1, N + 1)]
    assert np.abs(X[:, other]).max() < 1e-6 and np.abs(X[:, [1, N + 1]]).max() > 1e-3
    dev = np.abs((q[0] @ pmid.T) - (q[0] @ pmean)[:, None]).max(axis=1)
    np.testing.assert_allclose(X[:, 2 * m.Np + 6], np.log10(dev / (50.0 * h * h)), rtol=1e-6)


def test_adam_closed_forms():
    """The bias-corrected Adam step (Kingma & Ba 2015): from m = v = 0 the first step moves each parameter by
    lr g / (|g| + eps), i.e. lr in the direction -sign(g) for |g| >> eps, and 0 where g = 0; with a constant
    gradient the corrected moments stay mhat = g, vhat = g^2, so every later step is the same.  A missing bias
    correction (step 1 would be lr (1 - b1) / sqrt(1 - b2) = 3.16 lr), a wrong beta or a flipped sign fails."""
    g = np.array([0.5, -2.0, 1e-3, 0.0, -7.5], np.float32)
    p = np.zeros(5, np.float32)
    m = np.zeros(5, np.float32)
    v = np.zeros(5, np.float32)
    lr = 1e-2
    for t in range(1, 6):
        p_new, m, v = mlp_ref.adam_step(p, m, v, g, t, lr)
        step = p_new - p
        want = -lr * g / (np.abs(g) + 1e-8)
        np.testing.assert_allclose(step, want, rtol=2e-5, atol=1e-9)
        np.testing.assert_allclose(m / (1 - 0.9 ** t), g, rtol=1e-5)
        np.testing.assert_allclose(v / (1 - 0.999 ** t), g * g, rtol=2e-4)
        p = p_new
    # a gradient that flips sign after one step: m2 = 0.9 (0.1 g) - 0.1 g = -0.01 g, mhat = m2 / (1 - 0.9^2);
    # v2 = 0.999 (0.001 g^2) + 0.001 g^2, vhat = v2 / (1 - 0.999^2)
    g1 = np.array([1.0], np.float32)
    p1, m1, v1 = mlp_ref.adam_step(np.zeros(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32), g1, 1, lr)
    p2, m2, v2 = mlp_ref.adam_step(p1, m1, v1, -g1, 2, lr)
    mh = -0.01 / (1 - 0.81)
    vh = (0.999 * 0.001 + 0.001) / (1 - 0.999 ** 2)
    assert float(p2[0] - p1[0]) == pytest.approx(-lr * mh / (np.sqrt(vh) + 1e-8), rel=1e-5)
