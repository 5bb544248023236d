"""Pins of the oracle TVB minmod limiter (CPU only).

The limiter is not in the reference (HopeFOAM's Trianglelimite.C is a different scheme), so it is
pinned by the mathematics of the published method (Cockburn & Shu 1998, triangle limiter):
  * the mean and edge-midpoint operators against an independent least-squares P1 fit on the cubature
    rule of the reference element;
  * the neighbour-pair coefficients reproduce m_i - b_0 with non-negative weights;
  * invariants: cell means preserved, globally linear data untouched, nothing limited under the
    M h^2 bound, only cells at a discontinuity flagged, limited cells are linear (P1), the fixedValue
    ghost uses the edge average of the boundary data.
"""
import numpy as np
import pytest

from meshgen import split_quads
from oracle_lib import OracleMesh, refelem_ops, tvb_ops


@pytest.mark.parametrize("N", [1, 2, 3, 6, 8])
def test_mean_and_midpoint_operators(N):
    ops = refelem_ops(N)
    pmean, pmid, phi = tvb_ops(N)
    r, s = ops["r"], ops["s"]
    rng = np.random.default_rng(N)
    u = rng.standard_normal(r.size)
    # independent: the cubature rule (exact to degree 3(N+1)) gives the mean and the L2 projection on P1
    uq = ops["Vq"] @ u
    w = ops["cw"]
    Vq = ops["Vq"]
    # cubature point coordinates from Vq applied to the nodal coordinates (exact: r, s are degree 1)
    rq, sq = Vq @ r, Vq @ s
    assert pmean @ u == pytest.approx(np.sum(w * uq) / np.sum(w), rel=1e-12, abs=1e-13)
    B = np.stack([np.ones_like(rq), rq, sq], axis=1)
    coef = np.linalg.solve(B.T @ (w[:, None] * B), B.T @ (w * uq))
    mids = np.array([[0.0, -1.0], [0.0, 0.0], [-1.0, 0.0]])
    want = coef[0] + coef[1] * mids[:, 0] + coef[2] * mids[:, 1]
    np.testing.assert_allclose(pmid @ u, want, rtol=1e-11, atol=1e-12)
    # P1 edge functions: phi_i(m_k) = delta_ik, read through the exact P1 interpolation at the nodes
    for i in range(3):
        fit = np.linalg.lstsq(np.stack([np.ones_like(r), r, s], axis=1), phi[:, i], rcond=None)[0]
        vals = fit[0] + fit[1] * mids[:, 0] + fit[2] * mids[:, 1]
        np.testing.assert_allclose(vals, np.eye(3)[i], atol=1e-13)


def test_neighbour_coefficients():
    vxy, vid = split_quads(5, 4, 0.0, 1.0, 0.0, 0.5, periodic=(True, False))
    m = OracleMesh(2, vxy, vid, np.full(vid.shape, 2, np.int32))
    coef, jsel, h = m.tvb_geometry()
    E, F = m.conn()
    bc = vxy.mean(axis=1)
    for e in range(m.K):
        for f in range(3):
            mi = 0.5 * (vxy[e, f] + vxy[e, (f + 1) % 3])

            def nb_centroid(k):
                en = E[e, k]
                mk = 0.5 * (vxy[e, k] + vxy[e, (k + 1) % 3])
                if en < 0:
                    return 2 * mk - bc[e]
                fn = F[e, k]
                mn = 0.5 * (vxy[en, fn] + vxy[en, (fn + 1) % 3])
                return mk + bc[en] - mn

            a1, a2 = coef[e, f]
            j = jsel[e, f]
            assert j != f and a1 >= -1e-12 and a2 >= -1e-12
            got = a1 * (nb_centroid(f) - bc[e]) + a2 * (nb_centroid(j) - bc[e])
            np.testing.assert_allclose(got, mi - bc[e], atol=1e-13)
    assert np.allclose(h, np.hypot(0.2, 0.125))


def step_mesh(N, n=8):
    vxy, vid = split_quads(n, n, 0.0, 1.0, 0.0, 1.0, periodic=(True, True))
    m = OracleMesh(N, vxy, vid)
    x, y = m.nodes()
    return m, x, y


@pytest.mark.parametrize("N", [1, 3, 5])
def test_means_preserved_and_limited_cells_linear(N):
    m, x, y = step_mesh(N)
    rng = np.random.default_rng(7)
    q = np.where(x + 0.3 * y < 0.55, 1.0, 0.125) + 0.01 * rng.standard_normal(x.shape)
    pmean, pmid, phi = tvb_ops(N)
    ql, tr = m.tvb_limit(q, M=0.0)
    np.testing.assert_allclose(ql @ pmean, q @ pmean, rtol=0, atol=1e-14)
    assert tr.any() and not tr.all()
    # limited cells hold a P1 function: nodal values reproduced by their own P1 reconstruction
    dev = ql[tr] @ pmid.T - (ql[tr] @ pmean)[:, None]
    rec = (ql[tr] @ pmean)[:, None] + dev @ phi.T
    np.testing.assert_allclose(ql[tr], rec, atol=1e-13)
    np.testing.assert_array_equal(ql[~tr], q[~tr])


def test_linear_data_untouched():
    vxy, vid = split_quads(8, 8, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    m = OracleMesh(3, vxy, vid, np.full(vid.shape, 2, np.int32))
    x, y = m.nodes()
    q = 0.3 + 2.0 * x - 1.5 * y
    ql, tr = m.tvb_limit(q, M=0.0)
    E, _ = m.conn()
    interior = (E >= 0).all(axis=1)
    assert not tr[interior].any()
    np.testing.assert_array_equal(ql[interior], q[interior])


def test_nothing_limited_under_bound():
    m, x, y = step_mesh(4)
    q = 1e-3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    ql, tr = m.tvb_limit(q, M=1e4)     # M h^2 = 1e4 * (sqrt(2)/8)^2 >> 1e-3
    assert not tr.any()
    np.testing.assert_array_equal(ql, q)
    ql, tr = m.tvb_limit(q, M=0.0)     # plain minmod flags the extrema of the wave
    assert tr.any()


def test_step_flags_only_cells_at_the_discontinuity():
    m, x, y = step_mesh(2, n=16)
    line = x + 0.3 * y - 0.53                     # not aligned with the mesh
    q = np.where(line < 0, 2.0, 1.0)
    _, tr = m.tvb_limit(q, M=0.0)
    cut = (line.min(axis=1) < 0) & (line.max(axis=1) > 0)
    # a cell is flagged only at the discontinuity: it is cut, or a neighbour of a cut cell
    E, _ = m.conn()
    near = cut | cut[E].any(axis=1)
    assert tr.any() and not tr[~near].any()
    # mesh-aligned piecewise constants are not touched: every cell is constant
    q2 = np.where(x.mean(axis=1, keepdims=True) < 0.5, 2.0, 1.0) * np.ones_like(x)
    q2l, tr2 = m.tvb_limit(q2, M=0.0)
    assert not tr2.any() and np.array_equal(q2l, q2)


def test_systems_flag_per_cell_and_ghosts():
    vxy, vid = split_quads(6, 6, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    bct = np.ones(vid.shape, np.int32)
    m = OracleMesh(2, vxy, vid, bct)
    x, y = m.nodes()
    q = np.stack([np.ones_like(x), np.where(x + 0.2 * y < 0.5, 1.0, 0.0), np.zeros_like(x), 2.5 * np.ones_like(x)])
    ql, tr = m.tvb_limit(q, M=0.0)
    assert tr.any()
    # every variable of a troubled cell is replaced by its P1 reconstruction (constants stay constant)
    np.testing.assert_allclose(ql[0], 1.0, atol=1e-14)
    np.testing.assert_allclose(ql[3], 2.5, atol=1e-14)
    # fixedValue ghost mean = the edge average of the boundary data, placed at the mirrored barycentre:
    # data equal to the linear extension of a linear field (2 u(m) - ubar) -> nothing limited anywhere;
    # the same data shifted by 5 -> only boundary cells are limited.  M h^2 ~ 5e-8 lets through the
    # round-off of deviations that vanish exactly (gradient normal to m_i - b_0), which plain minmod
    # (M = 0) would flag on sign noise.
    pmean, pmid, _ = tvb_ops(2)
    lin = 1.0 + 0.5 * x - 0.25 * y
    q2 = np.stack([lin, np.ones_like(x), np.ones_like(x), np.ones_like(x)])
    ghost = 2 * (lin @ pmid.T) - (lin @ pmean)[:, None]          # [K, 3]
    bcq = np.ones((4, m.K, 3, m.Nfp))
    bcq[0] = ghost[:, :, None]
    _, tr2 = m.tvb_limit(q2, M=1e-6, bcq=bcq)
    assert not tr2.any()
    bcq[0] += 5.0
    _, tr3 = m.tvb_limit(q2, M=1e-6, bcq=bcq)
    E, _ = m.conn()
    assert tr3.any() and not tr3[(E >= 0).all(axis=1)].any()
