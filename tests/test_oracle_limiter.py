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


# ---- closed-form pins of the minmod factor nu, the h of the M h^2 bound, and the pos/neg rescaling ----------
def _steepened_cell(N, s, nx=8):
    """Globally linear data u = 1 + 0.3 x + 0.7 y on a structured split-quad mesh (dx = dy, so the edges of
    every triangle are h0, h0 and sqrt(2) h0), with the slope of one interior cell e0 multiplied by s around its
    mean.  Its neighbours' means are the linear data at their centroids, so the neighbour difference of edge i
    is D_i = grad u . (m_i - b_0) exactly (the pair weights reproduce m_i - b_0), and the cell's own deviations
    are s D_i."""
    vxy, vid = split_quads(nx, nx, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    m = OracleMesh(N, vxy, vid, np.full(vid.shape, 2, np.int32))
    x, y = m.nodes()
    lin = 1.0 + 0.3 * x + 0.7 * y
    e0 = 2 * ((nx // 2) * nx + nx // 2)          # an interior cell
    b0 = vxy[e0].mean(axis=0)
    ubar = 1.0 + 0.3 * b0[0] + 0.7 * b0[1]
    q = lin.copy()
    q[e0] = ubar + s * (lin[e0] - ubar)
    mids = 0.5 * (vxy[e0] + np.roll(vxy[e0], -1, axis=0))      # m_i: edge i joins vertices i, i+1
    D = 0.3 * (mids[:, 0] - b0[0]) + 0.7 * (mids[:, 1] - b0[1])
    edges = np.linalg.norm(np.roll(vxy[e0], -1, axis=0) - vxy[e0], axis=1)
    return m, x, y, q, e0, b0, ubar, D, edges


@pytest.mark.parametrize("N", [1, 3])
def test_nu_closed_form(N):
    """Cockburn-Shu's nu = 1.5: a cell steepened by s = 1.3 < nu keeps its slope (minmod picks its own
    deviation), by s = 2 it is limited to exactly nu = 1.5 times the linear slope (every edge clipped to
    nu D_i, sum nu D_i = 0 so no rescaling).  A limiter that ignored nu (slope 1) or applied it to the own
    deviation fails."""
    for s, want in ((1.3, 1.3), (2.0, 1.5)):
        m, x, y, q, e0, b0, ubar, D, _ = _steepened_cell(N, s)
        ql, tr = m.tvb_limit(q, M=0.0, nu=1.5)
        exp = ubar + want * (0.3 * (x[e0] - b0[0]) + 0.7 * (y[e0] - b0[1]))
        np.testing.assert_allclose(ql[e0], exp, rtol=0, atol=1e-13)
        assert tr[e0] == (want != s)
    # and nu = 1.0 (plain minmod) limits the s = 1.3 cell to the linear slope
    m, x, y, q, e0, b0, ubar, D, _ = _steepened_cell(N, 1.3)
    ql, tr = m.tvb_limit(q, M=0.0, nu=1.0)
    np.testing.assert_allclose(ql[e0], ubar + 0.3 * (x[e0] - b0[0]) + 0.7 * (y[e0] - b0[1]), atol=1e-13)


@pytest.mark.parametrize("N", [1, 2])
def test_tvb_bound_uses_longest_edge(N):
    """The M h^2 bound takes h = the cell's longest edge (sqrt(2) h0 on these meshes).  With the cell's largest
    deviation dmax placed between M h0^2 and M (sqrt 2 h0)^2 the cell stays untouched (with the shortest edge it
    would be limited); the bound is inclusive (M = dmax / h^2 leaves it alone) and just below it the largest
    edge is clipped (to round-off of M h^2: a relative 1e-12 above and 1e-9 below)."""
    s = 2.0
    m, x, y, q, e0, b0, ubar, D, edges = _steepened_cell(N, s)
    hs, hl = edges.min(), edges.max()
    assert hl == pytest.approx(np.sqrt(2) * hs)
    dmax = np.abs(s * D).max()
    for M in (dmax / (1.5 * hs * hs), dmax / (hl * hl) * (1 + 1e-12)):
        assert M * hs * hs < dmax <= M * hl * hl
        ql, tr = m.tvb_limit(q, M=M, nu=1.5)
        assert not tr[e0]
        np.testing.assert_array_equal(ql[e0], q[e0])
    ql, tr = m.tvb_limit(q, M=dmax / (hl * hl) * (1 - 1e-9), nu=1.5)
    assert tr[e0]


def test_pos_neg_rescaling_properties():
    """When minmod clips one edge but not the others, sum d != 0 and the deviations are rescaled
    (Cockburn-Shu): the limited cell keeps its mean, each limited deviation has the sign of d_i and no larger
    magnitude, and the side (positive or negative deviations) with the smaller total is kept unchanged while
    the other is scaled to balance it.  Here d = (nu D_k on the clipped edge k, s D_j on the others)."""
    s = 2.0
    m, x, y, q, e0, b0, ubar, D, edges = _steepened_cell(2, s)
    hl = edges.max()
    a = s * D
    k = int(np.argmax(np.abs(a)))
    M = np.abs(a).max() / (hl * hl) * (1 - 1e-9)
    assert np.sort(np.abs(a))[-2] <= M * hl * hl      # only edge k exceeds the bound
    ql, tr = m.tvb_limit(q, M=M, nu=1.5)
    assert tr[e0]
    d = a.copy()
    d[k] = 1.5 * D[k]
    assert abs(d.sum()) > 1e-3 * np.abs(d).max()
    # limited deviations at the edge midpoints, read from the limited cell's P1 values
    pmean, pmid, _ = tvb_ops(2)
    dh = ql[e0] @ pmid.T - ql[e0] @ pmean
    assert ql[e0] @ pmean == pytest.approx(q[e0] @ pmean, abs=1e-14)
    assert abs(dh.sum()) < 1e-13
    assert np.all(dh * d >= -1e-15) and np.all(np.abs(dh) <= np.abs(d) + 1e-14)
    pos, neg = d[d > 0].sum(), -d[d < 0].sum()
    keep = d < 0 if pos > neg else d > 0
    np.testing.assert_allclose(dh[keep], d[keep], atol=1e-14)
    assert np.all(np.abs(dh[~keep]) < np.abs(d[~keep]))


# ---- closed-form pins of the Zhang-Shu scaling: theta is the largest admissible value --------------------------
def _linear_cell_state(m, e0, vert):
    """nodal values of the P1 data with the given vertex values vert[4][3] in element e0 (the reference nodes
    include the vertices, so the minimum over every evaluation point of linear data sits at a vertex)"""
    ops = refelem_ops(m.N)
    r, s = ops["r"], ops["s"]
    lam = np.stack([-(r + s) / 2, (r + 1) / 2, (s + 1) / 2])        # barycentric coordinates of the nodes
    return np.asarray(vert) @ lam


@pytest.mark.parametrize("N", [1, 3, 6])
def test_positivity_density_theta_closed_form(N):
    """rho linear with vertex values (-0.1, 1.5, 1.6): mean 1, minimum -0.1 at a vertex; m = 0, E = 10
    (pressure 4).  With eps = 0.2: theta1 = (1 - 0.2) / (1 - (-0.1)) exactly, so the scaled minimum equals eps
    (the largest admissible theta); a smaller theta would leave it above eps, theta = 0 flattens the cell."""
    vxy, vid = split_quads(3, 3, 0.0, 1.0, 0.0, 1.0)
    m = OracleMesh(N, vxy, vid)
    x, _ = m.nodes()
    q = np.stack([np.ones_like(x), np.zeros_like(x), np.zeros_like(x), np.full_like(x, 10.0)])
    rv = np.array([-0.1, 1.5, 1.6])
    q[0, 0] = _linear_cell_state(m, 0, rv[None])[0]
    ql, sc = m.positivity(q, gamma=1.4, eps=0.2)
    assert sc[0] and not sc[1:].any()
    th = (1.0 - 0.2) / (1.0 - (-0.1))
    np.testing.assert_allclose(ql[0, 0], 1.0 + th * (q[0, 0] - 1.0), atol=1e-14)
    assert ql[0, 0].min() == pytest.approx(0.2, abs=1e-14)
    np.testing.assert_array_equal(ql[1:, 0], q[1:, 0])


@pytest.mark.parametrize("N", [1, 4])
def test_positivity_pressure_theta_closed_form(N):
    """rho = 1 and E = 2.5 constant, momentum linear: mean (0.3, 0), vertex offsets (2.5, 0.5), (-1, 1),
    (-1.5, -1.5).  p(t) at a vertex along Ubar + t (U_v - Ubar) is (gamma - 1)(E - |m + t dm|^2 / 2), so
    p = eps has the root of |dm|^2 t^2 + 2 (m.dm) t + |m|^2 - R^2 = 0, R^2 = 2 (E - eps / (gamma - 1)); theta2
    is the smallest such root over the vertices (the admissible set is convex, so the vertices decide), and
    the scaled state is Ubar + theta2 (U - Ubar) with the violating vertex's pressure at eps."""
    g, eps, E = 1.4, 0.1, 2.5
    vxy, vid = split_quads(3, 3, 0.0, 1.0, 0.0, 1.0)
    m = OracleMesh(N, vxy, vid)
    x, _ = m.nodes()
    q = np.stack([np.ones_like(x), np.zeros_like(x), np.zeros_like(x), np.full_like(x, E)])
    mbar = np.array([0.3, 0.0])
    dm = np.array([[2.5, 0.5], [-1.0, 1.0], [-1.5, -1.5]])
    q[1, 0] = _linear_cell_state(m, 0, (mbar[0] + dm[:, 0])[None])[0]
    q[2, 0] = _linear_cell_state(m, 0, (mbar[1] + dm[:, 1])[None])[0]
    R2 = 2.0 * (E - eps / (g - 1.0))
    roots = []
    for v in range(3):
        a, b, c = dm[v] @ dm[v], 2.0 * (mbar @ dm[v]), mbar @ mbar - R2
        if c + b + a > 0:                                   # the vertex violates p >= eps
            roots.append((-b + np.sqrt(b * b - 4 * a * c)) / (2 * a))
    assert len(roots) == 1
    th = min(roots)
    ql, sc = m.positivity(q, gamma=g, eps=eps)
    assert sc[0] and not sc[1:].any()
    np.testing.assert_allclose(ql[1, 0], mbar[0] + th * (q[1, 0] - mbar[0]), atol=1e-13)
    np.testing.assert_allclose(ql[2, 0], mbar[1] + th * (q[2, 0] - mbar[1]), atol=1e-13)
    np.testing.assert_array_equal(ql[0, 0], q[0, 0])
    pv = (g - 1.0) * (ql[3, 0] - 0.5 * (ql[1, 0] ** 2 + ql[2, 0] ** 2) / ql[0, 0])
    assert pv.min() == pytest.approx(eps, abs=1e-12)
