"""Pins of the oracle of the reference's Triangle limiter (Trianglelimite.C:61-852; CPU only).

The reference holds no test, fixture or published number for this limiter (its only user, the doubleMach
tutorial, publishes none), so the oracle is pinned by what the mathematics of the scheme fixes:
  * linear exactness: away from the boundary (cells whose neighbours have no boundary face) a state with
    linear density and pressure and constant velocity is reproduced exactly, and with constant density
    and linear velocity and pressure the density and momentum are exact and the energy is the closed-form
    linearisation E = Ebar + grad p.dx/(gamma-1) + rho (ubar grad u.dx + vbar grad v.dx);
  * every cell mean is preserved (the P1 part is built about the centroid);
  * a constant state at rest is left alone with every boundary type;
  * mirror symmetry: a half domain with a reflective wall equals the symmetric full domain outside the
    wall cells (pins the mirrored ghost centroid, the wall states and the boundary A_2);
  * rotation by 90 degrees commutes with the limiter (pins the x / y handling);
  * the density floor: floored cells have rho >= tol at every node, one halving step less would fall
    below tol, and the shortened du enters the momentum;
  * a second, brute-force evaluation on a tiny mesh built differently (centroids from the vertices, face
    gradients by solving the two diagonal equations, neighbour nodes matched by position).
"""
import numpy as np
import pytest

from meshgen import split_quads
from oracle_lib import OracleMesh, tvb_ops

G = 1.4


def face_tags(vxy, tags, box):
    """K x 3 boundary tags from the side of the box each face lies on (tags: dict W/E/S/N -> tag)"""
    x0, x1, y0, y1 = box
    K = vxy.shape[0]
    bct = np.zeros((K, 3), np.int32)
    for e in range(K):
        for f in range(3):
            a, b = vxy[e, f], vxy[e, (f + 1) % 3]
            mx, my = 0.5 * (a + b)
            for side, hit in (("W", abs(mx - x0) < 1e-12 and abs(a[0] - b[0]) < 1e-12),
                              ("E", abs(mx - x1) < 1e-12 and abs(a[0] - b[0]) < 1e-12),
                              ("S", abs(my - y0) < 1e-12 and abs(a[1] - b[1]) < 1e-12),
                              ("N", abs(my - y1) < 1e-12 and abs(a[1] - b[1]) < 1e-12)):
                if hit:
                    bct[e, f] = tags[side]
    return bct


def box_mesh(N, nx, ny, tags, box=(0.0, 1.0, 0.0, 1.0)):
    vxy, vid = split_quads(nx, ny, *box, periodic=(False, False))
    bct = face_tags(vxy, tags, box)
    return OracleMesh(N, vxy, vid, bct), vxy, vid, bct


def euler(rho, u, v, p):
    return np.stack([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)])


def deep_cells(om):
    E, _ = om.conn()
    b1 = np.any(E < 0, axis=1)
    b2 = np.array([np.any(b1[E[e][E[e] >= 0]]) for e in range(om.K)])
    return ~b1 & ~b2


def bc_trace(om, q):
    """fixedValue data = the state's own nodal values on each face (4 x K x 3 x Nfp)"""
    from oracle_lib import refelem_ops
    fm = refelem_ops(om.N)["fmask"]
    return np.ascontiguousarray(q[:, :, fm])


@pytest.mark.parametrize("N", [1, 2, 4])
def test_linear_density_pressure_exact(N):
    om, *_ = box_mesh(N, 10, 8, dict(W=2, E=2, S=2, N=2))
    x, y = om.nodes()
    q = euler(1 + 0.3 * x - 0.2 * y, 0.5 + 0 * x, -0.25 + 0 * x, 1 + 0.4 * x + 0.1 * y)
    out, fl = om.triangle_limit(q)
    deep = deep_cells(om)
    assert deep.sum() > 20 and np.all(fl & 1) and not np.any(fl & 6)
    np.testing.assert_allclose(out[:, deep], q[:, deep], rtol=0, atol=2e-14)
    # boundary cells are changed (the ghost copies break linearity there)
    assert np.abs(out[:, ~deep] - q[:, ~deep]).max() > 1e-6


@pytest.mark.parametrize("N", [2, 3])
def test_linear_velocity_energy_closed_form(N):
    om, vxy, *_ = box_mesh(N, 9, 9, dict(W=2, E=2, S=2, N=2))
    x, y = om.nodes()
    gu, gv, gp = np.array([0.5, -0.2]), np.array([0.3, 0.4]), np.array([0.4, 0.1])
    u = 0.3 + gu[0] * x + gu[1] * y
    v = -0.1 + gv[0] * x + gv[1] * y
    q = euler(1.2 + 0 * x, u, v, 2 + gp[0] * x + gp[1] * y)
    out, fl = om.triangle_limit(q)
    deep = deep_cells(om)
    np.testing.assert_allclose(out[:3, deep], q[:3, deep], rtol=0, atol=2e-14)
    pmean, _, _ = tvb_ops(N)
    c = vxy.mean(axis=1)                        # centroid from the vertices
    dx, dy = x - c[:, :1], y - c[:, 1:]
    uc = 0.3 + gu[0] * c[:, :1] + gu[1] * c[:, 1:]
    vc = -0.1 + gv[0] * c[:, :1] + gv[1] * c[:, 1:]
    Ebar = (q[3] @ pmean)[:, None]
    want = Ebar + (gp[0] * dx + gp[1] * dy) / (G - 1) + 1.2 * (uc * (gu[0] * dx + gu[1] * dy) +
                                                                 vc * (gv[0] * dx + gv[1] * dy))
    np.testing.assert_allclose(out[3, deep], want[deep], rtol=0, atol=5e-14)
    # and it is not the input energy (the quadratic part is dropped)
    assert np.abs(out[3, deep] - q[3, deep]).max() > 1e-4


def smooth_state(x, y):
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(np.pi * y)
    u = 0.4 * np.cos(np.pi * x) * np.sin(2 * np.pi * y)
    v = -0.3 + 0.2 * np.sin(np.pi * (x + y))
    p = 1.0 + 0.2 * np.cos(2 * np.pi * x * y)
    return euler(rho, u, v, p)


@pytest.mark.parametrize("N", [1, 3, 5])
def test_means_preserved_all_boundary_types(N):
    om, *_ = box_mesh(N, 8, 6, dict(W=3, E=2, S=1, N=3))
    x, y = om.nodes()
    q = smooth_state(x, y)
    bcq = bc_trace(om, smooth_state(x + 0.01, y))
    out, fl = om.triangle_limit(q, bcq=bcq)
    assert not np.any(fl & 2)
    pmean, _, _ = tvb_ops(N)
    np.testing.assert_allclose(out @ pmean, q @ pmean, rtol=1e-13, atol=1e-14)
    assert np.abs(out - q).max() > 1e-3


def test_constant_state_at_rest_unchanged():
    om, *_ = box_mesh(3, 6, 5, dict(W=3, E=2, S=1, N=0))
    x, y = om.nodes()
    q = euler(1.3 + 0 * x, 0 * x, 0 * x, 2.1 + 0 * x)
    out, fl = om.triangle_limit(q, bcq=bc_trace(om, q))
    np.testing.assert_allclose(out, q, rtol=0, atol=1e-14)
    assert np.all(fl == 1)


def mirror_meshes(N, nx, ny):
    """half mesh on [0,1]x[0,1] (wall at x = 0) and the full mesh on [-1,1]x[0,1] made of the half mesh
    (same elements first, same vertex order) and its mirror image"""
    vxy, _ = split_quads(nx, ny, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))
    mir = vxy[:, ::-1].copy()                     # reversed order keeps the mirror image anticlockwise
    mir[..., 0] *= -1.0
    full = np.concatenate([vxy, mir])
    key = {}

    def ids(v):
        out = np.zeros(v.shape[:2], np.int32)
        for e in range(v.shape[0]):
            for k in range(3):
                t = (round(v[e, k, 0] * nx), round(v[e, k, 1] * ny))
                out[e, k] = key.setdefault(t, len(key))
        return out

    vid_full = ids(full)
    vid_half = vid_full[:vxy.shape[0]].copy()
    half = OracleMesh(N, vxy, vid_half, face_tags(vxy, dict(W=3, E=2, S=2, N=2), (0.0, 1.0, 0.0, 1.0)))
    whole = OracleMesh(N, full, vid_full, face_tags(full, dict(W=2, E=2, S=2, N=2), (-1.0, 1.0, 0.0, 1.0)))
    return half, whole


@pytest.mark.parametrize("N", [1, 3])
def test_mirror_symmetry_with_reflective_wall(N):
    half, whole = mirror_meshes(N, 6, 5)

    def state(x, y):   # even in x, no x-velocity
        rho = 1.0 + 0.3 * np.cos(2.5 * x) * (1 + 0.5 * y * y) + 0.2 * np.exp(-8 * (x * x + (y - 0.4) ** 2))
        return euler(rho, 0 * x, 0.3 * np.cos(2 * x) * np.sin(3 * y), 1.0 + 0.3 * x * x * np.cos(y))

    xh, yh = half.nodes()
    xf, yf = whole.nodes()
    oh, _ = half.triangle_limit(state(xh, yh))
    of, _ = whole.triangle_limit(state(xf, yf))
    E, _ = half.conn()
    tags = face_tags(split_quads(6, 5, 0.0, 1.0, 0.0, 1.0, periodic=(False, False))[0],
                     dict(W=3, E=0, S=0, N=0), (0.0, 1.0, 0.0, 1.0))
    wall_cell = np.any(tags == 3, axis=1)
    keep = ~wall_cell
    assert wall_cell.sum() == 5 and np.abs(oh - state(xh, yh)).max() > 1e-3
    np.testing.assert_allclose(oh[:, keep], of[:, :half.K][:, keep], rtol=0, atol=1e-13)
    # the wall cells do differ: the ghost takes the face gradient, the mirror cell its own
    assert np.abs(oh[:, wall_cell] - of[:, :half.K][:, wall_cell]).max() > 1e-8


@pytest.mark.parametrize("N", [2, 4])
def test_rotation_commutes(N):
    vxy, vid = split_quads(7, 5, 0.0, 1.4, 0.0, 1.0, periodic=(False, False))
    bct = face_tags(vxy, dict(W=3, E=2, S=1, N=3), (0.0, 1.4, 0.0, 1.0))
    rot = np.stack([-vxy[..., 1], vxy[..., 0]], axis=-1)       # (x, y) -> (-y, x), exact
    a = OracleMesh(N, vxy, vid, bct)
    b = OracleMesh(N, rot, vid, bct)
    x, y = a.nodes()
    xr, yr = b.nodes()
    assert np.array_equal(xr, -y) and np.array_equal(yr, x)
    q = smooth_state(x, y)
    qr = np.stack([q[0], -q[2], q[1], q[3]])                   # momentum rotated with the mesh
    bcq = bc_trace(a, smooth_state(x + 0.02, y))
    bcr = np.stack([bcq[0], -bcq[2], bcq[1], bcq[3]])
    oa, fa = a.triangle_limit(q, bcq=bcq)
    ob, fb = b.triangle_limit(qr, bcq=bcr)
    want = np.stack([oa[0], -oa[2], oa[1], oa[3]])
    np.testing.assert_allclose(ob, want, rtol=0, atol=1e-13 * np.abs(q).max())
    assert np.array_equal(fa, fb)


def test_density_floor():
    N = 3
    om, *_ = box_mesh(N, 10, 4, dict(W=2, E=2, S=2, N=2))
    x, y = om.nodes()
    u0 = 0.7
    rho = np.where(x < 0.5, 0.0105, 1.0)
    q = euler(rho, u0 + 0 * x, 0 * x, np.where(x < 0.5, 0.005, 0.05))
    out, fl = om.triangle_limit(q, tol=1e-2)
    pmean, _, _ = tvb_ops(N)
    rb = (q[0] @ pmean)[:, None]
    floored = (fl & 2) != 0
    assert floored.sum() >= 2
    r = out[0]
    assert np.all(r[floored] >= 1e-2)
    # one halving step less would have been below the floor at some node of every floored cell
    assert np.all(np.any(rb + 2 * (r - rb) < 1e-2, axis=1)[floored])
    # the shortened du enters the momentum: rhou = m1bar + rhobar du1 + du ubar with du1 = 0 (u constant)
    m1b = (q[1] @ pmean)[:, None]
    np.testing.assert_allclose(out[1] - m1b, u0 * (r - rb), rtol=0, atol=1e-14)


def test_density_floor_below_tol_mean_terminates():
    om, *_ = box_mesh(2, 6, 2, dict(W=2, E=2, S=2, N=2))
    x, y = om.nodes()
    q = euler(0.004 + 0.002 * x, 0 * x, 0 * x, 0.01 + 0 * x)
    out, fl = om.triangle_limit(q, tol=1e-2)
    pmean, _, _ = tvb_ops(2)
    assert np.all(fl & 2)
    np.testing.assert_allclose(out[0], np.repeat((q[0] @ pmean)[:, None], om.Np, axis=1), rtol=0, atol=1e-17)


def brute_force(om, vxy, bct, q, bcq, N):
    """Second evaluation of the scheme, derived differently: centroids from the vertices, areas from cross
    products, face gradients by solving grad.(c_n - c_o) = dphibar and grad.(P1 - P0) = dphi, neighbour
    vertex values matched by position, every face evaluated from both sides."""
    K, Np = om.K, om.Np
    pmean, _, _ = tvb_ops(N)
    E, F = om.conn()
    xn, yn = om.nodes()
    vnode = [0, N, Np - 1]
    c = vxy.mean(axis=1)
    area = 0.5 * np.abs((vxy[:, 1, 0] - vxy[:, 0, 0]) * (vxy[:, 2, 1] - vxy[:, 0, 1]) -
                        (vxy[:, 2, 0] - vxy[:, 0, 0]) * (vxy[:, 1, 1] - vxy[:, 0, 1]))
    A0 = area / 3
    mean = q @ pmean                                     # [4][K]

    def prim(m):
        return np.array([m[0], m[1] / m[0], m[2] / m[0], (G - 1) * (m[3] - 0.5 * (m[1] ** 2 + m[2] ** 2) / m[0])])

    def endpoint(e, f, s):
        k = f if s == 0 else (f + 1) % 3
        return q[:, e, vnode[k]], vxy[e, k]

    def face(e, f):
        """face gradient [4][2], A_2, and the ghost gradient flag"""
        P = [endpoint(e, f, s) for s in (0, 1)]
        en = E[e, f]
        ex = vxy[e, (f + 1) % 3] - vxy[e, f]
        n = np.array([ex[1], -ex[0]]) / np.hypot(*ex)
        if en >= 0:
            vals = []
            for s in (0, 1):
                d = np.hypot(*(vxy[en] - P[s][1]).T)
                k = int(np.argmin(d))
                assert d[k] < 1e-12
                vals.append(q[:, en, vnode[k]])
            cn, mn, A2 = c[en], mean[:, en], A0[e] + A0[en]
        else:
            tag = bct[e, f]
            vals = []
            for s in (0, 1):
                qs = P[s][0].copy()
                if tag == 1:
                    qs = bcq[:, e, f, 0 if s == 0 else N]
                elif tag == 3:
                    qs[1:3] -= 2 * (qs[1:3] @ n) * n
                vals.append(qs)
            cn = c[e] - 2 * ((c[e] - P[0][1]) @ n) * n
            mn = mean[:, e].copy()
            if tag == 1:
                mn = bcq[:, e, f, 0].copy()
            elif tag == 3:
                mn[1:3] -= (mn[1:3] @ n) * n
            A2 = 2 * A0[e]
        pe = [prim(0.5 * (P[s][0] + vals[s])) for s in (0, 1)]
        po, pn = prim(mean[:, e]), prim(mn)
        M = np.array([cn - c[e], P[1][1] - P[0][1]])
        grads = np.array([np.linalg.solve(M, [pn[k] - po[k], pe[1][k] - pe[0][k]]) for k in range(4)])
        return grads, A2

    fg = [[face(e, f) for f in range(3)] for e in range(K)]
    CV = np.array([sum(fg[e][f][1] * fg[e][f][0] for f in range(3)) / sum(fg[e][f][1] for f in range(3))
                   for e in range(K)])                 # [K][4][2]
    out = q.copy()
    for e in range(K):
        nbg = [CV[E[e, f]] if E[e, f] >= 0 else fg[e][f][0] for f in range(3)]
        Lg = np.zeros((4, 2))
        for k in range(4):
            g = np.array([nbg[i][k] @ nbg[i][k] for i in range(3)])
            w = np.array([g[1] * g[2], g[0] * g[2], g[0] * g[1]]) + 1e-10
            w /= np.sum(g * g) + 3e-10
            Lg[k] = sum(w[i] * nbg[i][k] for i in range(3))
        pm = prim(mean[:, e])
        d = np.stack([xn[e] - c[e, 0], yn[e] - c[e, 1]], axis=1) @ Lg.T     # [Np][4]
        du = d[:, 0]
        assert np.all(pm[0] + du >= 1e-2)
        out[0, e] = pm[0] + du
        out[1, e] = mean[1, e] + pm[0] * d[:, 1] + du * pm[1]
        out[2, e] = mean[2, e] + pm[0] * d[:, 2] + du * pm[2]
        out[3, e] = (mean[3, e] + d[:, 3] / (G - 1) + 0.5 * du * (pm[1] ** 2 + pm[2] ** 2) +
                     pm[0] * (pm[1] * d[:, 1] + pm[2] * d[:, 2]))
    return out


@pytest.mark.parametrize("N", [1, 2, 3])
def test_brute_force_tiny_mesh(N):
    vxy, vid = split_quads(3, 2, 0.0, 1.0, 0.0, 0.8, periodic=(False, False))
    bct = face_tags(vxy, dict(W=3, E=2, S=1, N=3), (0.0, 1.0, 0.0, 0.8))
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    rng = np.random.default_rng(N)
    q = smooth_state(x, y) + 0.05 * rng.standard_normal((4, om.K, om.Np)) * np.array([1, 0.3, 0.3, 1])[:, None, None]
    bcq = bc_trace(om, smooth_state(x - 0.03, y))
    got, fl = om.triangle_limit(q, bcq=bcq)
    want = brute_force(om, vxy, bct, q, bcq, N)
    assert np.all(fl == 1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(q).max())


def test_selection_keeps_unselected_cells():
    om, *_ = box_mesh(2, 6, 4, dict(W=3, E=2, S=2, N=2))
    x, y = om.nodes()
    q = smooth_state(x, y)
    sel = (np.arange(om.K) % 3 == 0).astype(np.int32)
    all_, _ = om.triangle_limit(q)
    part, fl = om.triangle_limit(q, sel=sel)
    s = sel.astype(bool)
    np.testing.assert_array_equal(part[:, ~s], q[:, ~s])
    np.testing.assert_array_equal(part[:, s], all_[:, s])
    assert np.array_equal(fl & 1, sel)


def test_patch_first_ghost_mean():
    """The reference's fixedValue ghost mean is the first value of the whole patch (Trianglelimite.C:220-223),
    while the face endpoints keep each face's own data (:381-391).  With patch data whose faces all share the value
    at their first node, the patch map changes nothing (bitwise); after changing only the named first face's first
    value, the patch-map run moves every cell on the patch (each takes the new ghost mean), while the own-face run
    moves only the cell of that face."""
    from meshgen import split_quads
    N = 2
    box = (0.0, 1.0, 0.0, 1.0)
    vxy, vid = split_quads(6, 6, *box, periodic=(False, False))
    bct = face_tags(vxy, dict(W=1, E=2, S=3, N=2), box)
    om = OracleMesh(N, vxy, vid, bct)
    x, y = om.nodes()
    q = smooth_state(x, y)
    K, Nfp = om.K, om.Nfp
    E, _ = om.conn()
    west = (E < 0) & (bct == 1)
    cells = np.nonzero(west.any(axis=1))[0]
    assert len(cells) >= 4
    rng = np.random.default_rng(4)
    vals = np.array([1.3, 0.2, -0.1, 2.9])
    bcq = np.zeros((4, K, 3, Nfp))
    bcq[:, west, :] = vals[:, None, None] * (1 + 0.3 * rng.random((1, west.sum(), Nfp)))
    bcq[:, west, 0] = vals[:, None]                  # every face of the patch starts with the same value
    first = np.argwhere(west)[3]
    pf = np.where(west, 3 * first[0] + first[1], -1).astype(np.int32)
    a, _ = om.triangle_limit(q, bcq=bcq, patch_first=pf)
    b, _ = om.triangle_limit(q, bcq=bcq)
    np.testing.assert_array_equal(a, b)
    bcq2 = bcq.copy()
    bcq2[:, first[0], first[1], 0] *= 1.5
    a2, _ = om.triangle_limit(q, bcq=bcq2, patch_first=pf)
    b2, _ = om.triangle_limit(q, bcq=bcq2)
    moved_a = np.abs(a2 - a).max(axis=(0, 2)) > 1e-12
    moved_b = np.abs(b2 - b).max(axis=(0, 2)) > 1e-12
    assert moved_b[first[0]] and moved_b.sum() <= 4        # the face's cell (and at most its neighbours)
    assert moved_a[cells].all()                            # every cell on the patch
