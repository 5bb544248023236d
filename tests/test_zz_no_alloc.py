"""dgh_limit_noalloc (the limiter with its scratch in the library's static device array) against dgh_limit through
the C ABI: the same flags, counts and limited values, bit for bit, on two calls in a row (the second on a
smaller mesh), and its refusals (tables not built yet, Triangle kinds, K above the array's 131072 cells)."""
import os

import numpy as np
import pytest

from meshgen import split_quads

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")
cases = pytest.importorskip("hopefoam_amd.cases")
indicator = pytest.importorskip("hopefoam_amd.indicator")

WEIGHTS_N6 = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hopefoam_amd", "data",
                          "indicator_N6.npz")


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def riemann_mesh(ctx, N, nx, ny, seed):
    vxy, vid = split_quads(nx, ny, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, dgh.DGH_BC_ZERO_GRADIENT, np.int32)
    mesh = ctx.mesh(N, vxy, vid, bct)
    x, y = mesh.nodes()
    return mesh, cases.riemann4(x, y, seed)


def both(mesh, q, **kw):
    """the limiter on two copies of q, through dgh_limit and dgh_limit_noalloc"""
    fa, fb = mesh.field(4, q), mesh.field(4, q)
    ta, na = mesh.limit(fa, **kw)
    tb, nb = mesh.limit(fb, noalloc=True, **kw)
    return (ta, na, fa.download()), (tb, nb, fb.download())


@pytest.mark.parametrize("kind", ["tvb", "tvb+positivity", "mlp", "all"])
def test_noalloc_matches_limit(ctx, kind):
    net = indicator.load_net(ctx, WEIGHTS_N6)
    kw = {"M": 50.0}
    if kind == "tvb+positivity":
        kw.update(positivity=True, gamma=1.4, eps=1e-10)
    elif kind == "mlp":
        kw.update(mlp=net)
    elif kind == "all":
        kw.update(kind=dgh.DGH_LIM_ALL)
    # first call on a larger mesh, then a smaller one: the scratch keeps the first call's contents
    for nx, ny, seed in ((64, 16, 21), (24, 6, 22)):
        mesh, q = riemann_mesh(ctx, 6, nx, ny, seed)
        print("kind %s N=6 K=%d seed %d" % (kind, mesh.K, seed), flush=True)
        (ta, na, qa), (tb, nb, qb) = both(mesh, q, **kw)
        print("  troubled %d / %d (dgh_limit %d)" % (nb, mesh.K, na), flush=True)
        assert na == nb
        np.testing.assert_array_equal(ta, tb)
        np.testing.assert_array_equal(qa, qb)
        assert 0 < na


def test_noalloc_refusals(ctx):
    mesh, q = riemann_mesh(ctx, 3, 16, 4, 5)
    f = mesh.field(4, q)
    # the mesh's limiter tables do not exist before its first dgh_limit call
    with pytest.raises(dgh.DghError, match="limiter tables"):
        mesh.limit(f, M=50.0, noalloc=True)
    np.testing.assert_array_equal(f.download(), q)
    mesh.limit(mesh.field(4, q), M=50.0)
    with pytest.raises(dgh.DghError, match="Triangle"):
        d = mesh.limiter_desc(kind=dgh.DGH_LIM_TRIANGLE, eps=1e-2)
        st = ctx.L.dgh_limit_noalloc(ctx.h, mesh.h, d, f.h, None, None)
        ctx._ck(st, "dgh_limit_noalloc")
    big, qb = riemann_mesh(ctx, 1, 512, 130, 6)   # K = 133120 > 131072
    fb = big.field(4, qb)
    big.limit(big.field(4, qb), M=50.0)
    with pytest.raises(dgh.DghError, match="exceeds"):
        big.limit(fb, M=50.0, noalloc=True)
