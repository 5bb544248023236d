"""The reference's own Euler driver on the GPU: SSP-RK2 (dgEulerFoam.C:64-123), Roe flux, fixedValue
boundaries from the exact vortex at t_n (setBoundaryValues.H:9-50), mean-abs errors (eulerError.H:37-38).
The published known answers are the bar: User Guide 1.7-1.8 (vortex1024, N = 4, t = 2) and four-digit
entries of the presentation's slide 18."""
import numpy as np
import pytest

from meshgen import load_fixture_mesh

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")
cases = pytest.importorskip("hopefoam_amd.cases")


@pytest.fixture(scope="module")
def ctx():
    c = dgh.Context(0)
    yield c
    c.close()


def vortex_run(ctx, name, N, dt, nsteps):
    vxy, tri = load_fixture_mesh(name)
    mesh = ctx.mesh(N, vxy, tri, np.ones_like(tri))
    x, y = mesh.nodes()
    fm = dgh.refelem_get(N, "fmask").astype(int).reshape(3, N + 1)
    E, _ = mesh.conn()
    bnd = E < 0
    xb = np.stack([x[:, fm[f]] for f in range(3)], axis=1)        # [K, 3, Nfp]
    yb = np.stack([y[:, fm[f]] for f in range(3)], axis=1)
    q = mesh.field(4, cases.vortex(x, y, 0.0, x0=5.0, y0=0.0))
    q1, res = mesh.field(4), mesh.field(4)
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    for step in range(nsteps):
        bc = cases.vortex(xb, yb, step * dt, x0=5.0, y0=0.0) * bnd[None, :, :, None]
        mesh.set_bc_values(np.ascontiguousarray(bc))
        mesh.ssprk2_stage(d, q, q1, res, 0, dt)
        mesh.ssprk2_stage(d, q1, q, res, 1, dt)
    ex = cases.vortex(x, y, nsteps * dt, x0=5.0, y0=0.0)
    qq = q.download()
    er = np.mean(np.abs(ex[0] - qq[0]))
    eu = np.mean(np.hypot(ex[1] - qq[1], ex[2] - qq[2]))
    return er, eu


def test_user_guide_known_answer(ctx):
    er, eu = vortex_run(ctx, "vortex1024", 4, 0.004, 500)
    print("rhoError %.15e (UG 8.807979526797244e-06), rhoUError %.15e (UG 1.865574862711117e-05)" % (er, eu))
    assert abs(er / 8.807979526797244e-06 - 1) < 1e-9
    assert abs(eu / 1.865574862711117e-05 - 1) < 1e-9


@pytest.mark.parametrize("name,N,dt,erho,erhou", [("vortex0256", 2, 0.02, 3.481e-03, 6.384e-03),
                                                  ("vortex1024", 2, 0.01, 3.621e-04, 8.060e-04)])
def test_slide18_entries(ctx, name, N, dt, erho, erhou):
    er, eu = vortex_run(ctx, name, N, dt, int(round(2.0 / dt)))
    assert abs(er / erho - 1) < 6e-4 and abs(eu / erhou - 1) < 6e-4
