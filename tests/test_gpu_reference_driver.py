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


def test_user_guide_known_answer_library_loop(ctx):
    """The same run inside the library: dgh_ssprk2_run with the exact-vortex boundary function evaluated on the
    device at t^n before every step (the tutorial's setBoundaryValues.H at runTime - deltaT), no host round
    trips.  It must reproduce the User Guide answer as the host-driven loop does, and leave the mesh's time at
    t = 2."""
    vxy, tri = load_fixture_mesh("vortex1024")
    mesh = ctx.mesh(4, vxy, tri, np.ones_like(tri))
    x, y = mesh.nodes()
    q = mesh.field(4, cases.vortex(x, y, 0.0, x0=5.0, y0=0.0))
    mesh.set_bc_function(dgh.DGH_BCFN_ISENTROPIC_VORTEX, 0.0, x0=5.0, y0=0.0, beta=5.0, u_inf=1.0, v_inf=0.0)
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    mesh.ssprk2_run(d, q, 500, 0.004)
    assert abs(mesh.time() - 2.0) < 1e-12
    ex = cases.vortex(x, y, 2.0, x0=5.0, y0=0.0)
    qq = q.download()
    er = np.mean(np.abs(ex[0] - qq[0]))
    eu = np.mean(np.hypot(ex[1] - qq[1], ex[2] - qq[2]))
    print("library loop: rhoError %.15e, rhoUError %.15e" % (er, eu), flush=True)
    assert abs(er / 8.807979526797244e-06 - 1) < 1e-9
    assert abs(eu / 1.865574862711117e-05 - 1) < 1e-9


def test_boundary_function_matches_host_values(ctx):
    """The device-evaluated ghost states equal the host formula's (cases.vortex at the face nodes) to round-off:
    one right-hand side with each, on the reference's vortex mesh at t = 0.7."""
    vxy, tri = load_fixture_mesh("vortex0256")
    N = 3
    mesh = ctx.mesh(N, vxy, tri, np.ones_like(tri))
    x, y = mesh.nodes()
    fm = dgh.refelem_get(N, "fmask").astype(int).reshape(3, N + 1)
    E, _ = mesh.conn()
    xb = np.stack([x[:, fm[f]] for f in range(3)], axis=1)
    yb = np.stack([y[:, fm[f]] for f in range(3)], axis=1)
    t = 0.7
    q = mesh.field(4, cases.vortex(x, y, t, x0=5.0, y0=0.0) * 1.01)
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    r1, r2 = mesh.field(4), mesh.field(4)
    mesh.set_bc_values(np.ascontiguousarray(cases.vortex(xb, yb, t, x0=5.0, y0=0.0) * (E < 0)[None, :, :, None]))
    mesh.rhs(d, q, r1)
    mesh.set_bc_function(dgh.DGH_BCFN_ISENTROPIC_VORTEX, t)
    mesh.rhs(d, q, r2)
    a, b = r1.download(), r2.download()
    err = cases.rel_l2(b, a)
    print("device vs host boundary values: rhs rel L2 %.3e" % err, flush=True)
    assert err < 1e-13


def test_lserk4_run_boundary_function_stage_times(ctx):
    """dgh_lserk4_run with a boundary function evaluates it at every stage time t^n + c_s dt: equal, bit for bit,
    to per-stage calls with the function set to those times by the host."""
    vxy, tri = load_fixture_mesh("vortex0256")
    N, dt, nsteps = 3, 0.004, 6
    mesh = ctx.mesh(N, vxy, tri, np.ones_like(tri))
    x, y = mesh.nodes()
    q0 = cases.vortex(x, y, 0.0, x0=5.0, y0=0.0)
    d = dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_LF, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    c = [0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
         2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0]
    qa = mesh.field(4, q0)
    mesh.set_bc_function(dgh.DGH_BCFN_ISENTROPIC_VORTEX, 0.0)
    mesh.lserk4_run(d, qa, nsteps, dt)
    assert abs(mesh.time() - nsteps * dt) < 1e-15
    cur, nxt, res = mesh.field(4, q0), mesh.field(4), mesh.field(4, np.zeros_like(q0))
    for n in range(nsteps):
        for s in range(5):
            mesh.set_bc_function(dgh.DGH_BCFN_ISENTROPIC_VORTEX, (n + c[s]) * dt)
            mesh.lserk4_stage(d, cur, nxt, res, s, dt)
            cur, nxt = nxt, cur
    assert np.array_equal(qa.download(), cur.download())
