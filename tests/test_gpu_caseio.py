"""A reference tutorial case run from its own files: the isentropic-vortex case directory (Fluent mesh
vortex1024.msh, system/dgSolution baseOrder 4, system/controlDict endTime 2, system/dgSchemes godunovScheme
fluxScheme Roe, fixedValue Wall patches in 0/{rho,rhoU,Ener}) read by hopefoam_amd.caseio, its flux scheme taken
from the dictionary, advanced on the GPU by the library's SSP-RK2 loop with the exact-vortex boundary values
(setBoundaryValues.H:9-50) evaluated on the device, and compared with the User Guide's known answer; the final
fields are written in the reference's format and read back."""
import os
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

dgh = pytest.importorskip("hopefoam_amd.dgh")
cases = pytest.importorskip("hopefoam_amd.cases")
from hopefoam_amd import caseio  # noqa: E402

VORTEX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "isentropicVortex")


def test_vortex_case_known_answer(tmp_path):
    path = tmp_path / "isentropicVortex"
    shutil.copytree(VORTEX, path)
    case = caseio.Case(path, mesh="vortex1024.msh")
    m2 = case.mesh2d()
    # this mesh names its boundary zone WALL while the tutorial's field files say Wall: patch names are
    # case-sensitive (as in OpenFOAM), so the fields are rewritten for it as a user preparing the case would
    assert m2.K == 1024 and m2.patches == ["WALL"]
    with pytest.raises(caseio.CaseError):
        case.bc_tags(["rho", "rhoU", "Ener"])
    K, Np = m2.K, 15
    for name, comps in (("rho", 1), ("rhoU", 3), ("Ener", 1)):
        zero = np.zeros((K, Np)) if comps == 1 else np.zeros((3, K, Np))
        case.write_field("0", name, zero, boundary={"WALL": "fixedValue", "frontAndBackPlanes": "empty"})
    N, t_end = case.N, float(case.control["endTime"])
    assert (N, t_end) == (4, 2.0)
    # the flux from the case dictionary (godunovScheme { fluxScheme Roe; }); this solver application calls no
    # limiter (its Godunov.limite lines are commented out, isentropicVortex/dgEulerFoam/dgEulerFoam.C:92, :119),
    # although the dictionary names one for the godunovScheme constructor
    d = case.euler_rhs_desc()
    assert (d.eq, d.flux, d.gamma) == (dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, 1.4)
    assert case.limiter_desc().kind == dgh.DGH_LIM_TRIANGLE
    # the shipped controlDict says deltaT 0.008, above the stable step of User Guide table 1.1 for this mesh and
    # order (0.004), with which the guide's published error was computed (it reproduces it to 1e-9; 0.008 does not)
    assert float(case.control["deltaT"]) == 0.008
    dt = 0.004
    nsteps = int(round(t_end / dt))
    tag = case.bc_tags(["rho", "rhoU", "Ener"])
    ctx = dgh.Context(0)
    try:
        mesh = ctx.mesh(N, m2.vertices(), m2.tris, tag)
        x, y = mesh.nodes()
        q = mesh.field(4, cases.vortex(x, y, 0.0, x0=5.0, y0=0.0))
        mesh.set_bc_function(dgh.DGH_BCFN_ISENTROPIC_VORTEX, 0.0, gamma=case.gamma(), x0=5.0, y0=0.0, beta=5.0)
        mesh.ssprk2_run(d, q, nsteps, dt)
        qq = q.download()
    finally:
        ctx.close()
    ex = cases.vortex(x, y, nsteps * dt, x0=5.0, y0=0.0)
    er = np.mean(np.abs(ex[0] - qq[0]))
    eu = np.mean(np.hypot(ex[1] - qq[1], ex[2] - qq[2]))
    print("case files: rhoError %.15e rhoUError %.15e" % (er, eu), flush=True)
    assert abs(er / 8.807979526797244e-06 - 1) < 1e-9
    assert abs(eu / 1.865574862711117e-05 - 1) < 1e-9
    bnd_types = {"WALL": "fixedValue", "frontAndBackPlanes": "empty"}
    case.write_field(t_end, "rho", qq[0], "[1 -3 0 0 0 0 0]", bnd_types)
    case.write_field(t_end, "rhoU", qq[1:3], "[1 -2 -1 0 0 0 0]", bnd_types)
    case.write_field(t_end, "Ener", qq[3], "[1 -1 -2 0 0 0 0]", bnd_types)
    assert case.times() == ["0", "2"]
    rho, _ = case.read_field("rho", "2")
    rhou, _ = case.read_field("rhoU", "2")
    np.testing.assert_array_equal(rho[0], qq[0])
    np.testing.assert_array_equal(rhou[:2], qq[1:3])
    # the written time directory declares the same boundary types, so it tags the mesh as the initial one did
    assert case.bc_tags(["rho", "rhoU", "Ener"], "2").tolist() == tag.tolist()
