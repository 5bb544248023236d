"""Case I/O (hopefoam_amd/caseio.py) on the reference's own tutorial files (fixtures under tests/golden/cases,
copied by tests/golden/make_case_fixtures.py): the Fluent mesh of the isentropic vortex, the cylinder's
prism polyMesh, the dictionaries and DG field files; and field write/read round trips."""
import os

import numpy as np
import pytest

from hopefoam_amd import caseio
from hopefoam_amd.dgh import DGH_BC_FIXED_VALUE, DGH_BC_WALL, DGH_BC_ZERO_GRADIENT
from meshgen import load_fixture_mesh

CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")
VORTEX = os.path.join(CASES, "isentropicVortex")
CYLINDER = os.path.join(CASES, "cylinder")


def signed_areas(m):
    v = m.vertices()
    return 0.5 * ((v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) -
                  (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0]))


def check_conforming(m):
    """every face is either on exactly one boundary patch (used once) or interior (used twice, reversed)"""
    uses = {}
    for e, t in enumerate(m.tris):
        for k in range(3):
            a, b = int(t[k]), int(t[(k + 1) % 3])
            uses.setdefault((min(a, b), max(a, b)), []).append((e, k, a < b))
    for key, u in uses.items():
        if len(u) == 1:
            e, k, _ = u[0]
            assert m.face_patch[e, k] >= 0, key
        else:
            assert len(u) == 2 and u[0][2] != u[1][2], key
            for e, k, _ in u:
                assert m.face_patch[e, k] == -1


def test_fluent_mesh_matches_fixture():
    m = caseio.read_fluent(os.path.join(VORTEX, "vortex0256.msh"))
    vxy, tri = load_fixture_mesh("vortex0256")
    assert m.K == 256 and m.patches == ["Wall"]
    np.testing.assert_array_equal(m.tris, tri)
    np.testing.assert_array_equal(m.vertices(), vxy)
    assert (signed_areas(m) > 0).all()
    check_conforming(m)
    # the square [0,10] x [-5,5] of the tutorial, 8 boundary faces per side
    assert abs(signed_areas(m).sum() - 100.0) < 1e-9
    assert (m.face_patch == 0).sum() == 32


def test_polymesh_cylinder():
    d = os.path.join(CYLINDER, "constant", "polyMesh")
    m = caseio.read_polymesh(d)
    assert m.K == 1840
    bnd = caseio.parse_dict(caseio._foam_body(os.path.join(d, "boundary")))
    entries = next(v for v in bnd.values() if isinstance(v, list))
    nfaces = {str(entries[i]): int(entries[i + 1]["nFaces"]) for i in range(0, len(entries), 2)
              if entries[i + 1]["type"] != "empty"}
    assert m.patches == list(nfaces)
    for p, name in enumerate(m.patches):
        assert (m.face_patch == p).sum() == nfaces[name]
    assert (signed_areas(m) > 0).all()
    check_conforming(m)
    # the rule of dgPolyMesh.C:490-509: each triangle is its z = 0 face in stored order, or with points 1 and 2
    # swapped; checked against the raw face list
    pts = caseio._read_points(os.path.join(d, "points"))
    faces = caseio._read_faces(os.path.join(d, "faces"))
    z0 = [f for f in faces if len(f) == 3 and all(pts[q, 2] == 0 for q in f)]
    assert len(z0) == m.K
    base = {tuple(sorted(f)): f for f in z0}
    used = np.unique(np.array(z0).reshape(-1))
    for t in m.tris:
        orig = tuple(int(used[q]) for q in t)
        f = base[tuple(sorted(orig))]
        assert orig in ((f[0], f[1], f[2]), (f[0], f[2], f[1]))


def test_dictionaries():
    c = caseio.Case(VORTEX)
    assert c.N == 4
    assert c.control["deltaT"] == 0.008 and c.control["endTime"] == 2 and c.control["writePrecision"] == 16
    assert c.schemes["godunovScheme"] == {"fluxScheme": "Roe", "limiteScheme": "Triangle"}
    assert c.solution["solvers"]["dgrho"]["kspSolver"] == "preonly"
    cy = caseio.Case(CYLINDER)
    assert cy.N == 3 and cy.mesh_file().endswith("polyMesh")


def test_reference_field_files():
    c = caseio.Case(VORTEX)
    rho, b = c.read_field("rho")
    assert rho.shape == (1, 256, 15) and (rho == 0).all()
    assert b == {"Wall": {"type": "fixedValue"}, "frontAndBackPlanes": {"type": "empty"}}
    ru, _ = c.read_field("rhoU")
    assert ru.shape == (3, 256, 15) and (ru == 0).all()


def test_field_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    K, Np = 37, 10
    s = rng.standard_normal((K, Np)) * 10.0 ** rng.integers(-300, 300, (K, Np))
    v = rng.standard_normal((2, K, Np))
    caseio.write_dg_field(tmp_path / "0.5" / "rho", "rho", s, "[1 -3 0 0 0 0 0]",
                          {"Wall": "fixedValue", "frontAndBackPlanes": "empty"}, "0.5")
    caseio.write_dg_field(tmp_path / "0.5" / "rhoU", "rhoU", v, "[1 -2 -1 0 0 0 0]", {"Wall": "reflective"})
    got, b = caseio.read_dg_field(tmp_path / "0.5" / "rho", K, Np)
    np.testing.assert_array_equal(got[0], s)
    assert b["Wall"]["type"] == "fixedValue"
    gv, _ = caseio.read_dg_field(tmp_path / "0.5" / "rhoU", K, Np)
    np.testing.assert_array_equal(gv[:2], v)
    assert (gv[2] == 0).all()
    with pytest.raises(caseio.CaseError):
        caseio.read_dg_field(tmp_path / "0.5" / "rho", K + 1, Np)


def test_boundary_tags():
    c = caseio.Case(VORTEX)
    m = c.mesh2d()
    tag = c.bc_tags(["rho", "rhoU", "Ener"])
    assert (tag[m.face_patch == 0] == DGH_BC_FIXED_VALUE).all()
    cy = caseio.Case(CYLINDER)
    mc = cy.mesh2d()
    t = cy.bc_tags(["U"])
    for p, name in enumerate(mc.patches):
        want = DGH_BC_ZERO_GRADIENT if name == "outlet" else DGH_BC_FIXED_VALUE
        assert (t[mc.face_patch == p] == want).all()
    # velocity fixed and pressure zeroGradient on the same wall: no single device tag
    with pytest.raises(caseio.CaseError):
        cy.bc_tags(["U", "p"])


def test_wall_tag_from_reflective_momentum(tmp_path):
    import shutil
    shutil.copytree(VORTEX, tmp_path / "case")
    K, Np = 256, 15
    caseio.write_dg_field(tmp_path / "case" / "0" / "rhoU", "rhoU", np.zeros((3, K, Np)), boundary={
        "Wall": "reflective", "frontAndBackPlanes": "empty"})
    for name in ("rho", "Ener"):
        caseio.write_dg_field(tmp_path / "case" / "0" / name, name, np.ones((K, Np)), boundary={
            "Wall": "zeroGradient", "frontAndBackPlanes": "empty"})
    c = caseio.Case(tmp_path / "case")
    tag = c.bc_tags(["rho", "rhoU", "Ener"])
    assert (tag[c.mesh2d().face_patch == 0] == DGH_BC_WALL).all()


def test_scheme_words_dispatch():
    """The dictionaries' scheme words select the device descriptors: godunovScheme Roe / Triangle of the Euler
    tutorials, divSchemes LF / average of the projection solver (keywords with argument lists such as div(U,U)),
    and the dgSolution solver entries; unknown words raise."""
    iv = caseio.Case(VORTEX)
    d = iv.euler_rhs_desc()
    assert (d.flux, d.gamma) == (2, 1.4)                    # DGH_FLUX_ROE
    assert iv.limiter_desc().kind == 3                      # DGH_LIM_TRIANGLE
    pv = caseio.Case(os.path.join(CASES, "pearsonVortex"))
    assert pv.advection_rhs_desc("div(U,U)", 1.0, 0.0).flux == 0      # LF = upwind for a given velocity
    assert pv.advection_rhs_desc("div(U)", 1.0, 0.0).flux == 3        # average
    assert pv.advection_rhs_desc("div(T)", 1.0, 0.0).flux == 3        # the default entry
    assert pv.limiter_desc() is None
    s = pv.sipg_settings("p")
    assert s == dict(penalty=0, pc=1, rtol=0.0, abstol=1e-12, maxit=10000)   # cg + lu: block Jacobi
    with pytest.raises(caseio.CaseError):
        pv.euler_rhs_desc()
    dm = caseio.Case(os.path.join(CASES, "doubleMach"))
    assert dm.euler_rhs_desc().flux == 2 and dm.limiter_desc().kind == 3
    with pytest.raises(caseio.CaseError):
        dm.sipg_settings("rho1")                            # an OpenFOAM-style PCG entry, not a KSP
    dm.schemes["godunovScheme"]["fluxScheme"] = "HLLC"
    with pytest.raises(caseio.CaseError):
        dm.euler_rhs_desc()
    dm.schemes["godunovScheme"] = {"fluxScheme": "hipRoe", "limiteScheme": "hipTriangle", "troubledCell": "TVB"}
    assert dm.limiter_desc().kind == 4                      # DGH_LIM_TRIANGLE_TVB
    dm.schemes["godunovScheme"]["fluxScheme"] = ["hipLF", "bf16x3"]
    d = dm.euler_rhs_desc()
    assert (d.flux, d.prec) == (1, 2)
