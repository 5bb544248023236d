"""Case I/O of the reference's 2D DG cases: meshes, fields and dictionaries in the reference's own file formats.

A HopeFOAM 2D case directory holds
  * a mesh: a Fluent ASCII mesh (``*.msh`` beside ``system/``, as the tutorials ship them, converted by the
    reference with ``fluentMeshToFoam``) or an OpenFOAM ``constant/polyMesh`` of one layer of prisms, whose
    z = 0 faces are the triangles (``src/DG/dgMesh/dgPolyMesh.C:105-199``);
  * fields ``<time>/<name>``: ``dgScalarField`` / ``dgVectorField`` files in OpenFOAM dictionary syntax whose
    ``internalField`` lists the K Np nodal values element by element (``[K][Np]``, the reference's
    ``GeometricDofField`` order), vectors as ``(x y z)``;
  * dictionaries ``system/controlDict``, ``system/dgSolution`` (``DG { baseOrder N; }``), ``system/dgSchemes``.

This module reads and writes those files on the host and hands the mesh to the device path
(``Context.mesh``); it does no arithmetic of the scheme.  Triangles are anticlockwise with face k joining
vertices k and k+1, as the device path (and the reference) expects:
  * from a polyMesh, each triangle keeps the point order of its z = 0 face, with points 1 and 2 swapped when
    that order is clockwise (``dgPolyMesh.C:490-509``), so field files are interchangeable with the reference's;
  * from a Fluent mesh, each triangle starts at its lowest vertex id and is then made anticlockwise.
    The reference orders such a case by ``fluentMeshToFoam``'s output instead; the DG solution does not depend
    on the starting vertex (the nodes are symmetric under the triangle's rotations), but the node order of
    each element in a written field does.

Boundary conditions map onto the device path's face tags:
  ``fixedValue`` -> DGH_BC_FIXED_VALUE, ``zeroGradient`` -> DGH_BC_ZERO_GRADIENT, ``reflective`` and ``slip``
  -> DGH_BC_WALL (momentum mirrored, scalars copied), ``empty`` -> the front and back planes of the 2D mesh.
Other patch types (``cyclic``, ``processor``, ``mixed``, ``freestream``, ...) raise ``CaseError``.
"""
import os
import re

import numpy as np

from .dgh import (DGH_BC_FIXED_VALUE, DGH_BC_WALL, DGH_BC_ZERO_GRADIENT, DGH_EQ_ADVECT, DGH_EQ_EULER, DGH_FLUX_AVERAGE,
                  DGH_FLUX_LF, DGH_FLUX_ROE, DGH_FLUX_UPWIND, DGH_LIM_TRIANGLE, DGH_LIM_TRIANGLE_TVB, DGH_LIM_TVB,
                  DGH_PC_BLOCK_JACOBI, DGH_PC_JACOBI, DGH_PREC_BF16X3, DGH_PREC_FP32, DGH_PREC_FP64,
                  DGH_SIPG_PENALTY_REFERENCE)


class CaseError(RuntimeError):
    pass


# ---------------------------------------------------------------------------------------------- dictionaries
# a keyword may carry a parenthesised argument list without spaces (div(U,U), grad(p), "rho(1|2|3)")
_TOKEN = re.compile(r'\s+|//[^\n]*|/\*.*?\*/|#\{.*?#\}|"(?:[^"\\]|\\.)*"|[A-Za-z_][\w.:]*\([^\s{}();"]*\)|[{}()\[\];]|'
                    r'[^\s{}()\[\];"]+', re.S)


def _tokens(text):
    out = []
    for m in _TOKEN.finditer(text):
        t = m.group(0)
        if t[0].isspace() or t.startswith("//") or t.startswith("/*"):
            continue
        out.append(t)
    return out


def _number(t):
    try:
        return int(t)
    except ValueError:
        try:
            return float(t)
        except ValueError:
            return t


def _parse_list(toks, i, close):
    """toks[i] is just after the opening bracket; returns (list, index after the closing bracket)"""
    out = []
    while toks[i] != close:
        if toks[i] == "(":
            v, i = _parse_list(toks, i + 1, ")")
            out.append(v)
        elif toks[i] == "[":
            v, i = _parse_list(toks, i + 1, "]")
            out.append(v)
        elif toks[i] == "{":
            v, i = _parse_dict(toks, i + 1, "}")
            out.append(v)
        else:
            out.append(_number(toks[i]))
            i += 1
    return out, i + 1


def _parse_dict(toks, i, close=None):
    d = {}
    while i < len(toks) and toks[i] != close:
        key = toks[i]
        if key == ";":
            i += 1
            continue
        i += 1
        if i < len(toks) and toks[i] == "{":
            v, i = _parse_dict(toks, i + 1, "}")
            d[key.strip('"')] = v
            continue
        vals = []
        while i < len(toks) and toks[i] != ";":
            if toks[i] in ("(", "["):
                v, i = _parse_list(toks, i + 1, ")" if toks[i] == "(" else "]")
                vals.append(v)
            elif toks[i] == "{":
                v, i = _parse_dict(toks, i + 1, "}")
                vals.append(v)
            elif toks[i] == close:
                break
            else:
                vals.append(_number(toks[i]))
                i += 1
        if i < len(toks) and toks[i] == ";":
            i += 1
        d[key.strip('"')] = vals[0] if len(vals) == 1 else vals
    return d, i + 1


def parse_dict(text):
    """OpenFOAM dictionary syntax -> nested dict (entries ``key value...;`` and ``key { ... }``; lists as Python
    lists; a single value unwrapped).  Large numeric lists of fields and meshes are read by the dedicated
    functions below."""
    d, _ = _parse_dict(_tokens(text), 0)
    return d


def read_dict(path):
    with open(path) as f:
        return parse_dict(f.read())


# ---------------------------------------------------------------------------------------------- meshes
class Mesh2D:
    """Triangles of a 2D case: ``points`` [nv][2], ``tris`` [K][3] anticlockwise vertex ids, ``face_patch``
    [K][3] the boundary patch of face k (vertices k, k+1) or -1 for an interior face, ``patches`` the patch
    names in file order."""

    def __init__(self, points, tris, face_patch, patches, face_order=None):
        self.points = np.ascontiguousarray(points, np.float64)
        self.tris = np.ascontiguousarray(tris, np.int32)
        self.face_patch = np.ascontiguousarray(face_patch, np.int32)
        self.patches = list(patches)
        # [K][3] position of each boundary face in the file (the reference's patch face order), -1 inside
        self.face_order = None if face_order is None else np.ascontiguousarray(face_order, np.int64)

    def patch_first(self):
        """[K][3] e' 3 + f' of the first face (in file order) of the patch each boundary face belongs to, -1 for
        interior faces: the face whose first value the reference's Triangle limiter takes as the ghost mean of the
        whole patch (Trianglelimite.C:220-223)"""
        if self.face_order is None:
            raise CaseError("the mesh reader kept no boundary face order")
        out = np.full((self.K, 3), -1, np.int32)
        for p in range(len(self.patches)):
            sel = np.argwhere(self.face_patch == p)
            if len(sel) == 0:
                continue
            e, f = sel[np.argmin(self.face_order[sel[:, 0], sel[:, 1]])]
            out[self.face_patch == p] = 3 * e + f
        return out

    @property
    def K(self):
        return len(self.tris)

    def vertices(self):
        """[K][3][2] vertex coordinates per triangle"""
        return self.points[self.tris]

    def boundary_faces(self, patch):
        """(element, face) pairs of a patch"""
        p = self.patches.index(patch)
        return np.argwhere(self.face_patch == p)


def _orient(points, a, b, c):
    """anticlockwise (a, b, c): swap b and c when the order is clockwise"""
    (xa, ya), (xb, yb), (xc, yc) = points[a], points[b], points[c]
    if (xb - xa) * (yc - ya) - (yb - ya) * (xc - xa) < 0:
        b, c = c, b
    return a, b, c


def _attach_faces(tris, edge_patch, npatch):
    face_patch = np.full((len(tris), 3), -1, np.int32)
    for e, t in enumerate(tris):
        for k in range(3):
            key = (min(t[k], t[(k + 1) % 3]), max(t[k], t[(k + 1) % 3]))
            if key in edge_patch:
                face_patch[e, k] = edge_patch[key]
    return face_patch


_FLUENT_HDR = re.compile(r"\((10|12|13)\s*\(([0-9a-fA-F]+)\s+([0-9a-fA-F]+)\s+([0-9a-fA-F]+)([^)]*)\)\s*\(")
_FLUENT_ZONE = re.compile(r"\((?:45|39)\s*\(\s*(\d+)\s+([\w-]+)\s+([^\s)]+)[^)]*\)")


def read_fluent(path):
    """A 2D ASCII Fluent mesh of triangles: (10) nodes, (13) faces ``n0 n1 c0 c1`` (hexadecimal; c = 0 on the
    boundary; mixed sections carry an element type first), (45)/(39) zone names.  Boundary patches are the face
    zones that are not ``interior``."""
    with open(path) as f:
        text = f.read()
    names = {int(m.group(1)): (m.group(2), m.group(3)) for m in _FLUENT_ZONE.finditer(text)}
    nodes = {}
    faces = []    # (n0, n1, c0, c1, zone)
    for m in _FLUENT_HDR.finditer(text):
        sec, zone, first, last = m.group(1), int(m.group(2), 16), int(m.group(3), 16), int(m.group(4), 16)
        if zone == 0 or sec == "12":
            continue
        body = text[m.end():text.index(")", m.end())]
        toks = body.split()
        count = last - first + 1
        if sec == "10":
            rest = m.group(5).split()
            nd = int(rest[1], 16) if len(rest) >= 2 else 2
            if nd != 2:
                raise CaseError("%s: %d-dimensional nodes; the DG path is 2D" % (path, nd))
            vals = np.array(toks, np.float64)
            if vals.size != 2 * count:
                raise CaseError("%s: node section %d holds %d values, expected %d" % (path, zone, vals.size, 2 * count))
            for i in range(count):
                nodes[first + i] = (vals[2 * i], vals[2 * i + 1])
        else:
            rest = m.group(5).split()
            mixed = len(rest) >= 2 and int(rest[1], 16) == 0
            w = 5 if mixed else 4
            vals = [int(t, 16) for t in toks]
            if len(vals) != w * count:
                raise CaseError("%s: face section %d holds %d values, expected %d" % (path, zone, len(vals), w * count))
            for i in range(count):
                n0, n1, c0, c1 = vals[w * i + w - 4:w * i + w]
                faces.append((n0, n1, c0, c1, zone))
    if not nodes:
        raise CaseError("%s: no node section" % path)
    ids = sorted(nodes)
    index = {v: i for i, v in enumerate(ids)}
    points = np.array([nodes[v] for v in ids])
    cells = {}
    for n0, n1, c0, c1, _ in faces:
        for c in (c0, c1):
            if c:
                cells.setdefault(c, set()).update((index[n0], index[n1]))
    tris = []
    for c in sorted(cells):
        vs = sorted(cells[c])
        if len(vs) != 3:
            raise CaseError("%s: cell %d has %d vertices; the DG path takes triangles" % (path, c, len(vs)))
        tris.append(_orient(points, *vs))
    patch_zones = sorted({z for (_, _, c0, c1, z) in faces if c0 == 0 or c1 == 0})
    patches = [names.get(z, ("boundary", "zone%d" % z))[1] for z in patch_zones]
    edge_patch = {}
    edge_order = {}
    for i, (n0, n1, c0, c1, z) in enumerate(faces):
        if c0 == 0 or c1 == 0:
            a, b = index[n0], index[n1]
            edge_patch[(min(a, b), max(a, b))] = patch_zones.index(z)
            edge_order[(min(a, b), max(a, b))] = i
    tris = np.array(tris, np.int32)
    order = np.full((len(tris), 3), -1, np.int64)
    for e, t in enumerate(tris):
        for k in range(3):
            order[e, k] = edge_order.get((min(t[k], t[(k + 1) % 3]), max(t[k], t[(k + 1) % 3])), -1)
    return Mesh2D(points, tris, _attach_faces(tris, edge_patch, len(patches)), patches, order)


def _foam_body(path):
    """the text of an OpenFOAM list file after its FoamFile header"""
    with open(path) as f:
        text = f.read()
    i = text.find("FoamFile")
    if i >= 0:
        j = text.index("}", i)
        text = text[j + 1:]
    return re.sub(r"//[^\n]*|/\*.*?\*/", " ", text, flags=re.S)


def _read_points(path):
    body = _foam_body(path)
    n = int(re.search(r"\d+", body).group(0))
    vals = np.array(re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", body[body.index("(") + 1:]), np.float64)
    return vals[:3 * n].reshape(n, 3)


def _read_labels(path):
    body = _foam_body(path)
    m = re.search(r"(\d+)\s*\(", body)
    n = int(m.group(1))
    vals = np.array(body[m.end():].replace(")", " ").split()[:n], np.int64)
    return vals


def _read_faces(path):
    body = _foam_body(path)
    m = re.search(r"(\d+)\s*\(", body)
    n = int(m.group(1))
    faces = [[int(v) for v in g.split()] for g in re.findall(r"\d+\s*\(([\d\s]*)\)", body[m.end():])]
    if len(faces) != n:
        raise CaseError("%s: %d faces read, header says %d" % (path, len(faces), n))
    return faces


def read_polymesh(dirpath):
    """A one-layer prism polyMesh (``points``, ``faces``, ``owner``, ``neighbour``, ``boundary``): each cell's
    face with every point at z = 0 is its triangle (``dgPolyMesh.C:105-199``), in that face's point order with
    points 1 and 2 swapped when clockwise (``dgPolyMesh.C:490-509``); boundary patches other than ``empty`` give
    the triangles' boundary faces."""
    pts = _read_points(os.path.join(dirpath, "points"))
    faces = _read_faces(os.path.join(dirpath, "faces"))
    owner = _read_labels(os.path.join(dirpath, "owner"))
    neighbour = _read_labels(os.path.join(dirpath, "neighbour"))
    bnd = parse_dict(_foam_body(os.path.join(dirpath, "boundary")))
    # the boundary file is "n ( name { ... } ... )": parse_dict sees the count as a key holding the list
    blist = next(v for v in bnd.values() if isinstance(v, list))
    entries = []
    it = iter(blist)
    for name, spec in zip(it, it):
        entries.append((str(name), spec))
    ncell = int(max(owner.max(), neighbour.max() if neighbour.size else -1)) + 1
    cell_faces = [[] for _ in range(ncell)]
    for f, c in enumerate(owner):
        cell_faces[c].append(f)
    for f, c in enumerate(neighbour):
        cell_faces[c].append(f)
    points2 = pts[:, :2]
    tris = []
    for c in range(ncell):
        base = [f for f in cell_faces[c] if all(pts[p, 2] == 0 for p in faces[f])]
        if len(base) != 1:
            raise CaseError("%s: cell %d has %d faces at z = 0" % (dirpath, c, len(base)))
        fp = faces[base[0]]
        if len(fp) != 3:
            raise CaseError("%s: cell %d is a %d-gon; the DG path takes triangles" % (dirpath, c, len(fp)))
        tris.append(_orient(points2, fp[0], fp[1], fp[2]))
    patches, edge_patch = [], {}
    for name, spec in entries:
        if spec.get("type") == "empty":
            continue
        p = len(patches)
        patches.append(name)
        start, n = int(spec["startFace"]), int(spec["nFaces"])
        for f in range(start, start + n):
            ab = [q for q in faces[f] if pts[q, 2] == 0]
            if len(ab) != 2:
                raise CaseError("%s: boundary face %d of %s has %d points at z = 0" % (dirpath, f, name, len(ab)))
            edge_patch[(min(ab), max(ab))] = p
    # compact the vertex numbering to the z = 0 points
    used = np.unique(np.array(tris).reshape(-1))
    remap = -np.ones(len(pts), np.int64)
    remap[used] = np.arange(used.size)
    tris = remap[np.array(tris, np.int64)].astype(np.int32)
    edge_patch = {(int(remap[a]), int(remap[b])): p for (a, b), p in edge_patch.items()}
    edge_patch = {(min(a, b), max(a, b)): p for (a, b), p in edge_patch.items()}
    return Mesh2D(points2[used], tris, _attach_faces(tris, edge_patch, len(patches)), patches)


# ---------------------------------------------------------------------------------------------- fields
_FIELD_CLASSES = {"dgScalarField": 1, "dgVectorField": 3}


def _split_entry(text, key):
    """(start, end) of the value of top-level ``key`` in a field file (``key value ;``)"""
    m = re.search(r"(^|\n)\s*%s\s" % re.escape(key), text)
    if not m:
        return None
    return m.end()


def _parse_values(text, pos, ncomp, count_hint):
    """``uniform v`` / ``uniform (x y z)`` / ``nonuniform List<T> n ( ... )`` at text[pos:]; returns (array
    [n][ncomp] or [1][ncomp] for uniform, uniform flag, end position)"""
    m = re.compile(r"\s*(uniform|nonuniform)\s*").match(text, pos)
    if not m:
        raise CaseError("expected uniform or nonuniform")
    pos = m.end()
    num = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
    if m.group(1) == "uniform":
        if ncomp == 1:
            v = re.compile(r"(%s)\s*;" % num).match(text, pos)
            if not v:
                raise CaseError("bad uniform scalar")
            return np.array([[float(v.group(1))]]), True, v.end()
        v = re.compile(r"\(\s*(%s)\s+(%s)\s+(%s)\s*\)\s*;" % (num, num, num)).match(text, pos)
        if not v:
            raise CaseError("bad uniform vector")
        return np.array([[float(v.group(i)) for i in (1, 2, 3)]]), True, v.end()
    h = re.compile(r"List<\w+>\s*(\d+)\s*\(").match(text, pos)
    if not h:
        raise CaseError("bad nonuniform list")
    n = int(h.group(1))
    depth, j = 1, h.end()
    while depth:
        c = text[j]
        depth += c == "("
        depth -= c == ")"
        j += 1
    vals = np.array(re.findall(num, text[h.end():j - 1]), np.float64)
    if vals.size != n * ncomp:
        raise CaseError("list of %d entries holds %d numbers (%d components)" % (n, vals.size, ncomp))
    end = text.index(";", j)
    return vals.reshape(n, ncomp), False, end + 1


def read_dg_field(path, K, Np):
    """A ``dgScalarField`` / ``dgVectorField`` file -> (values [ncomp][K][Np] in the reference's element order,
    boundaryField dict with each patch's ``type`` and, where given, its ``value`` as text)"""
    with open(path) as f:
        text = f.read()
    hdr = parse_dict(text[text.index("FoamFile"):text.index("}", text.index("FoamFile")) + 1])["FoamFile"]
    cls = hdr.get("class")
    if cls not in _FIELD_CLASSES:
        raise CaseError("%s: class %s is not a DG field" % (path, cls))
    ncomp = _FIELD_CLASSES[cls]
    pos = _split_entry(text, "internalField")
    if pos is None:
        raise CaseError("%s: no internalField" % path)
    vals, uniform, end = _parse_values(text, pos, ncomp, K * Np)
    if uniform:
        vals = np.repeat(vals, K * Np, axis=0)
    elif len(vals) != K * Np:
        raise CaseError("%s: %d values, the mesh has K Np = %d" % (path, len(vals), K * Np))
    bpos = _split_entry(text, "boundaryField")
    boundary = {}
    if bpos is not None:
        toks = _tokens(text[bpos:])
        bd, _ = _parse_dict(toks, 1, "}")
        for name, spec in bd.items():
            boundary[name] = {"type": spec.get("type")} if isinstance(spec, dict) else {"type": None}
    out = np.ascontiguousarray(vals.T.reshape(ncomp, K, Np))
    return out, boundary


def _fmt(v):
    return "%.17g" % v   # enough digits to read back every double exactly


def write_dg_field(path, name, values, dimensions="[0 0 0 0 0 0 0]", boundary=None, time_dir=None):
    """Writes values [K][Np] (dgScalarField) or [3][K][Np] / [2][K][Np] (dgVectorField, z = 0 appended) in the
    reference's ASCII format; boundary: {patch: type} written as ``type <t>;`` entries (``value`` for fixedValue
    patches is written uniform 0, the device path takes boundary data separately)."""
    v = np.asarray(values, np.float64)
    if v.ndim == 2:
        cls, ncomp, flat = "dgScalarField", 1, v.reshape(-1)
    elif v.ndim == 3 and v.shape[0] in (2, 3):
        comps = [v[0], v[1], v[2] if v.shape[0] == 3 else np.zeros_like(v[0])]
        cls, ncomp, flat = "dgVectorField", 3, None
    else:
        raise CaseError("values must be [K][Np] or [2|3][K][Np]")
    lines = ["FoamFile", "{", "    version     0.1;", "    format      ascii;", "    class       %s;" % cls]
    if time_dir is not None:
        lines.append('    location    "%s";' % time_dir)
    lines += ["    object      %s;" % name, "}", "", "dimensions      %s;" % dimensions, ""]
    n = v.shape[-2] * v.shape[-1]
    lines.append("internalField   nonuniform List<%s> " % ("scalar" if ncomp == 1 else "vector"))
    lines.append(str(n))
    lines.append("(")
    if ncomp == 1:
        lines += [_fmt(x) for x in flat]
    else:
        a, b, c = (x.reshape(-1) for x in comps)
        lines += ["(%s %s %s)" % (_fmt(a[i]), _fmt(b[i]), _fmt(c[i])) for i in range(n)]
    lines += [")", ";", "", "boundaryField", "{"]
    for patch, t in (boundary or {}).items():
        lines += ["    %s" % patch, "    {", "        type            %s;" % t]
        if t == "fixedValue":
            lines.append("        value           uniform %s;" % ("0" if ncomp == 1 else "(0 0 0)"))
        lines.append("    }")
    lines += ["}", ""]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines))


# ---------------------------------------------------------------------------------------------- cases
# fixedGradient patches evaluate to the internal values (fixedGradientDgPatchField.C evaluate()), so fluxes see them
# as zeroGradient; their gradient data enter the SIPG right-hand side separately (dgh_sipg_rhs_neumann)
_TAGS = {"fixedValue": DGH_BC_FIXED_VALUE, "zeroGradient": DGH_BC_ZERO_GRADIENT, "fixedGradient": DGH_BC_ZERO_GRADIENT,
         "reflective": DGH_BC_WALL, "slip": DGH_BC_WALL}


# ---- scheme words of the case dictionaries -> the C ABI's descriptors ------------------------------------------
# The reference selects its schemes by run-time selection from dgSchemes (dgSchemes.C:485-501) and its solvers from
# dgSolution (dgMatrix.C:199-279).  Its words, and the "hip" words of the plugin that routes them to this library,
# map as follows; anything else raises rather than being approximated.
#   godunovScheme { fluxScheme Roe | hipRoe | hipLF; limiteScheme Triangle | hipTriangle | hipTVB | none;
#                   troubledCell TVB | MLP; }                    (godunovScheme.C:61-71, Trianglelimite.H:74)
#   divSchemes { div(U,T) default LF | default average | hip LF [fp64|fp32|bf16x3] | hip average; }
#                                                                  (LFFlux.H:71, averageFlux.H:71)
#   solvers { <field> { kspSolver cg | hipCG; kspPC jacobi | bjacobi | lu | ilu | DIC | none; tolerance; relTol;
#                       maxIter; } }                             (dgMatrix.C:199-279)
_GODUNOV_FLUX = {"Roe": DGH_FLUX_ROE, "hipRoe": DGH_FLUX_ROE, "hipLF": DGH_FLUX_LF}
_LIMITERS = {"Triangle": DGH_LIM_TRIANGLE, "hipTriangle": DGH_LIM_TRIANGLE, "hipTVB": DGH_LIM_TVB, "none": None}
_SIMPLE_FLUX = {"LF": DGH_FLUX_UPWIND, "average": DGH_FLUX_AVERAGE, "upwind": DGH_FLUX_UPWIND}
_PRECISION = {"fp64": DGH_PREC_FP64, "fp32": DGH_PREC_FP32, "bf16x3": DGH_PREC_BF16X3}
# PETSc preconditioners: point Jacobi stays point Jacobi; the factorisation-type ones (lu, ilu, DIC, bjacobi) become
# the exact element-block Jacobi (the block-diagonal part of the factorisation), a stronger preconditioner than
# point Jacobi and the strongest one without global fill
_KSP_PC = {"jacobi": DGH_PC_JACOBI, "none": DGH_PC_JACOBI, "bjacobi": DGH_PC_BLOCK_JACOBI, "lu": DGH_PC_BLOCK_JACOBI,
           "ilu": DGH_PC_BLOCK_JACOBI, "DIC": DGH_PC_BLOCK_JACOBI}


def _words(v):
    return [str(w) for w in v] if isinstance(v, list) else str(v).split()


class Case:
    """A reference 2D DG case directory: dictionaries, mesh, fields.  ``mesh``: the mesh file or polyMesh
    directory to read, relative to the case (default: ``constant/polyMesh``, else the first ``*.msh``)."""

    def __init__(self, path, mesh=None):
        self.path = os.path.abspath(path)
        self.mesh_name = mesh
        sysdir = os.path.join(self.path, "system")
        self.control = read_dict(os.path.join(sysdir, "controlDict"))
        self.solution = read_dict(os.path.join(sysdir, "dgSolution"))
        sch = os.path.join(sysdir, "dgSchemes")
        self.schemes = read_dict(sch) if os.path.exists(sch) else {}
        dg = self.solution.get("DG", {})
        if int(dg.get("meshDimension", 2)) != 2:
            raise CaseError("%s: meshDimension %s; the DG path is 2D" % (self.path, dg.get("meshDimension")))
        self.N = int(dg["baseOrder"])
        self._mesh = None

    def mesh_file(self):
        """the mesh this case is read from: the one named at construction, else constant/polyMesh, else the
        first *.msh of the case directory"""
        if self.mesh_name is not None:
            return os.path.join(self.path, self.mesh_name)
        pm = os.path.join(self.path, "constant", "polyMesh")
        if os.path.isdir(pm):
            return pm
        msh = sorted(f for f in os.listdir(self.path) if f.endswith(".msh"))
        if not msh:
            raise CaseError("%s: neither constant/polyMesh nor a *.msh file" % self.path)
        return os.path.join(self.path, msh[0])

    def mesh2d(self, mesh_file=None):
        if mesh_file is not None:
            return read_polymesh(mesh_file) if os.path.isdir(mesh_file) else read_fluent(mesh_file)
        if self._mesh is None:
            f = self.mesh_file()
            self._mesh = read_polymesh(f) if os.path.isdir(f) else read_fluent(f)
        return self._mesh

    def times(self):
        """the time directories present, ascending"""
        out = []
        for d in os.listdir(self.path):
            try:
                out.append((float(d), d))
            except ValueError:
                pass
        return [d for _, d in sorted(out)]

    def read_field(self, name, time="0", Np=None):
        m = self.mesh2d()
        Np = Np or (self.N + 1) * (self.N + 2) // 2
        return read_dg_field(os.path.join(self.path, time, name), m.K, Np)

    def write_field(self, time, name, values, dimensions="[0 0 0 0 0 0 0]", boundary=None):
        tdir = time if isinstance(time, str) else "%.6g" % time
        write_dg_field(os.path.join(self.path, tdir, name), name, values, dimensions, boundary, tdir)

    def bc_tags(self, fields, time="0"):
        """[K][3] device face tags from the boundary types the given fields declare per patch.  A patch whose
        momentum is ``reflective``/``slip`` is a wall (its scalars must be ``zeroGradient``, ``reflective`` or
        ``slip``); otherwise every field must agree on ``fixedValue`` or ``zeroGradient``."""
        m = self.mesh2d()
        types = {}
        for name in fields:
            with open(os.path.join(self.path, time, name)) as f:
                text = f.read()
            pos = _split_entry(text, "boundaryField")
            if pos is None:
                raise CaseError("%s/%s: no boundaryField" % (time, name))
            bd, _ = _parse_dict(_tokens(text[pos:]), 1, "}")
            for patch, spec in bd.items():
                types.setdefault(patch, {})[name] = spec.get("type") if isinstance(spec, dict) else None
        tag = np.full((m.K, 3), 0, np.int32)
        for p, patch in enumerate(m.patches):
            declared = {t for t in types.get(patch, {}).values()}
            if not declared:
                raise CaseError("patch %s of the mesh has no boundary condition in %s (their patches: %s); "
                                "patch names are case-sensitive" % (patch, ", ".join(fields), sorted(types)))
            if declared & {"reflective", "slip"}:
                if not declared <= {"reflective", "slip", "zeroGradient"}:
                    raise CaseError("patch %s: %s cannot be combined on the device path" % (patch, sorted(declared)))
                t = DGH_BC_WALL
            elif len(declared) == 1 and next(iter(declared)) in _TAGS:
                t = _TAGS[next(iter(declared))]
            else:
                raise CaseError("patch %s: boundary types %s are not available on the device path" %
                                (patch, sorted(str(d) for d in declared)))
            tag[m.face_patch == p] = t
        return tag

    # ---- scheme dispatch
    def gamma(self):
        """ratio of specific heats from constant/transportProperties (1.4 when absent)"""
        p = os.path.join(self.path, "constant", "transportProperties")
        if not os.path.exists(p):
            return 1.4
        g = read_dict(p).get("gamma", 1.4)
        return float(g[-1] if isinstance(g, list) else g)

    def euler_rhs_desc(self):
        """dgh_rhs_desc of an Euler case: the godunovScheme's fluxScheme word (godunovScheme.C:61-65)"""
        from .dgh import RhsDesc
        god = self.schemes.get("godunovScheme")
        if not isinstance(god, dict) or "fluxScheme" not in god:
            raise CaseError("%s: dgSchemes has no godunovScheme { fluxScheme ...; }" % self.path)
        word = _words(god["fluxScheme"])
        if word[0] not in _GODUNOV_FLUX:
            raise CaseError("fluxScheme %s is not available (known: %s)" % (word[0], sorted(_GODUNOV_FLUX)))
        prec = _PRECISION[word[1]] if len(word) > 1 else DGH_PREC_FP64
        return RhsDesc(DGH_EQ_EULER, _GODUNOV_FLUX[word[0]], prec, 0, 0.0, 0.0, self.gamma())

    def advection_rhs_desc(self, term, ax, ay):
        """dgh_rhs_desc of the convection term named in divSchemes (e.g. "div(U,T)", dgSchemes.C:485-501) with the
        velocity (ax, ay): "default LF" / "hip LF [prec]" -> upwind (the LF flux for a constant velocity),
        "default average" / "hip average" -> the central average flux"""
        from .dgh import RhsDesc
        div = self.schemes.get("divSchemes", {})
        spec = div.get(term, div.get("default"))
        if spec is None:
            raise CaseError("divSchemes has no %s and no default" % term)
        w = _words(spec)
        if len(w) < 2 or w[0] not in ("default", "hip") or w[1] not in _SIMPLE_FLUX:
            raise CaseError("divSchemes %s: '%s' is not available" % (term, " ".join(w)))
        prec = _PRECISION[w[2]] if w[0] == "hip" and len(w) > 2 else DGH_PREC_FP64
        return RhsDesc(DGH_EQ_ADVECT, _SIMPLE_FLUX[w[1]], prec, 0, float(ax), float(ay), 0.0)

    def limiter_desc(self, eps=1e-2):
        """dgh_lim_desc of the godunovScheme's limiteScheme (None for "none"): Triangle -> the reference's
        Triangle limiter (density floor tol = 1e-2, Trianglelimite.C:802-848), with troubledCell TVB / MLP as its
        detection step; hipTVB -> TVB minmod"""
        from .dgh import LimDesc
        god = self.schemes.get("godunovScheme", {})
        word = _words(god.get("limiteScheme", "none"))[0]
        if word not in _LIMITERS:
            raise CaseError("limiteScheme %s is not available (known: %s)" % (word, sorted(_LIMITERS)))
        kind = _LIMITERS[word]
        if kind is None:
            return None
        tc = _words(god.get("troubledCell", "none"))[0]
        if kind == DGH_LIM_TRIANGLE and tc == "TVB":
            kind = DGH_LIM_TRIANGLE_TVB
        elif tc not in ("none", "TVB"):
            raise CaseError("troubledCell %s needs a trained network: pass it to Mesh.limiter_desc(mlp=...)" % tc)
        return LimDesc(kind, 0, 50.0, 1.5, -1, 0, 0.5, self.gamma(), eps)

    def sipg_settings(self, field):
        """keyword arguments of Mesh.sipg for the field's dgSolution solver entry (dgMatrix.C:199-279): kspSolver cg
        (or hipCG), kspPC -> point or element-block Jacobi, relTol, tolerance (PETSc abstol), maxIter"""
        solvers = self.solution.get("solvers", {})
        ent = solvers.get(field)
        if ent is None:
            for k, v in solvers.items():    # "rho(1|2|3)"-style keys are regular expressions
                try:
                    if re.fullmatch(k.strip('"'), field):
                        ent = v
                        break
                except re.error:
                    continue
        if not isinstance(ent, dict):
            raise CaseError("dgSolution has no solver entry for %s" % field)
        ksp = str(ent.get("kspSolver", "gmres"))
        if ksp not in ("cg", "hipCG", "preonly"):
            raise CaseError("kspSolver %s: the device solver is conjugate gradients" % ksp)
        pc = str(ent.get("kspPC", ent.get("kspPc", "jacobi")))
        if pc not in _KSP_PC:
            raise CaseError("kspPC %s is not available (known: %s)" % (pc, sorted(_KSP_PC)))
        lap = _words(self.schemes.get("laplacianSchemes", {}).get("default", "default default 20"))
        if lap[0] not in ("default", "hip"):
            raise CaseError("laplacianSchemes default %s is not available" % " ".join(lap))
        return dict(penalty=DGH_SIPG_PENALTY_REFERENCE, pc=_KSP_PC[pc], rtol=float(ent.get("relTol", 1e-5)),
                    abstol=float(ent.get("tolerance", 1e-50)), maxit=int(ent.get("maxIter", 10000)))

    def device_mesh(self, ctx, fields=None, time="0"):
        """the case's mesh on the device (``Context.mesh``), order ``baseOrder``, with face tags from the
        boundary types of ``fields`` (all faces on periodic or untagged meshes when None)"""
        m = self.mesh2d()
        tag = self.bc_tags(fields, time) if fields else None
        if tag is None and (m.face_patch >= 0).any():
            raise CaseError("the mesh has boundary patches: name the fields whose boundary types tag them")
        return ctx.mesh(self.N, m.vertices(), m.tris, tag)
