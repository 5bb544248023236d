"""Convert the reference's isentropic-vortex Fluent meshes into small fixtures.

Source: HopeFOAM-0.1/tutorials/DG/2D/isentropicVortex/vortex{0256,1024,4096}.msh
(ASCII Fluent, section ids and counts in hexadecimal: (10 ...) nodes,
(12 ...) cells, (13 ...) faces "n0 n1 c0 c1", c = 0 on the boundary).
Each fixture vortexNNNN.txt holds
    nverts ntris
    x y              (nverts lines)
    v0 v1 v2         (ntris lines, 0-based, anticlockwise)
Triangles are rebuilt from their three faces and oriented anticlockwise, as
dgPolyMesh.C:490-509 does; the vertex a triangle starts from does not change
the DG solution (the warp & blend node set is invariant under the triangle's
symmetries).  Run only where the reference is present; outputs are committed.
"""
import os
import re
import sys

SRC = "/root/reference/HopeFOAM-0.1/tutorials/DG/2D/isentropicVortex"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_fluent(text):
    nodes = []
    faces = []
    pos = 0
    hdr = re.compile(r"\((10|13)\s*\(([0-9a-fA-F]+)\s+([0-9a-fA-F]+)\s+([0-9a-fA-F]+)([^)]*)\)\s*\(")
    for m in hdr.finditer(text):
        sec, zone, first, last = m.group(1), int(m.group(2), 16), int(m.group(3), 16), int(m.group(4), 16)
        if zone == 0:
            continue
        body = text[m.end():text.index(")", m.end())]
        toks = body.split()
        count = last - first + 1
        if sec == "10":
            vals = [float(t) for t in toks]
            assert len(vals) == 2 * count
            while len(nodes) < last:
                nodes.append(None)
            for i in range(count):
                nodes[first - 1 + i] = (vals[2 * i], vals[2 * i + 1])
        else:
            vals = [int(t, 16) for t in toks]
            rest = m.group(5).split()
            mixed = len(rest) >= 2 and int(rest[1], 16) == 0   # element type 0: "2 n0 n1 c0 c1"
            w = 5 if mixed else 4
            assert len(vals) == w * count, (len(vals), count)
            for i in range(count):
                faces.append(tuple(vals[w * i + w - 4:w * i + w]))
    return nodes, faces


def build(nodes, faces):
    cells = {}
    for n0, n1, c0, c1 in faces:
        for c in (c0, c1):
            if c:
                cells.setdefault(c, set()).update((n0, n1))
    tris = []
    for c in sorted(cells):
        vs = sorted(cells[c])
        assert len(vs) == 3, (c, vs)
        a, b, d = [v - 1 for v in vs]
        (xa, ya), (xb, yb), (xd, yd) = nodes[a], nodes[b], nodes[d]
        if (xb - xa) * (yd - ya) - (yb - ya) * (xd - xa) < 0:
            b, d = d, b
        tris.append((a, b, d))
    return tris


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    for name in ("vortex0256", "vortex1024", "vortex4096"):
        text = open(os.path.join(src, name + ".msh")).read()
        nodes, faces = parse_fluent(text)
        tris = build(nodes, faces)
        with open(os.path.join(HERE, name + ".txt"), "w") as f:
            f.write("%d %d\n" % (len(nodes), len(tris)))
            for x, y in nodes:
                f.write("%.17g %.17g\n" % (x, y))
            for t in tris:
                f.write("%d %d %d\n" % t)
        print(name, len(nodes), len(tris))


if __name__ == "__main__":
    main()
