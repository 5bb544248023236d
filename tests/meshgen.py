"""Mesh inputs for the tests: structured split-quad meshes and the reference's vortex meshes.

A mesh is given as per-element vertex coordinates vxy [K,3,2] (anticlockwise,
unwrapped across periodic boundaries) plus vertex ids vid [K,3] used for
connectivity (periodic images share one id) and optional bctag [K,3].
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def split_quads(nx, ny, x0, x1, y0, y1, periodic=(True, True)):
    """nx x ny quads, each split along the (i,j)-(i+1,j+1) diagonal into two anticlockwise
    triangles; element 2*(j*nx+i) is the lower-right one."""
    px, py = periodic
    hx = (x1 - x0) / nx
    hy = (y1 - y0) / ny
    nvx = nx if px else nx + 1
    nvy = ny if py else ny + 1

    def vid(i, j):
        return (j % nvy if py else j) * nvx + (i % nvx if px else i)

    K = 2 * nx * ny
    vxy = np.zeros((K, 3, 2))
    ids = np.zeros((K, 3), np.int32)
    e = 0
    for j in range(ny):
        for i in range(nx):
            c = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            for tri in ((0, 1, 2), (0, 2, 3)):
                for k, t in enumerate(tri):
                    ci, cj = c[t]
                    vxy[e, k] = (x0 + ci * hx, y0 + cj * hy)
                    ids[e, k] = vid(ci, cj)
                e += 1
    return vxy, ids


def load_fixture_mesh(name):
    with open(os.path.join(GOLDEN, name + ".txt")) as f:
        nv, nt = [int(t) for t in f.readline().split()]
        xy = np.array([[float(t) for t in f.readline().split()] for _ in range(nv)])
        tri = np.array([[int(t) for t in f.readline().split()] for _ in range(nt)], np.int32)
    return xy[tri], tri
