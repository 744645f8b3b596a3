"""Triangle meshes for fix mesh/surface: seeded-free generators of the
geometries the workloads use, and an ASCII STL writer/reader in the format
fix mesh/surface file reads (input_mesh_tri.cpp:318: facet / outer loop /
three vertex lines / endloop / endfacet).

Every generator returns an (ntri, 3, 3) float64 array of node coordinates,
node order as written to the STL file.
"""
import math

import numpy as np


def grid(x0, x1, y0, y1, z, nx, ny, diag="alt"):
    """Flat rectangle at height z split into nx*ny quads, two triangles each."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    tris = []
    for i in range(nx):
        for j in range(ny):
            a = (xs[i], ys[j], z)
            b = (xs[i + 1], ys[j], z)
            c = (xs[i + 1], ys[j + 1], z)
            d = (xs[i], ys[j + 1], z)
            if diag == "alt" and (i + j) % 2:
                tris += [(a, b, d), (b, c, d)]
            else:
                tris += [(a, b, c), (a, c, d)]
    return np.array(tris, dtype=np.float64)


def cube(lo, hi):
    """Closed axis-aligned box, 12 triangles, outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
         (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (1, 2, 6), (1, 6, 5), (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]
    return np.array([[v[a], v[b], v[c]] for a, b, c in f], dtype=np.float64)


def drum(radius=0.15, length=0.1, n_az=64, n_ax=16, center=(0.0, 0.0, 0.0)):
    """Closed cylinder with its axis along y: n_az x n_ax quads on the mantle,
    each split into two triangles, plus fan-triangulated end caps
    (2*n_az*n_ax + 2*n_az triangles; 2,176 for 64 x 16)."""
    cx, cy, cz = center
    th = [2.0 * math.pi * k / n_az for k in range(n_az)]
    ys = [cy - 0.5 * length + length * j / n_ax for j in range(n_ax + 1)]
    ring = [(cx + radius * math.cos(t), cz + radius * math.sin(t)) for t in th]
    tris = []
    for k in range(n_az):
        k1 = (k + 1) % n_az
        for j in range(n_ax):
            a = (ring[k][0], ys[j], ring[k][1])
            b = (ring[k1][0], ys[j], ring[k1][1])
            c = (ring[k1][0], ys[j + 1], ring[k1][1])
            d = (ring[k][0], ys[j + 1], ring[k][1])
            tris += [(a, b, c), (a, c, d)]
    for y in (ys[0], ys[-1]):
        ctr = (cx, y, cz)
        for k in range(n_az):
            k1 = (k + 1) % n_az
            tris.append((ctr, (ring[k][0], y, ring[k][1]), (ring[k1][0], y, ring[k1][1])))
    return np.array(tris, dtype=np.float64)


def write_stl(path, tris, name="mesh"):
    tris = np.asarray(tris, dtype=np.float64)
    with open(path, "w") as f:
        f.write(f"solid {name}\n")
        for t in tris.tolist():
            n = np.cross(np.subtract(t[1], t[0]), np.subtract(t[2], t[0]))
            nn = np.linalg.norm(n)
            n = (n / nn) if nn > 0 else n
            f.write(f"  facet normal {float(n[0])!r} {float(n[1])!r} {float(n[2])!r}\n")
            f.write("    outer loop\n")
            for p in t:
                f.write(f"      vertex {p[0]!r} {p[1]!r} {p[2]!r}\n")
            f.write("    endloop\n  endfacet\n")
        f.write(f"endsolid {name}\n")


def read_stl(path):
    verts = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if tok and tok[0] == "vertex":
                verts.append([float(tok[1]), float(tok[2]), float(tok[3])])
    return np.array(verts, dtype=np.float64).reshape(-1, 3, 3)
