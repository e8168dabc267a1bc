"""Kuhn tetrahedral meshes of a periodic box and their .nas files.

The benchmark meshes are n^3 cubes of edge h, each split into the 6 Kuhn
tetrahedra around the (0,0,0)-(1,1,1) diagonal: for every permutation
(a,b,c) of the axes the tet (v0, v0+e_a, v0+e_a+e_b, v0+e_a+e_b+e_c),
ordered so that the signed volume is positive (the reference assemblers use
the signed Jacobian).  Nodes are numbered i-major with k fastest, which makes
the periodic DOF numbering of the reference equal to i*n*n + j*n + k.
Optional seeded jitter moves nodes strictly inside the box only, so periodic
images stay identical.  The tetrahedra of a cube stay positively oriented for
jitter up to 0.35h (the smallest vertex-face distance of a Kuhn tet is
h/sqrt(2)).
"""
import itertools

import numpy as np


def kuhn_mesh(n, h=1.0, jitter=0.0, seed=0, shape=None):
    nx, ny, nz = (n, n, n) if shape is None else shape
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coords = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * h
    if jitter > 0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=coords.shape) * h
        inner = ((i.ravel() > 0) & (i.ravel() < nx) & (j.ravel() > 0) & (j.ravel() < ny)
                 & (k.ravel() > 0) & (k.ravel() < nz))
        coords[inner] += d[inner]

    def nid(a, b, c):
        return (a * (ny + 1) + b) * (nz + 1) + c

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    tets = []
    for perm in itertools.permutations(range(3)):
        off = np.zeros((4, 3), dtype=np.int64)
        for s in range(3):
            off[s + 1] = off[s]
            off[s + 1, perm[s]] += 1
        t = np.stack([nid(ci + o[0], cj + o[1], ck + o[2]) for o in off], axis=1)
        sign = np.linalg.det(off[1:] - off[0])
        if sign < 0:
            t = t[:, [0, 1, 3, 2]]
        tets.append(t)
    tets = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return coords, tets


def write_nas(path, coords, tets, float_fmt="{:.17g}"):
    """Free-field NASTRAN bulk data the reference reader accepts; node and
    element ids start at 1."""
    with open(path, "w") as fh:
        fh.write("$ Kuhn tetrahedral mesh\nBEGIN BULK\n")
        for a, (x, y, z) in enumerate(coords):
            fh.write(f"GRID,{a + 1},,{float_fmt.format(x)},{float_fmt.format(y)},{float_fmt.format(z)}\n")
        for e, t in enumerate(tets):
            fh.write(f"CTETRA,{e + 1},1,{t[0] + 1},{t[1] + 1},{t[2] + 1},{t[3] + 1}\n")
        fh.write("ENDDATA\n")
