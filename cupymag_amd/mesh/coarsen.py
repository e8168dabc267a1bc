"""Coarse Kuhn meshes for the multigrid hierarchy.

A Kuhn mesh of n^3 cells (``kuhn_mesh``: node (i,j,k) at index
(i*(n+1) + j)*(n+1) + k, six tetrahedra per cube around the (0,0,0)-(1,1,1)
diagonal) refines the Kuhn mesh of its even nodes, (n/2)^3 cells: every
coarse tetrahedron is the union of eight fine ones, so P1 functions of the
coarse mesh are P1 functions of the fine mesh.  The fine node 2I + d
(d in {0,1}^3) is coarse node I (d = 0) or the midpoint of the coarse edge
from I to I + d.
"""
import numpy as np
import scipy.sparse as sp

from .kuhn import kuhn_mesh


def kuhn_cells(coords, elements):
    """n when (coords, elements) is an n^3 Kuhn mesh in kuhn_mesh order
    (node positions within 0.45 of a cell of the regular grid), else None."""
    coords = np.asarray(coords, dtype=np.float64)
    N = len(coords)
    n = int(round(N ** (1.0 / 3.0))) - 1
    if n < 1 or (n + 1) ** 3 != N:
        return None
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    h = (hi - lo) / n
    if np.any(h <= 0):
        return None
    i, j, k = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    grid = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    if np.max(np.abs((coords - lo) / h - grid)) > 0.45:
        return None
    tets = np.asarray(elements)[:, :4]
    if tets.shape[0] != 6 * n ** 3 or not np.array_equal(tets, kuhn_mesh(n)[1]):
        return None
    return n


def coarse_kuhn(coords, n):
    """(coords, tets) of the Kuhn mesh of the even nodes of an n^3 Kuhn mesh
    (n even)."""
    if n % 2:
        raise ValueError("the Kuhn mesh must have an even number of cells per dimension")
    C = np.asarray(coords, dtype=np.float64).reshape(n + 1, n + 1, n + 1, 3)[::2, ::2, ::2].reshape(-1, 3)
    return np.ascontiguousarray(C), kuhn_mesh(n // 2)[1]


def hierarchy_sizes(n, coarsest_max=8):
    """Cells per dimension of each level: halve while even, down to <= coarsest_max."""
    sizes = [n]
    while sizes[-1] > coarsest_max and sizes[-1] % 2 == 0:
        sizes.append(sizes[-1] // 2)
    return sizes


def prolongation(dims_f, bs=1, anchor=True):
    """P (fine x coarse, scipy CSR) of the nested P1 interpolation on the
    periodic DOF grids dims_f and dims_f / 2 (DOF (i*ny + j)*nz + k,
    block size bs, node-major interleaved).  anchor: node 0 is left out of
    both grids (its row and column are zero)."""
    nf = np.asarray(dims_f)
    nc = nf // 2
    rows, cols, vals = [], [], []
    I, J, K = np.meshgrid(np.arange(nf[0]), np.arange(nf[1]), np.arange(nf[2]), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    f = (I * nf[1] + J) * nf[2] + K
    d = np.stack([I & 1, J & 1, K & 1], axis=1)
    base = np.stack([I >> 1, J >> 1, K >> 1], axis=1)

    def cid(a):
        a = a % nc
        return (a[:, 0] * nc[1] + a[:, 1]) * nc[2] + a[:, 2]

    c0 = cid(base)
    even = ~d.any(axis=1)
    rows.append(f[even]), cols.append(c0[even]), vals.append(np.ones(even.sum()))
    c1 = cid(base + d)
    odd = ~even
    rows += [f[odd], f[odd]]
    cols += [c0[odd], c1[odd]]
    vals += [np.full(odd.sum(), 0.5), np.full(odd.sum(), 0.5)]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    if anchor:
        keep = (r != 0) & (c != 0)
        r, c, v = r[keep], c[keep], v[keep]
    Nf, Nc = int(np.prod(nf)), int(np.prod(nc))
    P = sp.csr_matrix((v, (r, c)), shape=(Nf, Nc))
    if bs > 1:
        P = sp.kron(P, sp.identity(bs), format="csr")
    return P
