"""Coarse Kuhn meshes for the multigrid hierarchy.

A Kuhn mesh of n^3 cells (``kuhn_mesh``: node (i,j,k) at index
(i*(n+1) + j)*(n+1) + k, six tetrahedra per cube around the (0,0,0)-(1,1,1)
diagonal) has a coarse mesh of ceil(n/2)^3 cells.  For even n it is the
Kuhn mesh of the even nodes, which the fine mesh refines: every coarse
tetrahedron is the union of eight fine ones, so P1 functions of the coarse
mesh are P1 functions of the fine mesh, and the fine node 2I + d
(d in {0,1}^3) is coarse node I (d = 0) or the midpoint of the coarse edge
from I to I + d.  For odd n it is the regular Kuhn mesh of (n+1)/2 cells of
the same periodic box (not nested); the transfer is the P1 interpolation of
the coarse mesh at the fine nodes' nominal positions.
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
    """(coords, tets) of the coarse mesh of an n^3 Kuhn mesh: the Kuhn mesh
    of its even nodes (n even, node positions kept) or the regular Kuhn mesh
    of (n+1)/2 cells spanning the same box (n odd)."""
    coords = np.asarray(coords, dtype=np.float64)
    if n % 2 == 0:
        C = coords.reshape(n + 1, n + 1, n + 1, 3)[::2, ::2, ::2].reshape(-1, 3)
        return np.ascontiguousarray(C), kuhn_mesh(n // 2)[1]
    m = (n + 1) // 2
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    grid, tets = kuhn_mesh(m)
    return np.ascontiguousarray(lo + grid * ((hi - lo) / m)), tets


def hierarchy_sizes(n, coarsest_max=8):
    """Cells per dimension of each level: halve (rounding up) down to <= coarsest_max."""
    sizes = [n]
    while sizes[-1] > coarsest_max:
        sizes.append((sizes[-1] + 1) // 2)
    return sizes


def kuhn_weights(idx, nf, nc):
    """P1 interpolation of the periodic coarse Kuhn grid (nc cells per
    dimension) at fine nodes idx (m x 3 integer indices of a grid of nf
    cells, nominal position idx*nc/nf in coarse cells): (vertices m x 4 x 3,
    weights m x 4).  Cell c and fractions t, sorted t_p0 >= t_p1 >= t_p2:
    vertices c, c + e_p0, c + e_p0 + e_p1, c + (1,1,1), weights 1 - t_p0,
    t_p0 - t_p1, t_p1 - t_p2, t_p2 (barycentric coordinates in the Kuhn
    tetrahedron holding the point)."""
    idx = np.asarray(idx, dtype=np.int64)
    nf = np.asarray(nf, dtype=np.int64)
    nc = np.asarray(nc, dtype=np.int64)
    q = idx * nc
    cell = q // nf
    t = (q - cell * nf) / nf
    m = len(idx)
    order = np.argsort(-t, axis=1, kind="stable")
    ts = np.take_along_axis(t, order, axis=1)
    V = np.empty((m, 4, 3), dtype=np.int64)
    W = np.empty((m, 4))
    o = np.zeros((m, 3), dtype=np.int64)
    rows = np.arange(m)
    V[:, 0] = cell
    W[:, 0] = 1.0 - ts[:, 0]
    o[rows, order[:, 0]] = 1
    V[:, 1] = cell + o
    W[:, 1] = ts[:, 0] - ts[:, 1]
    o[rows, order[:, 1]] = 1
    V[:, 2] = cell + o
    W[:, 2] = ts[:, 1] - ts[:, 2]
    V[:, 3] = cell + 1
    W[:, 3] = ts[:, 2]
    return V % nc, W


def prolongation(dims_f, bs=1, anchor=True, dims_c=None):
    """P (fine x coarse, scipy CSR) of the P1 interpolation from the periodic
    DOF grid dims_c (default ceil(dims_f / 2)) to dims_f (DOF
    (i*ny + j)*nz + k, block size bs, node-major interleaved).  anchor: node
    0 is left out of both grids (its row and column are zero)."""
    nf = np.asarray(dims_f, dtype=np.int64)
    nc = (nf + 1) // 2 if dims_c is None else np.asarray(dims_c, dtype=np.int64)
    I, J, K = np.meshgrid(np.arange(nf[0]), np.arange(nf[1]), np.arange(nf[2]), indexing="ij")
    idx = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    f = (idx[:, 0] * nf[1] + idx[:, 1]) * nf[2] + idx[:, 2]
    V, W = kuhn_weights(idx, nf, nc)
    c = (V[..., 0] * nc[1] + V[..., 1]) * nc[2] + V[..., 2]
    r = np.repeat(f, 4)
    c = c.ravel()
    v = W.ravel()
    keep = v != 0.0
    if anchor:
        keep &= (r != 0) & (c != 0)
    r, c, v = r[keep], c[keep], v[keep]
    Nf, Nc = int(np.prod(nf)), int(np.prod(nc))
    P = sp.csr_matrix((v, (r, c)), shape=(Nf, Nc))
    if bs > 1:
        P = sp.kron(P, sp.identity(bs), format="csr")
    return P
