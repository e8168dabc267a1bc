"""Periodic node -> DOF map.

Same numbering as the reference's dictionary loop (gridHex.py:179-213,
gridTet.py:180-225): a coordinate within ``tol`` of the upper bound of its
axis is wrapped to the lower bound, keys are the wrapped coordinates rounded
to 12 decimals, and DOF ids are handed out in order of first appearance over
the nodes.  Vectorised with ``np.unique`` so that it scales to 2M+ nodes.
"""
import numpy as np


def periodic_node_map(coords, lower, upper, tol=1e-12):
    coords = np.asarray(coords, dtype=np.float64)
    w = coords.copy()
    for d in range(3):
        w[np.abs(coords[:, d] - upper[d]) < tol, d] = float(lower[d])
    key = np.round(w, 12) + 0.0          # + 0.0 folds -0.0 into 0.0
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(order), dtype=np.int64)
    rank[order] = np.arange(len(order))
    return rank[inv.ravel()].astype(np.int32)


def structured_dims(coords, gid):
    """(nx, ny, nz) when the periodic DOFs form an nx x ny x nz grid numbered
    (i*ny + j)*nz + k (the numbering HexGrid and the Kuhn generator give),
    else None.  Verified on every DOF's first node, so jittered interior
    nodes are fine as long as they stay within half a cell."""
    coords = np.asarray(coords, dtype=np.float64)
    gid = np.asarray(gid)
    n = int(gid.max()) + 1
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    L = hi - lo
    if np.any(L <= 0):
        return None
    h = (np.prod(L) / n) ** (1.0 / 3.0)
    dims = np.rint(L / h).astype(np.int64)
    if np.prod(dims) != n or np.any(dims < 1):
        return None
    first = np.full(n, len(gid), dtype=np.int64)
    np.minimum.at(first, gid, np.arange(len(gid)))
    ijk = np.rint((coords[first] - lo) / (L / dims)).astype(np.int64)
    if np.any(ijk < 0) or np.any(ijk >= dims[None, :]):
        return None
    expect = (ijk[:, 0] * dims[1] + ijk[:, 1]) * dims[2] + ijk[:, 2]
    if not np.array_equal(expect, np.arange(n)):
        return None
    return tuple(int(d) for d in dims)
