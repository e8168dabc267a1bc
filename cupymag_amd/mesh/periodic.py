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
