"""Tetrahedral mesh from a NASTRAN .nas file (cupymag/mesh/gridTet.py:72-225).

Reader rules of the reference: lines starting with '$' and blank lines are
skipped, reading stops at ENDDATA, GRID and CTETRA cards are comma-separated
free-field (``GRID,ID,CP,X,Y,Z`` / ``CTETRA,EID,PID,N1,N2,N3,N4``), and node ids
are remapped to row indices through argsort/searchsorted.  Defect flag =
``shape_fn(centroid - centre)`` with the centre defaulting to the midpoint of
the node bounding box.  The periodic map wraps at the bounding box.
"""
from io import StringIO

import numpy as np

from .periodic import periodic_node_map


def read_nas(path):
    grids, tets = [], []
    with open(path, "r") as fh:
        for raw in fh:
            if raw.startswith("$") or raw.isspace():
                continue
            if raw.lstrip().startswith("ENDDATA"):
                break
            line = raw.rstrip()
            if line.startswith("GRID"):
                grids.append(line)
            elif line.startswith("CTETRA"):
                tets.append(line)
    if not grids or not tets:
        raise ValueError("File contains no GRID or CTETRA cards")
    import pandas as pd
    opts = dict(header=None, engine="c", sep=",", skipinitialspace=True, dtype=str)
    g = pd.read_csv(StringIO("\n".join(grids)), names=["GRID", "ID", "CP", "X", "Y", "Z"], **opts)
    t = pd.read_csv(StringIO("\n".join(tets)), names=["CTETRA", "EID", "PID", "N1", "N2", "N3", "N4"], **opts)
    node_ids = g["ID"].astype(np.int32).to_numpy()
    xyz = g[["X", "Y", "Z"]].astype(np.float64).to_numpy()
    en = t[["N1", "N2", "N3", "N4"]].astype(np.int32).to_numpy()
    order = np.argsort(node_ids)
    idx = order[np.searchsorted(node_ids[order], en, side="left")]
    return node_ids, xyz, t["EID"].astype(np.int32).to_numpy(), idx.astype(np.int32)


class TetraGrid:
    def __init__(self, nas_filename=None, defect_shape=None, defect_center=None, arrays=None):
        """Read ``nas_filename``, or take ``arrays = (coords, tets)`` already
        in memory (the same arrays the reader returns)."""
        self.nas_filename = nas_filename
        self.defect_shape = defect_shape
        if arrays is None:
            _, self.node_coords, _, self.element_nodes = read_nas(nas_filename)
        else:
            self.node_coords = np.ascontiguousarray(arrays[0], dtype=np.float64)
            self.element_nodes = np.ascontiguousarray(arrays[1], dtype=np.int32)
        if defect_center is None:
            defect_center = (self.node_coords.min(axis=0) + self.node_coords.max(axis=0)) / 2
        self.defect_center = np.asarray(defect_center, dtype=np.float64)

    def std_fem_mesh(self):
        ne = len(self.element_nodes)
        el = np.zeros((ne, 5), dtype=np.int32)
        el[:, :4] = self.element_nodes
        centers = self.node_coords[el[:, :4]].mean(axis=1)
        el[:, 4] = self.defect_shape(centers - self.defect_center).astype(np.int32)
        return self.node_coords, el

    def build_periodic_node_map(self, node_coords, tol=1e-12):
        lo = node_coords.min(axis=0)
        hi = node_coords.max(axis=0)
        return periodic_node_map(node_coords, lo, hi, tol)
