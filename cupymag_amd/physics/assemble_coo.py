"""The reference's COO assembler classes (cupymag/physics/assemble_demag.py,
assemble_Gauss_Seidel.py, assemble_elasticity.py) on top of the native
element integrals.  Same constructor arguments, method names and COO layout
(row = global_id[corner a] (+ d*nNodeGlobal), col = global_id[corner b]
(+ db*nNodeGlobal, or 6*gid + v for F_el)), same anchors.  They materialise
npe^2 triplets per element like the reference; the simulation itself uses
the direct-to-CSR assembler in host_ops.py instead.
"""
import numpy as np

from .host_ops import element_matrices, voigt_C


def _scalar_coo(elements, gid, Ke):
    npe = elements.shape[1] - 1
    g = gid[elements[:, :npe]].astype(np.int32)
    rows = np.repeat(g[:, :, None], npe, axis=2).ravel()
    cols = np.repeat(g[:, None, :], npe, axis=1).ravel()
    return rows, cols, np.ascontiguousarray(Ke).ravel().copy()


def _anchor(rows, cols, vals, anchors):
    vals = vals.copy()
    for a in anchors:
        vals[(rows == a) | (cols == a)] = 0.0
        rows = np.append(rows, np.int32(a))
        cols = np.append(cols, np.int32(a))
        vals = np.append(vals, 1.0)
    return rows, cols, vals


class AssembleDemag:
    def __init__(self, node_coords, elements, global_id=None):
        self.node_coords, self.elements, self.global_id = node_coords, elements, global_id
        self.anchor_row = 0

    def build_coo_matrix_A_numba(self):
        K = element_matrices(self.node_coords, self.elements, want=("K",))["K"]
        return _anchor(*_scalar_coo(self.elements, self.global_id, K), [self.anchor_row])

    def build_coo_matrices_F_numba(self):
        F = element_matrices(self.node_coords, self.elements, want=("F",))["F"]
        out = []
        for k in range(3):
            out.extend(_scalar_coo(self.elements, self.global_id, F[k]))
        return tuple(out)


class AssembleGaussSeidel:
    def __init__(self, node_coords, elements, global_id):
        self.node_coords, self.elements, self.global_id = node_coords, elements, global_id

    def build_coo_matrices_numba(self):
        e = element_matrices(self.node_coords, self.elements, want=("K", "M"))
        r, c, vK = _scalar_coo(self.elements, self.global_id, e["K"])
        _, _, vM = _scalar_coo(self.elements, self.global_id, e["M"])
        return r, c, vK, r.copy(), c.copy(), vM


class AssembleElasticity:
    def __init__(self, node_coords, elements, global_id, c11, c12, c44, lambda100, lambda111, rot111=False):
        self.node_coords, self.elements, self.global_id = node_coords, elements, global_id
        self.nNodeGlobal = int(global_id.max()) + 1
        self.C_voigt = voigt_C(c11, c12, c44, rot111)
        self.anchor_row = 0

    def _rowd(self):
        npe = self.elements.shape[1] - 1
        g = self.global_id[self.elements[:, :npe]].astype(np.int64)
        return g, (g[:, :, None] + self.nNodeGlobal * np.arange(3)[None, None, :]).reshape(len(g), -1)

    def build_coo_matrix_A_numba(self):
        Ke = element_matrices(self.node_coords, self.elements, self.C_voigt, want=("Kel",))["Kel"]
        _, rowd = self._rowd()
        w = rowd.shape[1]
        rows = np.repeat(rowd[:, :, None], w, axis=2).ravel().astype(np.int32)
        cols = np.repeat(rowd[:, None, :], w, axis=1).ravel().astype(np.int32)
        n = self.nNodeGlobal
        return _anchor(rows, cols, Ke.ravel().copy(), [0, n, 2 * n])

    def build_coo_matrices_F_numba(self):
        Fe = element_matrices(self.node_coords, self.elements, self.C_voigt, want=("Fel",))["Fel"]
        g, rowd = self._rowd()
        cold = (6 * g[:, :, None] + np.arange(6)[None, None, :]).reshape(len(g), -1)
        rows = np.repeat(rowd[:, :, None], cold.shape[1], axis=2).ravel().astype(np.int32)
        cols = np.repeat(cold[:, None, :], rowd.shape[1], axis=1).ravel().astype(np.int32)
        vals = Fe.ravel().copy()
        n = self.nNodeGlobal
        vals[np.isin(rows, [0, n, 2 * n])] = 0.0
        return rows, cols, vals
