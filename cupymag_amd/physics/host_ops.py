"""Python side of the native host assembler (cpm_host_assemble).

Assembles every operator of the GSPM step straight into CSR on the CPU, in
parallel over DOF rows, and exposes the arrays as zero-copy numpy views of
the library-owned buffers until ``free()``.
"""
import ctypes as C
import os

import numpy as np

from .._lib import HA, check, lib, ptr

_DT = {"PTR": np.int64, "GS_PTR": np.int64, "F1_PTR": np.int64, "EL_PTR": np.int64, "FE_PTR": np.int64,
       "IDX": np.int32, "GS_IDX": np.int32, "F1_IDX": np.int32, "EL_IDX": np.int32, "FE_IDX": np.int32}


def voigt_C(c11, c12, c44, rot111=False):
    """6x6 cubic stiffness, rotated C' = M C M^T when rot111
    (assemble_elasticity.py:61-98,176-192)."""
    Cm = np.zeros((6, 6))
    Cm[:3, :3] = c12
    Cm[0, 0] = Cm[1, 1] = Cm[2, 2] = c11
    Cm[3, 3] = Cm[4, 4] = Cm[5, 5] = c44
    if rot111:
        from ..utils.rot_111_matrices import get_M_matrix
        M = get_M_matrix()
        tmp = np.zeros((6, 6))
        out = np.zeros((6, 6))
        for i in range(6):
            for j in range(6):
                s = 0.0
                for k in range(6):
                    s += M[i, k] * Cm[k, j]
                tmp[i, j] = s
        for i in range(6):
            for j in range(6):
                s = 0.0
                for k in range(6):
                    s += tmp[i, k] * M[j, k]
                out[i, j] = s
        Cm = out
    return Cm


class HostOps:
    def __init__(self, coords, elements, gid, ndof, defect_dofs, alpha1, alpha2, Cvoigt=None, nthreads=None,
                 anchors=True):
        """anchors=False: A_demag and A_el stay the singular periodic
        stiffness (no identity rows at DOF 0), as the floating multigrid
        levels use them."""
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        elements = np.ascontiguousarray(elements, dtype=np.int32)
        gid = np.ascontiguousarray(gid, dtype=np.int32)
        npe = elements.shape[1] - 1
        mask = np.zeros(ndof, dtype=np.uint8)
        if defect_dofs is not None and len(defect_dofs):
            mask[np.asarray(defect_dofs)] = 1
        coef = np.zeros(38)
        coef[0], coef[1] = alpha1, alpha2
        self.elastic = Cvoigt is not None
        if self.elastic:
            coef[2:] = np.asarray(Cvoigt, dtype=np.float64).ravel()
        nt = nthreads or int(os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)
        h = C.c_void_p()
        check(lib().cpm_host_assemble(ptr(coords), coords.shape[0], ptr(elements), elements.shape[0], npe,
                                      elements.shape[1], ptr(gid), int(ndof), ptr(mask), ptr(coef),
                                      int(self.elastic) | (0 if anchors else 2), int(nt), C.byref(h)),
              "cpm_host_assemble")
        self.h = h
        self.n = int(ndof)

    def array(self, name):
        p = C.c_void_p()
        n = C.c_int64()
        check(lib().cpm_host_ops_array(self.h, HA[name], C.byref(p), C.byref(n)), "cpm_host_ops_array")
        dt = _DT.get(name, np.float64)
        if n.value == 0:
            return np.zeros(0, dtype=dt)
        buf = (C.c_char * (n.value * np.dtype(dt).itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=dt)

    def csr(self, which):
        """(indptr, indices, values, shape) of one operator (views)."""
        n = self.n
        if which in ("A_demag", "Fx", "Fy", "Fz"):
            v = {"A_demag": "ADEMAG", "Fx": "FX", "Fy": "FY", "Fz": "FZ"}[which]
            return self.array("PTR"), self.array("IDX"), self.array(v), (n, n)
        if which in ("A1", "A2"):
            return self.array("GS_PTR"), self.array("GS_IDX"), self.array(which), (n, n)
        if which == "F1":
            return self.array("F1_PTR"), self.array("F1_IDX"), self.array("F1"), (n, n)
        if which == "A_el":
            return self.array("EL_PTR"), self.array("EL_IDX"), self.array("AEL"), (3 * n, 3 * n)
        if which == "F_el":
            return self.array("FE_PTR"), self.array("FE_IDX"), self.array("FEL"), (3 * n, 6 * n)
        raise KeyError(which)

    def scipy(self, which):
        import scipy.sparse as sp
        ip, ix, v, shape = self.csr(which)
        return sp.csr_matrix((v.copy(), ix.copy(), ip.copy()), shape=shape)

    def free(self):
        if getattr(self, "h", None):
            lib().cpm_host_ops_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def element_matrices(coords, elements, Cvoigt=None, want=("K", "M", "F")):
    """Dense per-element matrices from the native element integrals."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    elements = np.ascontiguousarray(elements, dtype=np.int32)
    ne, npe = elements.shape[0], elements.shape[1] - 1
    out = {}
    K = np.zeros((ne, npe, npe)) if "K" in want else None
    M = np.zeros((ne, npe, npe)) if "M" in want else None
    F = np.zeros((3, ne, npe, npe)) if "F" in want else None
    Kel = np.zeros((ne, 3 * npe, 3 * npe)) if "Kel" in want else None
    Fel = np.zeros((ne, 3 * npe, 6 * npe)) if "Fel" in want else None
    Cv = None if Cvoigt is None else np.ascontiguousarray(Cvoigt, dtype=np.float64)
    check(lib().cpm_host_element_matrices(ptr(coords), coords.shape[0], ptr(elements), ne, npe, elements.shape[1],
                                          ptr(Cv), ptr(K), ptr(M), ptr(F), ptr(Kel), ptr(Fel)),
          "cpm_host_element_matrices")
    for k, v in (("K", K), ("M", M), ("F", F), ("Kel", Kel), ("Fel", Fel)):
        if v is not None:
            out[k] = v
    return out
