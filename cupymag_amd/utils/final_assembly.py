"""Final assembly helpers (cupymag/utils/final_assembly.py:27-266).

``extract_defect_dofs``, the CSR builders with defect surgery, and
``build_E0_from_m``.  Operators are device-resident ``DeviceCSR``; the COO
-> CSR summation and the row/column surgery run on the host (SciPy), the
spontaneous strain on the device (cpm_build_e0).
"""
import numpy as np


def extract_defect_dofs(elements, global_id):
    elements = np.asarray(elements)
    mask = elements[:, -1] == 1
    if not mask.any():
        return np.zeros(0, dtype=np.int32)
    nodes = np.unique(elements[mask, :-1].ravel())
    return np.unique(np.asarray(global_id)[nodes]).astype(np.int32)


def _surgery_A(A, defect_dofs):
    from scipy.sparse import coo_matrix
    if defect_dofs is None or len(defect_dofs) == 0:
        return A
    c = A.tocoo()
    isdef = np.zeros(A.shape[0], dtype=bool)
    isdef[np.asarray(defect_dofs)] = True
    keep = ~(isdef[c.row] | isdef[c.col])
    d = np.asarray(defect_dofs)
    return coo_matrix((np.concatenate([c.data[keep], np.ones(len(d))]),
                       (np.concatenate([c.row[keep], d]), np.concatenate([c.col[keep], d]))),
                      shape=A.shape).tocsr()


def _surgery_F(F, defect_dofs):
    from scipy.sparse import coo_matrix
    if defect_dofs is None or len(defect_dofs) == 0:
        return F
    c = F.tocoo()
    isdef = np.zeros(F.shape[0], dtype=bool)
    isdef[np.asarray(defect_dofs)] = True
    keep = ~isdef[c.row]
    return coo_matrix((c.data[keep], (c.row[keep], c.col[keep])), shape=F.shape).tocsr()


def enforce_defect_region_A(A, defect_dofs=None):
    from ..device import DeviceCSR
    if defect_dofs is None or len(defect_dofs) == 0:
        return A
    host = A.to_scipy() if isinstance(A, DeviceCSR) else A
    out = _surgery_A(host, defect_dofs)
    return DeviceCSR.from_scipy(out) if isinstance(A, DeviceCSR) else out


def enforce_defect_region_F(F, defect_dofs=None):
    from ..device import DeviceCSR
    if defect_dofs is None or len(defect_dofs) == 0:
        return F
    host = F.to_scipy() if isinstance(F, DeviceCSR) else F
    out = _surgery_F(host, defect_dofs)
    return DeviceCSR.from_scipy(out) if isinstance(F, DeviceCSR) else out


def assemble_stiffness_matrix(rows_np, cols_np, vals_np, nDOFx, defect_dofs=None, nDOFy=None):
    from scipy.sparse import coo_matrix

    from ..device import DeviceCSR
    A = coo_matrix((vals_np, (rows_np, cols_np)), shape=(nDOFx, nDOFy or nDOFx)).tocsr()
    return DeviceCSR.from_scipy(_surgery_A(A, defect_dofs))


def assemble_mass_matrix(rows_np, cols_np, vals_np, nDOFx, defect_dofs=None, nDOFy=None):
    from scipy.sparse import coo_matrix

    from ..device import DeviceCSR
    F = coo_matrix((vals_np, (rows_np, cols_np)), shape=(nDOFx, nDOFy or nDOFx)).tocsr()
    return DeviceCSR.from_scipy(_surgery_F(F, defect_dofs))


def build_E0_from_m(lam100, lam111, m_nodes, R=None, M=None):
    """E0 (6N,) Voigt with engineering shear, interleaved [6n+v]; with R and M
    the rot111 variant of Micromagnetics.py:112-119 (m@R, then @M^T)."""
    import ctypes as C

    import torch

    from .._lib import check, lib, ptr
    from ..device import as_device, context
    ctx = context()
    m = as_device(m_nodes, ctx)
    if m.ndim == 1:
        m = m.reshape(-1, 3)
    n = m.shape[0]
    E0 = torch.empty(6 * n, dtype=torch.float64, device=ctx.device)
    Rh = None if R is None else np.ascontiguousarray(R, dtype=np.float64)
    Mh = None if M is None else np.ascontiguousarray(M, dtype=np.float64)
    check(lib().cpm_build_e0(ctx.h, n, float(lam100), float(lam111), ptr(m), ptr(Rh), ptr(Mh), ptr(E0)),
          "cpm_build_e0")
    return E0
