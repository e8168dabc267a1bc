"""Matrix-free F_el (cpm_fel_*): ``F_el @ E0`` without the 3n x 6n matrix, and
the derivative operators Fx, Fy, Fz without their CSR form.

Stands in for the CSR F_el of ``SimulationOperators.csr_elasticity``
(assemble_elasticity.py:268-348,507-527) on structured periodic Kuhn meshes;
the element geometry is recomputed on the device from the node coordinates.
"""
import ctypes as C

import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr
from ..device import as_device


class MatrixFreeFel:
    """Cvoigt None: only the derivative operators (derivs) are used."""

    def __init__(self, coords, n, Cvoigt, ctx):
        self.ctx = ctx
        self.n = int(n)
        N = self.n ** 3
        self.shape = (3 * N, 6 * N)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        if coords.shape != ((self.n + 1) ** 3, 3):
            raise ValueError("matrix-free F_el needs the (n+1)^3 node coordinates of an n^3 Kuhn mesh")
        Cm = None if Cvoigt is None else np.ascontiguousarray(Cvoigt, dtype=np.float64).reshape(36)
        h = C.c_void_p()
        check(lib().cpm_fel_create(ctx.h, self.n, ptr(coords), ptr(Cm), C.byref(h)), "cpm_fel_create")
        self.h = h
        self.has_C = Cvoigt is not None

    def derivs(self):
        """Fx, Fy, Fz of the mesh as matrix-free operators."""
        return tuple(FelDerivView(self, k) for k in range(3))

    def strain(self, u, R=None):
        N = self.n ** 3
        u = as_device(u, self.ctx).reshape(-1)
        if u.numel() != 3 * N:
            raise ValueError("u must have 3*n entries")
        E = torch.empty((N, 6), dtype=torch.float64, device=self.ctx.device)
        Rh = None if R is None else np.ascontiguousarray(np.asarray(R, dtype=np.float64))
        check(lib().cpm_fel_strain(self.h, ptr(u), ptr(Rh), ptr(E)), "cpm_fel_strain")
        return E

    def hd(self, U):
        N = self.n ** 3
        U = as_device(U, self.ctx).reshape(-1)
        if U.numel() != N:
            raise ValueError("U must have n entries")
        H = torch.empty(3 * N, dtype=torch.float64, device=self.ctx.device)
        check(lib().cpm_fel_hd(self.h, ptr(U), ptr(H)), "cpm_fel_hd")
        return H[:N], H[N:2 * N], H[2 * N:]

    def matvec(self, E0, out=None):
        if not self.has_C:
            raise ValueError("this matrix-free operator was made without the elastic constants (derivatives only)")
        E0 = as_device(E0, self.ctx).reshape(-1)
        if E0.numel() != self.shape[1]:
            raise ValueError(f"dimension mismatch: {self.shape} @ {tuple(E0.shape)}")
        y = out if out is not None else torch.empty(self.shape[0], dtype=torch.float64, device=self.ctx.device)
        check(lib().cpm_fel_apply(self.h, ptr(E0), ptr(y)), "cpm_fel_apply")
        return y

    def __matmul__(self, x):
        return self.matvec(x)

    def __del__(self, _release=_lib.release):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            _release("cpm_fel_destroy", h)


class FelDerivView:
    """Fx, Fy or Fz (k = 0, 1, 2) of a MatrixFreeFel: ``F @ x`` on the device
    without the CSR form (assemble_demag.py:235-305 element values)."""

    def __init__(self, fel, k):
        self.fel, self.k = fel, int(k)
        N = fel.n ** 3
        self.shape = (N, N)
        self.ctx = fel.ctx

    def __matmul__(self, x):
        x = as_device(x, self.ctx).reshape(-1)
        if x.numel() != self.shape[1]:
            raise ValueError(f"dimension mismatch: {self.shape} @ {tuple(x.shape)}")
        y = torch.empty(self.shape[0], dtype=torch.float64, device=self.ctx.device)
        check(lib().cpm_fel_dmul(self.fel.h, self.k, ptr(x), ptr(y)), "cpm_fel_dmul")
        return y
