"""Matrix-free F_el (cpm_fel_*): ``F_el @ E0`` without the 3n x 6n matrix.

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
    def __init__(self, coords, n, Cvoigt, ctx):
        self.ctx = ctx
        self.n = int(n)
        N = self.n ** 3
        self.shape = (3 * N, 6 * N)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        if coords.shape != ((self.n + 1) ** 3, 3):
            raise ValueError("matrix-free F_el needs the (n+1)^3 node coordinates of an n^3 Kuhn mesh")
        Cm = np.ascontiguousarray(Cvoigt, dtype=np.float64).reshape(36)
        h = C.c_void_p()
        check(lib().cpm_fel_create(ctx.h, self.n, ptr(coords), ptr(Cm), C.byref(h)), "cpm_fel_create")
        self.h = h

    def matvec(self, E0, out=None):
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
