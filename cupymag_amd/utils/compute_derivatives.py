"""Weak-form derivative operators (cupymag/utils/compute_derivatives.py:31-125).

Fx, Fy, Fz share one sparsity pattern; when they are the three value sets
of one device operator, ``compute_Hd_from_U`` (3 SpMVs) and
``compute_E_from_u`` (6 or 9 SpMVs plus the rotation mixing) each run as a
single fused pass over the pattern (cpm_hd_from_u / cpm_strain_from_u);
matrix-free Fx, Fy, Fz (physics.fel.FelDerivView) walk the elements instead
(cpm_fel_hd / cpm_fel_strain).
"""
import numpy as np
import torch

from .._lib import check, lib, ptr
from ..device import as_device


class DerivView:
    """One value set of a multi-value device operator (behaves like Fx)."""

    def __init__(self, base, valset):
        self.base, self.valset = base, valset
        self.shape, self.nnz, self.ctx = base.shape, base.nnz, base.ctx

    def __matmul__(self, x):
        return self.base.matvec(x, valset=self.valset)

    def __neg__(self):
        ip, ix, v = self.base.to_arrays()
        from ..device import DeviceCSR
        return DeviceCSR(ip, ix, -v[self.valset], self.shape, self.ctx)

    def to_scipy(self):
        return self.base.to_scipy(self.valset)


class ComputeDerivatives:
    def __init__(self, Fx, Fy, Fz):
        from ..physics.fel import FelDerivView
        self.Dx, self.Dy, self.Dz = Fx, Fy, Fz
        fused = all(isinstance(F, DerivView) for F in (Fx, Fy, Fz)) and \
            Fx.base is Fy.base is Fz.base and (Fx.valset, Fy.valset, Fz.valset) == (0, 1, 2)
        self.F3 = Fx.base if fused else None
        # matrix-free Fx, Fy, Fz of one mesh (cpm_fel_strain / cpm_fel_hd)
        mf = all(isinstance(F, FelDerivView) for F in (Fx, Fy, Fz)) and Fx.fel is Fy.fel is Fz.fel and \
            (Fx.k, Fy.k, Fz.k) == (0, 1, 2)
        self.fel = Fx.fel if mf else None

    def compute_Hd_from_U(self, U):
        if self.fel is not None:
            return self.fel.hd(U)
        if self.F3 is None:
            return -(self.Dx @ U), -(self.Dy @ U), -(self.Dz @ U)
        ctx = self.F3.ctx
        U = as_device(U, ctx)
        n = self.F3.shape[0]
        H = torch.empty(3 * n, dtype=torch.float64, device=ctx.device)
        check(lib().cpm_hd_from_u(ctx.h, self.F3.h, ptr(U), ptr(H)), "cpm_hd_from_u")
        return H[:n], H[n:2 * n], H[2 * n:]

    def compute_E_from_u(self, n, u, R=None):
        if self.fel is not None:
            if self.fel.n ** 3 != n:
                raise ValueError("u must have 3*n entries")
            return self.fel.strain(u, R)
        if self.F3 is None:
            raise NotImplementedError("compute_E_from_u needs Fx, Fy, Fz on one device pattern")
        ctx = self.F3.ctx
        u = as_device(u, ctx)
        if u.numel() != 3 * n or self.F3.shape[0] != n:
            raise ValueError("u must have 3*n entries")
        E = torch.empty((n, 6), dtype=torch.float64, device=ctx.device)
        Rh = None if R is None else np.ascontiguousarray(np.asarray(R, dtype=np.float64))
        check(lib().cpm_strain_from_u(ctx.h, self.F3.h, ptr(u), ptr(Rh), ptr(E)), "cpm_strain_from_u")
        return E
