"""Device context and device-resident sparse operators.

PyTorch-ROCm provides the device buffers and the stream; every computation
on them goes through libcupymag_hip.so.  ``DeviceCSR`` plays the role of
``cupyx.scipy.sparse.csr_matrix`` in the reference (final_assembly.py:
176-177,200-201): ``A @ x`` runs the library's CSR SpMV.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr

_contexts = {}


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("cupymag_amd needs a ROCm GPU: torch.cuda.is_available() is False")


class Context:
    """A library context bound to a device and torch's current stream on it."""

    def __init__(self, device=0):
        _require_gpu()
        self.device = torch.device("cuda", device)
        self.stream = torch.cuda.current_stream(self.device)
        h = C.c_void_p()
        check(lib().cpm_ctx_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(h)), "cpm_ctx_create")
        self.h = h

    def sync(self):
        check(lib().cpm_ctx_synchronize(self.h), "cpm_ctx_synchronize")

    def __del__(self, _release=_lib.release):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            _release("cpm_ctx_destroy", h)


def context(device=None):
    """The per-process context of ``device`` (default: current torch device)."""
    _require_gpu()
    d = torch.cuda.current_device() if device is None else int(device)
    stream = torch.cuda.current_stream(d).cuda_stream
    c = _contexts.get(d)
    if c is None or c.stream.cuda_stream != stream:
        with torch.cuda.device(d):
            c = Context(d)
        _contexts[d] = c
    return c


def as_device(x, ctx=None, dtype=torch.float64):
    """Contiguous float64 device tensor from a numpy array / tensor."""
    ctx = ctx or context()
    if isinstance(x, torch.Tensor):
        return x.to(device=ctx.device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=ctx.device)


class DeviceCSR:
    """A CSR operator uploaded to the device (int32 indices, FP64 values).

    ``values`` may hold several value sets on one pattern (shape (k, nnz)),
    as the Fx/Fy/Fz derivative triple does.
    """

    def __init__(self, indptr, indices, values, shape, ctx=None):
        self.ctx = ctx or context()
        indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        indices = np.ascontiguousarray(indices, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        nsets = 1 if values.ndim == 1 else values.shape[0]
        nnz = int(indptr[-1])
        if values.size != nsets * nnz:
            raise ValueError("values do not match the pattern")
        self.shape = (int(shape[0]), int(shape[1]))
        self.nnz = nnz
        self.nsets = nsets
        h = C.c_void_p()
        check(lib().cpm_mat_create_csr(self.ctx.h, self.shape[0], self.shape[1], nnz, ptr(indptr), ptr(indices),
                                       ptr(values), nsets, C.byref(h)), "cpm_mat_create_csr")
        self.h = h
        self.dtype = torch.float64
        # host copies are kept only on request (to_scipy); the device owns the data
        self._host = (indptr, indices, values) if nnz <= 50_000_000 else None

    @classmethod
    def from_scipy(cls, A, ctx=None):
        A = A.tocsr()
        return cls(A.indptr, A.indices, A.data, A.shape, ctx)

    def add_sell(self, bs, indptr=None, indices=None, values=None, grid_dims=None):
        """Attach the SELL-64 form (FP64 + FP32 values) used by the fast
        solvers; bs = 3 for the 3x3-block elasticity operator.  grid_dims:
        DOF grid (nx, ny, nz) of a structured periodic mesh numbered
        (i*ny + j)*nz + k; the library then stores no column indices when the
        pattern is a neighbour stencil (and keeps them otherwise)."""
        if indptr is None:
            indptr, indices, values = self.to_arrays()
        indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        indices = np.ascontiguousarray(indices, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        dims = None if grid_dims is None else np.ascontiguousarray(grid_dims, dtype=np.int32)
        check(lib().cpm_mat_add_sell(self.h, int(bs), ptr(indptr), ptr(indices), ptr(values), ptr(dims)),
              "cpm_mat_add_sell")
        self.sell_bs = int(bs)
        return self

    def sell_info(self):
        bs, st, sl = C.c_int(), C.c_int(), C.c_int64()
        check(lib().cpm_mat_sell_info(self.h, C.byref(bs), C.byref(st), C.byref(sl)), "cpm_mat_sell_info")
        return dict(bs=bs.value, stencil=bool(st.value), kuhn=st.value >= 2, half32=st.value >= 3,
                    half64=st.value == 4, slots=sl.value)

    def matvec(self, x, out=None, valset=0):
        x = as_device(x, self.ctx)
        if x.numel() != self.shape[1]:
            raise ValueError(f"dimension mismatch: {self.shape} @ {tuple(x.shape)}")
        y = out if out is not None else torch.empty(self.shape[0], dtype=torch.float64, device=self.ctx.device)
        check(lib().cpm_spmv(self.ctx.h, self.h, valset, ptr(x), ptr(y)), "cpm_spmv")
        return y

    def __matmul__(self, x):
        return self.matvec(x)

    def __neg__(self):
        ip, ix, v = self.to_arrays()
        return DeviceCSR(ip, ix, -v, self.shape, self.ctx)

    def to_arrays(self):
        if self._host is None:
            raise RuntimeError("host copy of this operator was not kept")
        return self._host

    def to_scipy(self, valset=0):
        import scipy.sparse as sp
        ip, ix, v = self.to_arrays()
        vv = v if v.ndim == 1 else v[valset]
        return sp.csr_matrix((vv, ix, ip), shape=self.shape)

    def __del__(self, _release=_lib.release):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            _release("cpm_mat_destroy", h)
