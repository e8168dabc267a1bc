"""CG wrapper (cupymag/solvers/linear_solvers.py:20-74) on the device CG of
libcupymag_hip.so (cpm_cg), which follows CuPy's cg: x0 = 0 unless
use_init; b returned when ||b|| = 0; atol = tol*||b||; at least one
iteration; convergence tested after each update.  A preconditioner M is
forwarded as the reference forwards it to cupyx cg (cpm_cg_precond:
z = M r every iteration).  Non-convergence raises RuntimeError with the
reference's message, including the true residual ||b - A x||."""
import ctypes as C

import torch

from .._lib import check, lib, ptr
from ..device import as_device


def _as_operator(M, ctx, n):
    """The device operator of a preconditioner: a DeviceCSR (or one value set
    of it), or a SciPy sparse / dense square matrix, uploaded once.  Other
    kinds (a LinearOperator with only a matvec) have no device form here."""
    from ..device import DeviceCSR
    Mb = getattr(M, "base", M)
    if isinstance(Mb, DeviceCSR):
        if getattr(M, "valset", 0) != 0 or Mb.nsets != 1:
            raise ValueError("the preconditioner must be a single-valued operator")
        return Mb
    import numpy as np
    import scipy.sparse as sp
    if sp.issparse(M) or isinstance(M, np.ndarray):
        Mc = sp.csr_matrix(M, dtype=np.float64)
        if Mc.shape != (n, n):
            raise ValueError(f"the preconditioner must be {n} x {n}, got {Mc.shape}")
        return DeviceCSR(Mc.indptr.astype(np.int32), Mc.indices.astype(np.int32), Mc.data, Mc.shape, ctx)
    raise TypeError("solve_cg: M must be a sparse or dense matrix (applied as z = M r); "
                    f"got {type(M).__name__}")


def solve_cg(A, b, M=None, x0=None, tol=1e-7, maxiter=5000, use_init=False, system=None, return_info=False):
    base = getattr(A, "base", A)
    if getattr(A, "valset", 0) != 0 or base.nsets != 1:
        raise ValueError("solve_cg needs a single-valued square operator")
    ctx = base.ctx
    b = as_device(b, ctx)
    if use_init and x0 is not None:
        x = as_device(x0, ctx).clone()
    else:
        x = torch.empty_like(b)
        use_init = False
    it = C.c_int32()
    res = C.c_double()
    if M is None:
        info = check(lib().cpm_cg(ctx.h, base.h, ptr(b), ptr(x), int(bool(use_init)), float(tol), int(maxiter),
                                  C.byref(it), C.byref(res)), "cpm_cg")
    else:
        Md = _as_operator(M, ctx, base.shape[0])
        info = check(lib().cpm_cg_precond(ctx.h, base.h, Md.h, ptr(b), ptr(x), int(bool(use_init)), float(tol),
                                          int(maxiter), C.byref(it), C.byref(res)), "cpm_cg_precond")
    if info != 0:
        rnorm = torch.linalg.norm(b - base.matvec(x)).item()
        who = f"CG for {system}" if system is not None else "CG"
        raise RuntimeError(f"Error! {who} did not converge. info={info}, residual={rnorm}.")
    if return_info:
        return x, int(it.value), float(res.value)
    return x


def solve_cg_sell(A, b, x0=None, tol=1e-7, maxiter=5000, precision=64, rel_delta=0.3, nrhs=1):
    """CG on the SELL-64 form of ``A`` (see DeviceCSR.add_sell) for ``nrhs``
    interleaved right-hand sides (b[i*nrhs + j]); bs-3 operators take
    node-major interleaved vectors.  precision 64: CuPy's recurrence per
    column; 32: FP32 CG refined in FP64 (reliable updates).  Returns
    (x, iters[nrhs], resid[nrhs]); raises RuntimeError on non-convergence."""
    ctx = A.ctx
    b = as_device(b, ctx)
    use_init = x0 is not None
    x = as_device(x0, ctx).clone() if use_init else torch.empty_like(b)
    it = (C.c_int32 * nrhs)()
    res = (C.c_double * nrhs)()
    info = check(lib().cpm_cg_sell(ctx.h, A.h, int(nrhs), ptr(b), ptr(x), int(use_init), float(tol), int(maxiter),
                                   int(precision), float(rel_delta), it, res), "cpm_cg_sell")
    if info != 0:
        raise RuntimeError(f"Error! CG did not converge. info={info}, residual={max(res)}.")
    return x, list(it), list(res)
