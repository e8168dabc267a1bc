"""Geometric multigrid for the elasticity and demag solves (cpm_mg_*).

The reference solves both systems with unpreconditioned CG
(solvers/linear_solvers.py:20-74, ``M=None``); their iteration counts grow
with the mesh size (about 2 600 for the elasticity system and 700 for the
demag system per step at 128^3).  On a structured periodic Kuhn mesh the
operators have a natural hierarchy: the Kuhn mesh of the even nodes (or,
for an odd number of cells, the regular Kuhn mesh of half as many cells,
rounded up, over the same box), assembled again by the host assembler
(rediscretisation), down to <= 8 cells per dimension, where a dense inverse
finishes the V-cycle.
``build_hierarchies`` builds it for a ``SimulationOperators`` object;
``KuhnMultigrid.solve`` runs the preconditioned CG (same stopping rule as
``solve_cg``).
"""
import ctypes as C

import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr
from ..device import DeviceCSR, as_device
from ..mesh.coarsen import coarse_kuhn, hierarchy_sizes, kuhn_cells
from ..mesh.periodic import periodic_node_map, structured_dims


class KuhnMultigrid:
    """A multigrid hierarchy (levels[0] is the operator being solved)."""

    def __init__(self, levels, coarse_dense, precision=32, degree=2, ratio=10.0, floating=False,
                 coarse_smooth=None, coarse_degree=0):
        """floating: levels[0] is anchored, levels[1:] are the periodic
        stiffness and coarse_dense its pseudo-inverse (cpm_mg_create tells
        the kinds apart from the operators themselves; the flag is kept for
        info()).  coarse_smooth = (degree, ratio), with coarse_dense None:
        the last level is solved by that many Chebyshev steps over
        [lmax / ratio, lmax] (cpm_mg_create_smoothed; operators without an
        anchor).  coarse_degree: the Chebyshev degree of the levels between
        the finest and the coarsest (0: ``degree``; cpm_mg_set_coarse_degree)."""
        self.floating = bool(floating)
        self.levels = list(levels)
        self.ctx = levels[0].ctx
        self.precision = int(precision)
        self.degree = int(degree)
        self.ratio = float(ratio)
        self.coarse_smooth = None if coarse_smooth is None else (int(coarse_smooth[0]), float(coarse_smooth[1]))
        arr = (C.c_void_p * len(levels))(*[L.h.value for L in levels])
        h = C.c_void_p()
        if self.coarse_smooth is not None:
            check(lib().cpm_mg_create_smoothed(self.ctx.h, len(levels), arr, self.precision, self.degree, self.ratio,
                                               self.coarse_smooth[0], self.coarse_smooth[1], C.byref(h)),
                  "cpm_mg_create_smoothed")
        else:
            cinv = np.ascontiguousarray(coarse_dense, dtype=np.float64)
            check(lib().cpm_mg_create(self.ctx.h, len(levels), arr, ptr(cinv), self.precision, self.degree,
                                      self.ratio, C.byref(h)), "cpm_mg_create")
        self.h = h
        self.coarse_degree = int(coarse_degree)
        if self.coarse_degree:
            check(lib().cpm_mg_set_coarse_degree(h, self.coarse_degree), "cpm_mg_set_coarse_degree")

    def info(self):
        n = C.c_int()
        lm = (C.c_double * 16)()
        check(lib().cpm_mg_info(self.h, C.byref(n), lm, 16), "cpm_mg_info")
        hs = (C.c_int * 16)()
        check(lib().cpm_mg_forms(self.h, hs, 16), "cpm_mg_forms")
        return dict(levels=n.value, lmax=list(lm[:n.value]), dims=[getattr(L, "grid_dims", None) for L in self.levels],
                    floating=self.floating, half_stencil=[bool(v) for v in hs[:n.value]],
                    coarsest="dense" if self.coarse_smooth is None else ("chebyshev",) + self.coarse_smooth,
                    coarse_degree=self.coarse_degree or self.degree)

    def set_graphs(self, on):
        """Replay the PCG iterations from captured hipGraph chunks (default)
        or launch every kernel; the results are the same bit for bit."""
        check(lib().cpm_mg_set_graphs(self.h, int(bool(on))), "cpm_mg_set_graphs")

    def apply(self, r):
        """One V-cycle: z = M^-1 r (vectors in the SELL order)."""
        r = as_device(r, self.ctx)
        z = torch.empty_like(r)
        check(lib().cpm_mg_apply(self.h, ptr(r), ptr(z)), "cpm_mg_apply")
        return z

    def solve(self, b, x0=None, tol=1e-7, maxiter=5000, use_init=False, nrhs=1, delta=0.0):
        """(x, iterations, residual) of the preconditioned CG; raises the
        reference's error when maxiter is reached.  nrhs = 3 (scalar
        operators): b holds three interleaved columns (entry i*3 + j) and
        iterations / residual are lists of three.  delta > 0: FP32 operator
        products with FP64 residual replacements (stopping test on the FP64
        residual)."""
        b = as_device(b, self.ctx)
        x = x0.clone() if (use_init and x0 is not None) else torch.empty_like(b)
        it, res = (C.c_int32 * nrhs)(), (C.c_double * nrhs)()
        info = check(lib().cpm_cg_mg(self.h, int(nrhs), ptr(b), ptr(x), int(bool(use_init and x0 is not None)),
                                     float(tol), int(maxiter), float(delta), it, res), "cpm_cg_mg")
        if info > 0:
            raise RuntimeError(f"Error! CG did not converge. info={info}, residual={max(res)}.")
        if nrhs == 1:
            return x, it[0], res[0]
        return x, list(it), list(res)

    def __del__(self, _release=_lib.release):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            _release("cpm_mg_destroy", h)


def solve_batch(mgs, rhs, x0s=None, tol=1e-7, maxiter=5000, nrhs=1):
    """``KuhnMultigrid.solve`` on many independent systems at once
    (cpm_cg_mg_batch): one launch per phase for the whole batch.  mgs:
    hierarchies (one block size, precision and degree; distinct), rhs: their
    right-hand sides (nrhs interleaved columns each), x0s: optional initial
    guesses (used where not None).  Each system's (x, iterations, residual)
    equals ``solve``'s bit for bit.  Raises the reference's error when a
    system reaches maxiter."""
    if not mgs:
        return []
    ctx = mgs[0].ctx
    rhs = [as_device(b, ctx) for b in rhs]
    x0s = list(x0s) if x0s is not None else [None] * len(mgs)
    xs = [x0.clone() if x0 is not None else torch.empty_like(b) for b, x0 in zip(rhs, x0s)]
    n = len(mgs)
    h = (C.c_void_p * n)(*[m.h.value for m in mgs])
    bp = (C.c_void_p * n)(*[b.data_ptr() for b in rhs])
    xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
    ux = (C.c_int32 * n)(*[int(x0 is not None) for x0 in x0s])
    it, res, fl = (C.c_int32 * (n * nrhs))(), (C.c_double * (n * nrhs))(), (C.c_int32 * n)()
    check(lib().cpm_cg_mg_batch(ctx.h, n, int(nrhs), h, bp, xp, ux, float(tol), int(maxiter), it, res, fl),
          "cpm_cg_mg_batch")
    out = []
    for j, x in enumerate(xs):
        its, rs = list(it[j * nrhs:(j + 1) * nrhs]), list(res[j * nrhs:(j + 1) * nrhs])
        if fl[j]:
            raise RuntimeError(f"Error! CG did not converge. info={max(its)}, residual={max(rs)}.")
        out.append((x, its[0], rs[0]) if nrhs == 1 else (x, its, rs))
    return out


def _interleave_perm(n, bs):
    """perm[i*bs + d] = d*n + i: SoA (reference) order -> node-major order."""
    i = np.arange(n)
    return (np.arange(bs)[None, :] * n + i[:, None]).ravel()


def _pinv_periodic(A, bs):
    """Pseudo-inverse of a singular periodic stiffness whose null space is
    the bs translations (node-major interleaved order): with Q their
    orthonormal basis, A^+ = (A + s Q Q^T)^-1 - Q Q^T / s for any s > 0
    (s = the mean diagonal entry, for conditioning)."""
    n = A.shape[0] // bs
    Q = np.zeros((A.shape[0], bs))
    for d in range(bs):
        Q[d::bs, d] = 1.0 / np.sqrt(n)
    sc = float(np.mean(np.diag(A)))
    QQ = Q @ Q.T
    return np.linalg.inv(A + sc * QQ) - QQ / sc


# A1 = M + alpha1 K: the stiffness part shrinks against the mass part as the
# mesh coarsens (alpha1 K / M ~ 1/h^2), so a coarse level of A1 is well
# conditioned and a few Chebyshev steps solve it as well as the levels and
# the dense solve below it would.  The first coarse level whose Jacobi-scaled
# condition number is at most A1_KAPPA_MAX becomes the coarsest, solved by
# A1_CHEB Chebyshev steps over [lmax / A1_CHEB_RATIO, lmax].  At the 128^3
# headline (kappa 131 on level 0, 33 on the 64^3 level, 8.4 on 32^3) this
# keeps two levels; a CPU probe of the PCG (SciPy, same smoother) took 15
# iterations with it against 16 with the five levels down to the dense 8^3
# solve.  On the GPU at 128^3 (20 steps) each A1 solve took 19-20
# iterations with 4 coarsest steps, 16 with 6 and 15 with 8 (44.8, 42.6 and
# 42.3 ms per step): 8, the most the library takes.
A1_KAPPA_MAX = 40.0
# The elasticity V-cycle's levels below the finest smooth with Chebyshev
# degree 3 (the finest keeps `degree`): a CPU probe of the periodic
# elasticity PCG at 128^3 took 18 iterations against 20 with degree 2 on
# every level (16 with degree 3 on every level, at 2 more level-0 passes
# per iteration), for about 5% more V-cycle work per iteration.
EL_COARSE_DEGREE = 3
A1_CHEB = 8
A1_CHEB_RATIO = 20.0


def jacobi_kappa(ip, ix, v, n):
    """Condition number of D^-1/2 A D^-1/2 for a host CSR operator (Lanczos,
    loose tolerance), or inf when the estimate does not converge."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as sla
    A = sps.csr_matrix((np.asarray(v, dtype=np.float64), np.asarray(ix), np.asarray(ip)), shape=(n, n))
    d = 1.0 / np.sqrt(A.diagonal())
    B = sps.diags(d) @ A @ sps.diags(d)
    try:
        lo = sla.eigsh(B, k=1, which="SA", tol=1e-2, maxiter=20 * n, return_eigenvectors=False)[0]
        hi = sla.eigsh(B, k=1, which="LA", tol=1e-2, maxiter=20 * n, return_eigenvectors=False)[0]
    except sla.ArpackNoConvergence:
        return float("inf")
    return float(hi / lo) if lo > 0 else float("inf")


def scalar_levels(sizes, coarsest_scalar=8):
    """Levels of a scalar operator's hierarchy: the first level of at most
    coarsest_scalar cells per dimension is its coarsest (dense) level, and
    there are at least two."""
    k = next((i for i, m in enumerate(sizes) if m <= coarsest_scalar), len(sizes) - 1)
    return sizes[:max(2, k + 1)]


def build_hierarchies(SimOp, precision=32, degree=2, ratio=10.0, coarsest_max=8, nthreads=None, floating=True,
                      coarsest_scalar=8, a1_coarse="auto", el_coarse_degree=EL_COARSE_DEGREE):
    """{"A_demag": KuhnMultigrid, "A1": KuhnMultigrid (Gauss-Seidel operator
    M + alpha1 K; coarse levels without the defect surgery), "A_el":
    KuhnMultigrid (with coupling)} for a SimulationOperators object on an
    n^3 Kuhn mesh with n > coarsest_max,
    or None when the mesh has no such hierarchy (not a Kuhn mesh in
    kuhn_mesh order, or too small to coarsen).

    floating (default): the coarse levels of the anchored operators (A_demag,
    A_el) are the singular periodic stiffness, without the anchor, and the
    coarsest solve is its pseudo-inverse; the V-cycle then solves the
    periodic problem and the anchor is applied around it (cpm_mg_create),
    which keeps the iteration counts flat as the mesh grows.  False: anchored
    coarse levels (node 0 left out of the transfers).

    The elasticity hierarchy coarsens to <= coarsest_max cells per
    dimension; the scalar ones (A_demag, A1) to <= coarsest_scalar
    (scalar_levels; 16 makes their dense level up to 16^3 unknowns, 64 MB in
    FP32, in place of the sweeps and transfers of the levels below: no
    faster at 128^3, and a longer set-up).

    a1_coarse: "auto" stops the A1 hierarchy at its first coarse level whose
    Jacobi-scaled condition number is at most A1_KAPPA_MAX and solves that
    level with A1_CHEB Chebyshev steps (cpm_mg_create_smoothed); "dense"
    keeps the levels down to the dense solve.

    el_coarse_degree: the Chebyshev degree of the elasticity hierarchy's
    levels below the finest (0: ``degree``)."""
    if a1_coarse not in ("auto", "dense"):
        raise ValueError(f"a1_coarse must be 'auto' or 'dense', not {a1_coarse!r}")
    from ..physics.host_ops import HostOps, voigt_C
    n = kuhn_cells(SimOp.node_coords_np, SimOp.elements_np)
    if n is None or SimOp.grid_dims != (n, n, n):
        return None
    sizes = hierarchy_sizes(n, coarsest_max)
    if len(sizes) < 2:
        return None
    P = SimOp.P
    ctx = SimOp.csr_demag_and_deriv()[0].ctx
    fine = {"A_demag": SimOp.csr_demag_and_deriv()[0], "A1": SimOp.csr_GS()[0]}
    if P.ME:
        fine["A_el"] = SimOp.csr_elasticity()[0]
    Cv = voigt_C(P.c11, P.c12, P.c44, bool(P.rot111)) if P.ME else None
    levels = {k: [v] for k, v in fine.items()}
    nlev = {k: (len(sizes) if k == "A_el" else len(scalar_levels(sizes, coarsest_scalar))) for k in fine}
    dense = {}
    smoothed = set()
    coords = SimOp.node_coords_np
    for li, m in enumerate(sizes[1:]):
        if all(len(levels[k]) == nlev[k] for k in fine):
            break
        coords, tets = coarse_kuhn(coords, sizes[li])
        el = np.zeros((tets.shape[0], 5), dtype=np.int32)
        el[:, :4] = tets
        gid = periodic_node_map(coords, coords.min(axis=0), coords.max(axis=0))
        ndof = int(gid.max()) + 1
        dims = structured_dims(coords, gid)
        if dims != (m, m, m):
            return None
        H = HostOps(coords, el, gid, ndof, None, SimOp.alpha1, SimOp.alpha2, Cv, nthreads, anchors=not floating)
        for name, bs in (("A_demag", 1), ("A_el", 3), ("A1", 1)):
            if name not in fine or len(levels[name]) == nlev[name]:
                continue
            last = len(levels[name]) == nlev[name] - 1
            ip, ix, v, shape = H.csr(name)
            if name == "A1" and a1_coarse == "auto" and jacobi_kappa(ip, ix, v, shape[0]) <= A1_KAPPA_MAX:
                # a well-conditioned coarse level: the last one, smoothed
                last = True
                nlev[name] = len(levels[name]) + 1
                smoothed.add(name)
            M = DeviceCSR(ip, ix, v, shape, ctx)
            # the host arrays are views of the assembler's buffers, freed below
            M._host = (ip.copy(), ix.copy(), np.array(v, copy=True))
            M.grid_dims = dims
            M.add_sell(bs, ip, ix, v, dims)
            levels[name].append(M)
            if last and name not in smoothed:
                import scipy.sparse as sps
                A = sps.csr_matrix((np.array(v), np.array(ix), np.array(ip)), shape=shape).toarray()
                perm = _interleave_perm(ndof, bs)
                Ai = A[np.ix_(perm, perm)]
                dense[name] = _pinv_periodic(Ai, bs) if floating and name != "A1" else np.linalg.inv(Ai)
        H.free()
    return {name: KuhnMultigrid(levels[name], dense.get(name), precision, degree, ratio,
                                floating=floating and name != "A1",
                                coarse_smooth=(A1_CHEB, A1_CHEB_RATIO) if name in smoothed else None,
                                coarse_degree=el_coarse_degree if (name == "A_el" and degree >= 2) else 0)
            for name in fine}
