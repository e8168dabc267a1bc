"""All operators of the GSPM step (cupymag/core/assemble_simulation_operators.py:27-117).

The reference assembles COO triplets per operator with Numba and converts
with SciPy.  Here one call of the native host assembler builds every
operator straight into CSR (in parallel over DOF rows, memory proportional
to the final nnz), and the operators are uploaded once.  Fx, Fy, Fz are
uploaded as three value sets of one device pattern so that the derivative
kernels read the pattern once.
"""
import numpy as np

from ..device import DeviceCSR, context
from ..mesh.setup_FEM_mesh import FEMMesh
from ..physics.host_ops import HostOps, voigt_C
from ..utils.compute_derivatives import ComputeDerivatives, DerivView
from ..utils.volume_average import VolumeAverage


class SimulationOperators(FEMMesh):
    def __init__(self, P=None, mesh_arrays=None, ctx=None, keep_host=None, nthreads=None, sell=True, autotune=False,
                 matrix_free_fel=False, matrix_free_deriv=False):
        """``P``: the simulation parameters; ``SimulationOperators()`` reads
        them, as the reference does at import (parameters.py:44-57), from
        the config named by ``sys.argv[1]``, else ``$CUPYMAG_CONFIG_PATH``,
        else ./config.yaml (core.parameters.load_parameters).
        ``sell``: also attach the SELL-64 forms the fast / mixed-precision
        solvers run on (A_demag, A1, A2 scalar; A_el in 3x3 blocks).
        ``autotune``: time the launch schedules of every operator after the
        upload and keep the fastest (cupymag_amd.autotune, cached per size
        class); ``self.tuning`` then maps operator names to their schedule
        and timings.  ``matrix_free_fel``: on a structured Kuhn mesh, F_el is
        not uploaded; ``csr_elasticity`` returns a matrix-free operator
        (cupymag_amd.physics.fel) in its place.  ``matrix_free_deriv``: likewise
        Fx, Fy, Fz (their 0.9 GB CSR at 128^3 is not uploaded;
        ``csr_demag_and_deriv`` returns matrix-free operators, and the step
        forms the strain, Htilde and the demag right-hand side by walking the
        elements)."""
        if P is None:
            from .parameters import load_parameters
            P = load_parameters()
        super().__init__(P, mesh_arrays)
        self.sell = sell
        self.autotune = autotune
        self.tuning = {}
        self.matrix_free_fel = matrix_free_fel
        self.matrix_free_deriv = matrix_free_deriv
        self.alpha1 = P.alpha
        self.alpha2 = P.alpha * P.alpha_damp
        self.ctx = ctx
        self.keep_host = keep_host
        self.nthreads = nthreads
        self._dev = None

    def _upload(self):
        if self._dev is not None:
            return self._dev
        P = self.P
        ctx = self.ctx or context()
        Cv = voigt_C(P.c11, P.c12, P.c44, bool(P.rot111)) if P.ME else None
        H = HostOps(self.node_coords_np, self.elements_np, self.global_id_np, self.n_dof, self.DefDOF,
                    self.alpha1, self.alpha2, Cv, self.nthreads)
        keep = self.keep_host if self.keep_host is not None else self.n_dof <= 300_000

        def up(name, vals=None, sell_bs=0):
            ip, ix, v, shape = H.csr(name)
            if vals is not None:
                v = vals
            M = DeviceCSR(ip, ix, v, shape, ctx)
            if sell_bs and self.sell:
                M.add_sell(sell_bs, ip, ix, v, self.grid_dims)
            if not keep:
                M._host = None
            else:
                M._host = (ip.copy(), ix.copy(), np.array(v, copy=True))
            return M

        d = {}
        d["A_demag"] = up("A_demag", sell_bs=1)
        mf = None
        if self.matrix_free_deriv or (P.ME and self.matrix_free_fel):
            mf = self._matrix_free_fel(ctx, Cv if (P.ME and self.matrix_free_fel) else None)
        d["fder"] = mf if (self.matrix_free_deriv and mf is not None) else None
        if d["fder"] is None:
            F3v = np.stack([H.array("FX"), H.array("FY"), H.array("FZ")])
            d["F3"] = up("Fx", F3v)
            del F3v
        else:
            d["F3"] = None
        for k in ("A1", "A2"):
            d[k] = up(k, sell_bs=1)
        d["F1"] = up("F1", sell_bs=1)
        if P.ME:
            d["A_el"] = up("A_el", sell_bs=3)
            d["F_el"] = mf if (self.matrix_free_fel and mf is not None) else up("F_el")
        H.free()
        if self.autotune:
            from ..autotune import tune
            for name, M in d.items():
                if isinstance(M, DeviceCSR):
                    self.tuning[name] = tune(M)
        self._dev = d
        return d

    def _matrix_free_fel(self, ctx, Cv):
        """A MatrixFreeFel for an n^3 Kuhn mesh in generator order, else None."""
        from ..mesh.coarsen import kuhn_cells
        from ..physics.fel import MatrixFreeFel
        n = kuhn_cells(self.node_coords_np, self.elements_np)
        if n is None or self.grid_dims != (n, n, n):
            return None
        return MatrixFreeFel(self.node_coords_np, n, Cv, ctx)

    def multigrid(self, precision=32, degree=2, ratio=10.0, floating=True, coarsest_scalar=8, a1_coarse="auto",
                  el_coarse_degree=None):
        """Multigrid hierarchies of A_demag, A1 and A_el (cached per
        settings), or None when the mesh has none
        (cupymag_amd.solvers.multigrid.build_hierarchies)."""
        from ..solvers.multigrid import EL_COARSE_DEGREE, build_hierarchies
        cache = self.__dict__.setdefault("_mg", {})
        ecd = EL_COARSE_DEGREE if el_coarse_degree is None else int(el_coarse_degree)
        key = (precision, degree, ratio, bool(floating), int(coarsest_scalar), a1_coarse, ecd)
        if key not in cache:
            cache[key] = build_hierarchies(self, precision, degree, ratio, nthreads=self.nthreads,
                                           floating=floating, coarsest_scalar=coarsest_scalar, a1_coarse=a1_coarse,
                                           el_coarse_degree=ecd)
        return cache[key]

    def csr_demag_and_deriv(self):
        d = self._upload()
        if d["F3"] is None:
            Fx, Fy, Fz = d["fder"].derivs()
            return d["A_demag"], Fx, Fy, Fz, ComputeDerivatives(Fx, Fy, Fz)
        Fx, Fy, Fz = (DerivView(d["F3"], k) for k in range(3))
        return d["A_demag"], Fx, Fy, Fz, ComputeDerivatives(Fx, Fy, Fz)

    def csr_GS(self):
        d = self._upload()
        return d["A1"], d["A2"], d["F1"]

    def csr_elasticity(self):
        if not self.P.ME:
            return None, None
        d = self._upload()
        return d["A_el"], d["F_el"]

    def volume_average(self):
        if getattr(self, "_avg", None) is None:
            self._avg = VolumeAverage(self.node_coords_np, self.elements_np, self.global_id_np)
        return self._avg
