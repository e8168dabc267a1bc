"""The device-resident GSPM LLG step (Micromagnetics.py:110-307) behind one
C call per step (cpm_sim_step).

``GSPMSimulation`` owns the device state of one simulation: m, the CG warm
starts of all ten solves, Htilde, E, U, u.  ``step`` returns the report
the reference's loop needs on the host (sumsq, avg_m) plus the CG
iteration count of every solve.
"""
import ctypes as C

import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr
from ..utils.rot_111_matrices import get_M_matrix, get_R_matrix
from ..utils.sigma_matrices import get_Ebar_sigma

SYSTEMS = ("elasticity", "demag", "Gauss-Seidel g1n", "Gauss-Seidel g2n", "Gauss-Seidel g3n",
           "Gauss-Seidel g1star", "Gauss-Seidel g2star", "Gauss-Seidel m1starstar",
           "Gauss-Seidel m2starstar", "Gauss-Seidel m3starstar")


class StepReport:
    def __init__(self, rep):
        self.sumsq = rep.sumsq
        self.avg_m = np.array(rep.avg_m[:], dtype=np.float64)
        self.iters = list(rep.iters)
        self.failed = rep.failed
        self.resid = rep.resid
        # hysteresis control (set_control): outcome of the step
        self.point_done = bool(rep.point_done)
        self.state = rep.state          # 0 running, 1 stop condition met, 2 field list done
        self.point = rep.point
        self.nstep = rep.nstep
        self.use_init = bool(rep.use_init)
        self.H = np.array(rep.H[:], dtype=np.float64)
        self.H_next = np.array(rep.H_next[:], dtype=np.float64)


class GSPMSimulation:
    def __init__(self, SimOp, P, mode="fp64", rel_delta=0.3, warm_start=False, precond=None, rel_delta_gs=1e-4,
                 mg_gs=True, mg_degree=2, mg_ratio=10.0, mg_precision=None, rel_delta_mg=1e-4, overlap=True):
        """mode: "fp64" (CSR operators, the reference's sequence of solves),
        "fast64" (SELL-64 operators, 3-RHS solves) or "mixed" (as fast64 with
        FP32 CG refined in FP64 by reliable updates: an FP64 residual update
        whenever the FP32 recurrence residual has dropped by ``rel_delta``;
        ``rel_delta_gs`` sets it for the Gauss-Seidel solves alone: these
        converge fast, so fewer FP64 updates suffice; at the 128^3 headline
        0.01, 1e-3 and 1e-4 need the same iterations and give the same
        magnetization to 1e-12, and 1e-4 the fewest FP64 products).
        ``warm_start`` (fast modes): every CG starts from the previous step's
        solution instead of zero (True or 1), or from the linear
        extrapolation in time of the last two steps' solutions, 2 x_n -
        x_{n-1} ("extrapolate" or 2); the reference starts from the previous
        solution only when use_init is set, so results differ at the level
        of the CG tolerance.
        ``precond="mg"`` (fast modes, structured Kuhn meshes): the
        elasticity and demag solves run CG preconditioned by a geometric
        multigrid V-cycle (FP32 V-cycle in the mixed mode, FP64 in fast64;
        cupymag_amd.solvers.multigrid), converged to the same tolerance;
        with ``mg_gs`` also the Gauss-Seidel solves on A1 = M + alpha1 K
        (g1n..g3n as one 3-column solve, g1*, g2*); ``mg_degree`` and
        ``mg_ratio``: Chebyshev smoother degree and eigenvalue ratio;
        ``mg_precision``: V-cycle arithmetic (default 16 in the mixed mode:
        FP32 vectors with FP16 operator values; 64 in fast64; 32 possible);
        ``rel_delta_mg`` (mixed mode): the multigrid solves' outer products
        use the FP32 operator with an FP64 residual replacement whenever the
        recurrence residual drops by this factor (0: FP64 products; at the
        128^3 headline 1e-4 needs the same iterations as 1e-3 with fewer
        FP64 passes, and 3e-5 no fewer).
        ``overlap`` (fast modes, multigrid on the elasticity and demag
        solves, magnetoelastic coupling on): the demag solve runs beside the
        elasticity solve on the context's second stream, with the same
        results bit for bit (single steps; batches solve one after the
        other)."""
        self.P = P
        self.SimOp = SimOp
        A_demag, Fx, Fy, Fz, self.Deriv = SimOp.csr_demag_and_deriv()
        A1, A2, F1 = SimOp.csr_GS()
        A_el, F_el = SimOp.csr_elasticity()
        self.Avg = SimOp.volume_average()
        from ..physics.fel import FelDerivView
        fder = Fx.fel if isinstance(Fx, FelDerivView) else None
        F3 = None if fder is not None else Fx.base
        self.ctx = Fx.ctx if fder is not None else F3.ctx
        self.n = n = SimOp.n_dof
        self._keep = (A_demag, F3, A1, A2, F1, A_el, F_el)
        self.mg = None
        if precond not in (None, "mg"):
            raise ValueError(f"unknown preconditioner {precond!r} (None or 'mg')")
        if precond == "mg":
            if mode == "fp64":
                raise ValueError("precond is a fast-mode option")
            prec = mg_precision or (16 if mode == "mixed" else 64)
            self.mg = SimOp.multigrid(precision=prec, degree=mg_degree, ratio=mg_ratio)
            if self.mg is None:
                raise ValueError("multigrid needs a structured Kuhn mesh of more than 8 cells per dimension")
        mg_el = self.mg["A_el"].h if (self.mg and "A_el" in self.mg) else None
        mg_dm = self.mg["A_demag"].h if self.mg else None
        mg_a1 = self.mg["A1"].h if (self.mg and mg_gs and "A1" in self.mg) else None
        from ..physics.fel import MatrixFreeFel
        mf = F_el if isinstance(F_el, MatrixFreeFel) else None
        ops = _lib.cpm_sim_ops(A_demag.h, F3.h if F3 is not None else None, A1.h, A2.h, F1.h,
                               A_el.h if A_el is not None else None,
                               F_el.h if (F_el is not None and mf is None) else None, mg_el, mg_dm,
                               mf.h if mf is not None else None, mg_a1, fder.h if fder is not None else None)
        self._keep = self._keep + (fder,)
        pr = _lib.cpm_sim_params()
        pr.n = n
        pr.me = int(bool(P.ME))
        pr.rot111 = int(bool(P.rot111))
        pr.maxiter = int(P.maxiter)
        pr.tol = float(P.tol)
        pr.dt = float(P.dt)
        pr.alpha_damp = float(P.alpha_damp)
        pr.K1 = float(P.K1)
        ec = (P.elasC1, P.elasC2, P.elasC3) if P.ME else (0.0, 0.0, 0.0)
        for k in range(3):
            pr.elasC[k] = float(ec[k])
        pr.lambda100 = float(P.lambda100)
        pr.lambda111 = float(P.lambda111)
        for k, v in enumerate((P.N_x, P.N_y, P.N_z)):
            pr.Nd[k] = float(v)
        for k, v in enumerate(get_Ebar_sigma(P)):
            pr.Ebar_sigma[k] = float(v)
        R, M = get_R_matrix(), get_M_matrix()
        MinvT = np.linalg.inv(M).T
        for k, v in enumerate(R.ravel()):
            pr.R[k] = float(v)
        for k, v in enumerate(M.ravel()):
            pr.Mv[k] = float(v)
        for k, v in enumerate(MinvT.ravel()):
            pr.MvinvT[k] = float(v)
        pr.mode = _lib.MODES[mode]
        pr.rel_delta = float(rel_delta)
        pr.rel_delta_gs = float(rel_delta_gs or 0.0)
        pr.rel_delta_mg = float(rel_delta_mg or 0.0)
        if warm_start and mode == "fp64":
            raise ValueError("warm_start is a fast-mode option")
        pr.warm_start = 2 if warm_start in ("extrapolate", 2) else int(bool(warm_start))
        pr.precond = ((_lib.PRECOND_EL_DEMAG | (_lib.PRECOND_GS if mg_a1 else 0)) if precond == "mg" else 0)
        pr.overlap = int(bool(overlap))
        self.mode = mode
        W = self.Avg.weights_device(self.ctx)
        dmask = np.zeros(n, dtype=np.uint8)
        dmask[np.asarray(SimOp.DefDOF, dtype=np.int64)] = 1
        h = C.c_void_p()
        check(lib().cpm_sim_create(self.ctx.h, C.byref(ops), C.byref(pr), ptr(W), float(self.Avg.volume),
                                   ptr(dmask), C.byref(h)), "cpm_sim_create")
        self.h = h
        self.avg_m = None

    def set_m(self, m):
        from ..device import as_device
        md = as_device(np.asarray(m, dtype=np.float64).reshape(-1, 3) if not isinstance(m, torch.Tensor) else m,
                       self.ctx).reshape(-1, 3).contiguous()
        out = np.zeros(3)
        check(lib().cpm_sim_set_m(self.h, ptr(md), ptr(out)), "cpm_sim_set_m")
        self.avg_m = out
        return out

    def set_control(self, H0=(0.0, 0.0, 0.0), dH=(0.0, 0.0, 0.0), fields=None, steps_per_point=0,
                    llg_accuracy=None, initial_factor=None, stop_dir=None, stop_value=None, llg_min_step=None):
        """Keep the reference's hysteresis control on the device
        (Micromagnetics.py:109,299-335, cpm_sim_control): from now on the
        steps take their field and use_init from it, and each StepReport says
        whether the step completed a field point (LLG convergence, or
        ``steps_per_point``), the point index and the state (1: the stop
        condition avg_m . stop_dir <= stop_value holds, 2: the ``fields``
        list is done).  Defaults from the parameters: H0 and dH the config's
        Hext and dHext (reduced), llg_accuracy = nDOF * llg_accuracy_factor,
        the config's initial_factor, stop condition and llg_min_step."""
        P = self.P
        cfg = _lib.cpm_sweep_cfg()
        cfg.llg_accuracy = float(self.n * P.LLG_accuracy_factor if llg_accuracy is None else llg_accuracy)
        cfg.initial_factor = float(P.use_init_factor if initial_factor is None else initial_factor)
        sd = np.asarray(P.stop_direction if stop_dir is None else stop_dir, dtype=np.float64)
        for k in range(3):
            cfg.H0[k] = float(H0[k])
            cfg.dH[k] = float(dH[k])
            cfg.stop_dir[k] = float(sd[k])
        cfg.stop_value = float(P.stop_value if stop_value is None else stop_value)
        cfg.llg_min_step = int((10 if P.ME else 100) if llg_min_step is None else llg_min_step)
        cfg.steps_per_point = int(steps_per_point)
        fl = None
        if fields is not None:
            fl = np.ascontiguousarray(np.asarray(fields, dtype=np.float64).reshape(-1, 3))
            cfg.npoints = fl.shape[0]
        check(lib().cpm_sim_control(self.h, C.byref(cfg), ptr(fl)), "cpm_sim_control")
        self.controlled = True

    def clear_control(self):
        check(lib().cpm_sim_control(self.h, None, None), "cpm_sim_control")
        self.controlled = False

    def step(self, H=None, use_init=False):
        """One LLG step at field H (reduced) with the reference's use_init;
        under a control (set_control) both come from the device state."""
        if H is None:
            if not getattr(self, "controlled", False):
                raise ValueError("step needs the field H unless a control is attached (set_control)")
            H = (0.0, 0.0, 0.0)
        Hh = np.ascontiguousarray(H, dtype=np.float64)
        rep = _lib.cpm_step_report()
        info = check(lib().cpm_sim_step(self.h, ptr(Hh), int(bool(use_init)), C.byref(rep)), "cpm_sim_step")
        r = StepReport(rep)
        if info != 0:
            sysname = SYSTEMS[rep.failed - 1] if rep.failed else "?"
            raise RuntimeError(f"Error! CG for {sysname} did not converge. info={info}, residual={rep.resid}.")
        self.avg_m = r.avg_m
        return r

    def get(self, field):
        n = self.n
        shape = {"M": (n, 3), "HTILDE": (3, n), "E": (n, 6), "U": (n,), "UEL": (3 * n,)}[field]
        out = torch.empty(shape, dtype=torch.float64, device=self.ctx.device)
        check(lib().cpm_sim_get(self.h, _lib.FIELD[field], ptr(out)), "cpm_sim_get")
        return out

    def __del__(self, _release=_lib.release):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            _release("cpm_sim_destroy", h)
