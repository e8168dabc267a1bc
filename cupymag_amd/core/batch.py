"""Many GSPM simulations stepped together (cpm_batch_*).

The reference runs one simulation per process (Micromagnetics.py:34).  A
``GSPMBatch`` holds any number of ``GSPMSimulation`` objects of any mesh
sizes (fast modes, all the same mode) and advances them by one LLG step
each per ``step`` call.  Each of the step's ten solves runs as one batched
CG over every member (one launch per CG phase for the whole batch); members
created with ``precond="mg"`` run their elasticity and demag solves as one
batched multigrid-preconditioned CG instead.  The field kernels run per
member.  Members can join and leave between steps,
which is what the serving loop uses for continuous batching.  A member's
results are bit-identical alone and in any batch.
"""
import ctypes as C

import numpy as np

from .. import _lib
from .._lib import check, lib, ptr
from .gspm import SYSTEMS, StepReport


class GSPMBatch:
    def __init__(self, ctx, graphs=False):
        """ctx: the device context every member uses.  graphs: replay
        captured chunks of CG iterations (hipGraph) instead of launching
        every kernel from the host."""
        self.ctx = ctx
        self.graphs = bool(graphs)
        h = C.c_void_p()
        check(lib().cpm_batch_create(ctx.h, _lib.BATCH_GRAPH if graphs else 0, C.byref(h)), "cpm_batch_create")
        self.h = h
        self.members = []

    def add(self, sim):
        if sim.ctx is not self.ctx:
            raise ValueError("batch members must share the batch's context")
        check(lib().cpm_batch_add(self.h, sim.h), "cpm_batch_add")
        self.members.append(sim)
        return sim

    def remove(self, sim):
        check(lib().cpm_batch_remove(self.h, sim.h), "cpm_batch_remove")
        self.members.remove(sim)

    def __len__(self):
        return len(self.members)

    def step(self, H=None, use_init=None, raise_on_fail=True):
        """One step of every member.  H: (len, 3) reduced external fields,
        use_init: (len,) warm-start flags, both in member order (ignored for
        members with a control attached, GSPMSimulation.set_control; None:
        zeros, when every member has one).  Returns
        one StepReport per member; a member whose solve did not converge
        raises the reference's error (raise_on_fail) or is reported with
        ``failed`` set."""
        nb = len(self.members)
        if nb == 0:
            return []
        Hh = np.ascontiguousarray(np.zeros((nb, 3)) if H is None else np.asarray(H, dtype=np.float64).reshape(nb, 3))
        ui = np.ascontiguousarray(np.zeros(nb, dtype=np.int32) if use_init is None else
                                  np.asarray(use_init, dtype=np.int32).reshape(nb))
        reps = (_lib.cpm_step_report * nb)()
        check(lib().cpm_batch_step(self.h, ptr(Hh), ptr(ui), reps), "cpm_batch_step")
        out = [StepReport(r) for r in reps]
        for sim, r in zip(self.members, out):
            if r.failed:
                if raise_on_fail:
                    raise RuntimeError(f"Error! CG for {SYSTEMS[r.failed - 1]} did not converge. "
                                       f"info={max(r.iters)}, residual={r.resid}.")
                continue
            sim.avg_m = r.avg_m
        return out

    def __del__(self, _release=_lib.release):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            _release("cpm_batch_destroy", h)
