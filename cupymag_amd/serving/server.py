"""Continuous-batching server for hysteresis-sweep requests.

A request is one hysteresis sweep on its own mesh: a list of external-field
points, at each of which the LLG step runs until the reference's
convergence rule fires (sumsq < LLG_accuracy after llg_min_step steps,
Micromagnetics.py:309-335) or a per-point step budget is spent.  The
server

* admits arrived requests in arrival order while their estimated device
  memory fits the HBM budget (and a member limit),
* prepares an admitted request (mesh, host assembly, upload, initial m) on
  a worker thread while the GPU keeps stepping the others,
* adds it to the running batch at the next step boundary, so requests that
  arrive while others are on the GPU join as memory frees up,
* steps every member of the batch with one ``GSPMBatch.step`` call per
  step (each solve a batched CG over all members), keeping m, u, the CG
  state and the sweep's control (field points, the convergence rule,
  use_init: GSPMSimulation.set_control) of every member resident on the
  device,
* retires a request when its last point is done and frees its memory,
* contains failures to their request: a request over the budget is
  rejected, and one whose preparation raises or whose CG does not converge
  (the reference's RuntimeError, linear_solvers.py:64-74) is retired with
  the error text while the others continue.

A member's arithmetic does not depend on the batch (``GSPMBatch``), so a
request's results are the same alone and in any mix.  ``run`` returns the
per-request results and the latency / throughput statistics: per-point
latency (wall time of a field point), per-loop latency (arrival to the end
of the sweep), queueing delay and aggregate node-steps/s.
"""
import concurrent.futures as cf
import contextlib
import os
import time
from collections import deque

import numpy as np

from ..core.gspm import SYSTEMS
from .workloads import initial_m

# device bytes per DOF of a member (operators in CSR and SELL-64 forms with
# the matrix-free F_el, the step's state vectors, the batched-CG workspace
# of the mixed mode; on Kuhn meshes the half FP64 copies of the symmetric
# SELL forms: 7 off-diagonal values of A_demag, A1 and A2, 7 3x3 blocks of
# A_el)
_BYTES_PER_DOF = {False: 2200.0 + 170.0, True: 5700.0 + 680.0}
# FP32 multigrid hierarchies per fine DOF (demag, Gauss-Seidel A1 with three
# columns, elasticity): coarse operators (1/8 + 1/64 + ... of the fine SELL
# forms, with their half FP64 copies), level vectors, the outer CG vectors
# (r, p, q)
_MG_BYTES_PER_DOF = {False: 500.0 + 20.0, True: 1200.0 + 90.0}


# FP64 hierarchies (the fast64 mode) add per fine DOF: the FP64 level vectors
# and coarse forms, and the FP64 copy of the level-0 periodic stiffness that
# the floating hierarchies keep (demag: 14 off-diagonal + 1 diagonal value;
# elasticity: 14 off-diagonal 3x3 blocks + the diagonal block)
_MG64_EXTRA_PER_DOF = {False: 500.0 + 120.0, True: 1200.0 + 1080.0 + 120.0}


def estimate_bytes(n_dof, me, multigrid_n=None, mode="mixed"):
    """Device memory of one member with ``n_dof`` DOFs (upper estimate);
    multigrid_n: cells per dimension when the member runs the multigrid
    solves (adds the hierarchies and their dense coarsest inverses: FP32 in
    the mixed mode, FP64 otherwise)."""
    b = 1.15 * _BYTES_PER_DOF[bool(me)] * n_dof
    if multigrid_n:
        from ..mesh.coarsen import hierarchy_sizes
        from ..solvers.multigrid import scalar_levels
        sizes = hierarchy_sizes(int(multigrid_n))
        nc = sizes[-1] ** 3                 # coarsest nodes of the elasticity hierarchy
        ns = scalar_levels(sizes)[-1] ** 3  # ... of the scalar ones (demag, A1)
        ts = 4.0 if mode == "mixed" else 8.0
        b += _MG_BYTES_PER_DOF[bool(me)] * n_dof + ts * (2 * ns ** 2 + (9 * nc ** 2 if me else 0))
        if mode != "mixed":
            b += _MG64_EXTRA_PER_DOF[bool(me)] * n_dof
    return int(b) + (64 << 20)


def _blas_threads(n):
    """Limit the BLAS / OpenMP thread pools of this process to n threads
    inside the block (threadpoolctl when importable, else no limit)."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        return contextlib.nullcontext()
    return threadpool_limits(limits=n)


def multigrid_ok(n, coarsest_max=8):
    """Whether an n^3 Kuhn mesh has a multigrid hierarchy (the rule of
    solvers.multigrid.build_hierarchies)."""
    return int(n) > coarsest_max


def _failed(req, t, error, t_admit=None, points=(), steps=0):
    """Result record of a request that was rejected or failed: the error
    text and whatever field points it finished."""
    return dict(rid=req["rid"], n=req["n"], failed=True, error=error, points=list(points), steps=steps, m=None,
                loop_latency=t - req["arrival"], queue_delay=None if t_admit is None else t_admit - req["arrival"],
                point_latency=[])


class SweepJob:
    """An admitted request: its simulation and the hysteresis control."""

    def __init__(self, req, sim, P, n_dof, est, t_arrival, t_admit):
        self.req = req
        self.rid = req["rid"]
        self.sim = sim
        self.P = P
        self.n_dof = n_dof
        self.est = est
        self.fields = np.asarray(req["fields"], dtype=np.float64)
        self.fields_red = self.fields / P.ms
        self.steps_per_point = int(req["steps_per_point"])
        self.llg_accuracy = n_dof * P.LLG_accuracy_factor
        self.llg_min_step = 10 if P.ME else 100
        self.point = 0
        self.nstep = 0
        self.steps = 0
        self.use_init = False
        self.points = []        # (H [A/m], avg_m, steps) per finished point
        self.point_lat = []
        self.t_arrival = t_arrival
        self.t_admit = t_admit
        self.t_point = None
        self.t_done = None

    def attach_control(self):
        """The sweep's control on the device: the field list, the per-point
        step cap and the reference's convergence rule and use_init."""
        self.sim.set_control(fields=self.fields_red, steps_per_point=self.steps_per_point,
                             llg_accuracy=self.llg_accuracy, initial_factor=self.P.use_init_factor,
                             llg_min_step=self.llg_min_step, stop_value=-2.0)

    def advance(self, rep, now):
        """Account one step from its report; returns True when the sweep is
        finished."""
        self.nstep += 1
        self.steps += 1
        if rep.point_done:
            self.points.append((self.fields[self.point].copy(), np.array(rep.avg_m), self.nstep))
            self.point_lat.append(now - self.t_point)
            self.t_point = now
            self.point += 1
            self.nstep = 0
        return rep.state == 2


class SweepServer:
    def __init__(self, ctx=None, hbm_budget=240e9, mode="mixed", graphs=False, max_members=256, prep_workers=2,
                 rel_delta=0.3, keep_m=True, precond=None, autotune=True):
        """precond="mg": requests whose mesh has a multigrid hierarchy run
        multigrid-preconditioned elasticity and demag solves (batched over
        those members); the others run plain CG.  autotune: each request's
        operators take the SpMV schedule of their size class, timed on the
        device the first time a class is seen (cupymag_amd.autotune)."""
        if precond not in (None, "mg"):
            raise ValueError(f"unknown preconditioner {precond!r} (None or 'mg')")
        from ..device import context
        self.ctx = ctx or context()
        self.hbm_budget = float(hbm_budget)
        self.mode = mode
        self.graphs = graphs
        self.max_members = int(max_members)
        self.prep_workers = int(prep_workers)
        self.rel_delta = rel_delta
        self.keep_m = keep_m
        self.precond = precond
        self.autotune = bool(autotune)

    def _uses_mg(self, n):
        return self.precond == "mg" and multigrid_ok(n)

    def _prepare(self, req):
        import torch

        from ..core.assemble_simulation_operators import SimulationOperators
        from ..core.gspm import GSPMSimulation
        from ..core.parameters import from_dict
        from ..mesh.kuhn import kuhn_mesh
        torch.cuda.set_device(self.ctx.device)
        t0 = time.perf_counter()
        P = from_dict(req["cfg"])
        arr = kuhn_mesh(req["n"])
        # the preparing threads share the host with each other and with the
        # stepping loop: few native threads each (assembly, BLAS)
        nth = max(1, (os.cpu_count() or 8) // (2 * self.prep_workers))
        with _blas_threads(nth):
            S = SimulationOperators(P, mesh_arrays=arr, ctx=self.ctx, keep_host=False, matrix_free_fel=True,
                                    nthreads=nth, autotune=self.autotune)
            S.csr_demag_and_deriv()
            t1 = time.perf_counter()
            sim = GSPMSimulation(S, P, mode=self.mode, rel_delta=self.rel_delta,
                                 precond="mg" if self._uses_mg(req["n"]) else None)
        t2 = time.perf_counter()
        f0 = np.asarray(req["fields"][0], dtype=np.float64)
        direction = f0 if np.linalg.norm(f0) > 0 else (1.0, 0.0, 0.0)
        sim.set_m(initial_m(S.n_dof, req["m_seed"], direction=direction))
        self.ctx.sync()
        # host-side preparation times (operators, simulation incl. hierarchies, initial state)
        sim.prep_s = (t1 - t0, t2 - t1, time.perf_counter() - t2)
        return sim, P, S.n_dof

    def run(self, requests, log=None):
        """Serve ``requests`` (dicts of workloads.sweep_requests, arrival
        times in seconds from the start); returns (results, stats)."""
        from ..core.batch import GSPMBatch
        say = log or (lambda *a: None)
        pending = deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = deque()
        preparing = {}
        active = []
        used = 0
        results = {}
        batch = GSPMBatch(self.ctx, graphs=self.graphs)
        node_steps = 0
        max_members = 0
        steps_run = 0
        pool = cf.ThreadPoolExecutor(max_workers=self.prep_workers)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        try:
            while pending or waiting or preparing or active:
                t = now()
                while pending and pending[0]["arrival"] <= t:
                    waiting.append(pending.popleft())
                # admission: arrival order, HBM budget, member limit
                while waiting:
                    req = waiting[0]
                    me = bool(req["cfg"]["physics"]["magnetoelastic_coupling"])
                    est = estimate_bytes(req["n"] ** 3, me, req["n"] if self._uses_mg(req["n"]) else None,
                                         self.mode)
                    if est > self.hbm_budget:
                        # rejected alone: the others keep being served
                        waiting.popleft()
                        results[req["rid"]] = _failed(req, t, f"needs ~{est / 1e9:.1f} GB, over the HBM budget "
                                                              f"of {self.hbm_budget / 1e9:.1f} GB")
                        say(f"[serve] t={t:.2f}s reject request {req['rid']}: {results[req['rid']]['error']}")
                        continue
                    if used + est > self.hbm_budget or len(active) + len(preparing) >= self.max_members:
                        break
                    waiting.popleft()
                    used += est
                    preparing[pool.submit(self._prepare, req)] = (req, est, t)
                    say(f"[serve] t={t:.2f}s admit request {req['rid']} (n={req['n']}, ~{est / 1e9:.2f} GB)")
                # prepared requests join the batch at this step boundary
                for fut in [f for f in preparing if f.done()]:
                    req, est, t_adm = preparing.pop(fut)
                    try:
                        sim, P, n_dof = fut.result()
                    except Exception as e:  # noqa: BLE001 -- any preparation error fails this request only
                        used -= est
                        results[req["rid"]] = _failed(req, now(), f"preparation failed: {type(e).__name__}: {e}",
                                                      t_admit=t_adm)
                        say(f"[serve] t={now():.2f}s request {req['rid']} failed: {results[req['rid']]['error']}")
                        continue
                    job = SweepJob(req, sim, P, n_dof, est, req["arrival"], t_adm)
                    job.attach_control()
                    job.t_point = now()
                    batch.add(sim)
                    active.append(job)
                    pt = getattr(sim, "prep_s", (0.0, 0.0, 0.0))
                    say(f"[serve] t={now():.2f}s request {job.rid} joins ({len(active)} members; prepared in "
                        f"{sum(pt):.2f}s: operators {pt[0]:.2f}s, simulation {pt[1]:.2f}s, state {pt[2]:.2f}s)")
                if not active:
                    if preparing:
                        cf.wait(list(preparing), timeout=0.05, return_when=cf.FIRST_COMPLETED)
                    elif pending:
                        time.sleep(max(0.0, min(0.05, pending[0]["arrival"] - now())))
                    continue
                max_members = max(max_members, len(active))
                # field, use_init and the point control of every member live on the device
                reps = batch.step(raise_on_fail=False)
                steps_run += 1
                t = now()
                finished, broken = [], []
                for job, rep in zip(active, reps):
                    node_steps += job.n_dof
                    if rep.failed:
                        # the reference raises here (linear_solvers.py:64-74); the server
                        # fails this request and keeps stepping the others
                        broken.append((job, f"Error! CG for {SYSTEMS[rep.failed - 1]} did not converge. "
                                            f"info={max(rep.iters)}, residual={rep.resid}."))
                    elif job.advance(rep, t):
                        finished.append(job)
                for job, err in broken:
                    batch.remove(job.sim)
                    active.remove(job)
                    used -= job.est
                    results[job.rid] = _failed(job.req, t, err, t_admit=job.t_admit, points=job.points,
                                               steps=job.steps + 1)
                    say(f"[serve] t={t:.2f}s request {job.rid} failed at point {job.point}: {err}")
                    job.sim = None
                for job in finished:
                    job.t_done = t
                    batch.remove(job.sim)
                    active.remove(job)
                    used -= job.est
                    results[job.rid] = dict(
                        rid=job.rid, n=job.req["n"], n_dof=job.n_dof, points=job.points, steps=job.steps,
                        m=job.sim.get("M").cpu().numpy() if self.keep_m else None,
                        loop_latency=job.t_done - job.t_arrival, queue_delay=job.t_admit - job.t_arrival,
                        point_latency=list(job.point_lat))
                    say(f"[serve] t={t:.2f}s request {job.rid} done ({job.steps} steps, "
                        f"{len(active)} members left)")
                    job.sim = None
        finally:
            pool.shutdown(wait=True)
        wall = now()
        ok = [r for r in results.values() if not r.get("failed")]
        pl = np.concatenate([np.asarray(r["point_latency"]) for r in ok]) if ok else np.zeros(0)
        ll = np.asarray([r["loop_latency"] for r in ok])
        qd = np.asarray([r["queue_delay"] for r in ok])
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else None  # noqa: E731
        stats = dict(requests=len(ok), failed=len(results) - len(ok), wall_s=wall, batch_steps=steps_run, node_steps=node_steps,
                     node_steps_per_s=node_steps / wall if wall > 0 else 0.0, max_members=max_members,
                     point_latency_s=dict(p50=pct(pl, 50), p99=pct(pl, 99)),
                     loop_latency_s=dict(p50=pct(ll, 50), p99=pct(ll, 99)),
                     queue_delay_s=dict(p50=pct(qd, 50), p99=pct(qd, 99)))
        return results, stats
