"""Benchmark of the GSPM LLG step on one MI355X.

Workload (the configuration the headline metric is quoted on): a 128^3 Kuhn
tetrahedral mesh (6 tets per cube, 12.6M tets, 2.1M DOFs), a rotated
ellipsoid defect with semi-axes (20h, 12h, 8h) and a seeded rotation,
magnetoelastic coupling on (c11 = 2.4e11, c12 = 1.4e11, c44 = 1.2e11 Pa,
lambda100 = 2.6e-4, lambda111 = -3e-5), crystal x axis along [111],
dt = 10 ps, alpha = 0.1, CG tolerance 1e-12 (every path; at 1e-10 the FP64
solves themselves leave m uncertain by ~1e-6 after 50 steps of this
strongly driven state).  Inputs are synthetic and seeded; every operator is
resident in HBM before the timed region.

Prints ONE JSON line: node-steps/s (DOFs x steps / wall time of the timed
steps) of the fastest mixed-precision path whose magnetization stays within
1e-7 of the FP64 path (FP64 CG with the reference's recurrence) after the
workload's 50 steps, the roofline of its dominant kernel and of its outer
elasticity product measured with HIP events, and a CPU baseline (the
NumPy/SciPy oracle on a bounded sample of the same workload).

    python bench.py [--steps K] [--warmup W] [--n 128] [--cpu-n 40] [--tol 1e-12]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
REF_STEPS = 50   # the single large instance runs 50 steps


def workload_config(n, tol=1e-12, seed=2024):
    from cupymag_amd.serving.workloads import headline_config
    return headline_config(n, tol=tol, seed=seed)


def initial_m(ndof, seed=7):
    from cupymag_amd.serving.workloads import initial_m as im
    return im(ndof, seed)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mode_kwargs(base, opts, mg_degree=2, mg_ratio=10.0):
    """GSPMSimulation options of a bench mode: '+warm' (warm start), '+xwarm'
    (warm start from the extrapolation of the last two solutions), '+mg'
    (multigrid for the elasticity, demag and Gauss-Seidel A1 solves; FP16
    level values in the mixed mode), '+mged' (multigrid for elasticity and
    demag only), '+mg32' (FP32 level values), '+seq' (the demag solve after
    the elasticity solve instead of beside it)."""
    mg = "mg" in opts or "mged" in opts or "mg32" in opts
    mgp = (32 if "mg32" in opts else 16) if base == "mixed" else 64
    warm = "extrapolate" if "xwarm" in opts else ("warm" in opts)
    return dict(warm_start=warm, precond="mg" if mg else None, mg_gs="mged" not in opts,
                mg_degree=mg_degree, mg_ratio=mg_ratio, mg_precision=mgp if mg else None,
                overlap="seq" not in opts)


def kernel_bytes(base, mg, mgp, nnz, rows, stencil=True, half=False, half32=False, half64=False):
    """Algorithmic bytes per launch of a mode's roofline kernel and its name
    (rocprofv3 spelling).  nnz, rows: of the elasticity operator A_el (3x3
    blocks, nb = rows / 3 block rows, nnz / 9 blocks of which nb diagonal).
    half: the level-0 FP16 form keeps half of the symmetric stencil (each
    off-diagonal block stored once, for the row and its mirror); half32 and
    half64: the FP32 and the FP64 forms of A_el do."""
    nb = rows // 3
    off = nnz / 9.0 - nb
    if base == "fp64":
        # CSR FP64: values + column indices + row pointers + p read + q written
        return ("k_cg_spmv (A_el, CSR FP64)", None, 12.0 * nnz + 4.0 * (rows + 1) + 16.0 * rows)
    if mg:
        # level-0 Chebyshev step of the V-cycle (r -= A d, d = c1 d + c2 D^-1 r,
        # x += d): off-diagonal blocks, diagonal blocks, d gathered, r and x
        # read and written, d written; the FP16 forms keep FP32 diagonal
        # blocks and invert them per node (no stored D^-1); the
        # pre-smoothing's step is MODE 0, the post-smoothing's MODE 4 (the
        # same bytes)
        s_ = 4 if base == "mixed" else 8
        t_ = "float" if s_ == 4 else "double"
        v_, vs_ = ("_Float16", 2) if mgp == 16 else (t_, s_)
        name = (f"k_mg_sweep<3,1,{t_},0|4,{v_}> (A_el level 0, Chebyshev step, SELL-64 3x3 blocks, stencil, "
                f"{v_} off-diagonal values)")
        roc = f"k_mg_sweep<3, 1, {t_}, 0, true, {v_}>"
        if mgp == 16:
            stored = off / 2 if half else off
            if half:
                name = name.replace("stencil,", "half-stored symmetric stencil,")
            return name, roc, stored * 9 * vs_ + nb * 36.0 + nb * 18.0 * s_
        return name, roc, off * 9 * vs_ + nb * 72.0 + nb * 27.0 * s_
    # SELL-64 3x3 blocks: off-diagonal blocks 9 values (+ 1 column index
    # unless the index-free stencil form is in use), diagonal blocks 9 FP64
    # values, p read, q written
    s_ = 8 if base == "fast64" else 4
    idx = 0 if stencil else 4
    hs = half32 if s_ == 4 else half64
    form = ("half-stored symmetric stencil" if hs else "stencil") if stencil else "indexed"
    name = f"k_sell_spmv<3,1,{'double' if s_ == 8 else 'float'}> (A_el, SELL-64 3x3 blocks, {form})"
    roc = (f"k_sell_spmv<3, 1, {'double' if s_ == 8 else 'float'}, true, false, false, "
           f"{'true' if stencil else 'false'}>")
    return name, roc, (off / 2 if hs else off) * (9 * s_ + idx) + nb * 72.0 + 2.0 * rows * s_


def outer_bytes(nnz, rows, half32=False):
    """Algorithmic bytes of the mixed multigrid PCG's outer product q = A p
    (k_pcg_spmv_rel, FP32 form): FP32 off-diagonal 3x3 blocks (each stored
    once when the FP32 form keeps half of the symmetric stencil), FP64
    compensated diagonal blocks, p read and q written in FP32."""
    nb = rows // 3
    off = nnz / 9.0 - nb
    return (off / 2 if half32 else off) * 36.0 + nb * 72.0 + 2.0 * rows * 4.0


def run_gpu(n, steps, warmup, steps64, modes, tol, ref_steps=REF_STEPS, rel_delta=0.3, autotune=True,
            rel_delta_gs=None, mg_degree=2, mg_ratio=10.0, matrix_free_deriv=False):
    """Run the plain FP64 path and the fast paths on the same operators and
    initial state; returns per-mode timings, CG iteration counts, the timed
    kernels and the magnetization after the steps the errors are taken at:
    the plain FP64 path's last step, the end of the timed run and
    ``ref_steps`` (the fast paths keep stepping, untimed, to ref_steps)."""
    import ctypes as C

    import torch

    from cupymag_amd._lib import check, lib
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh

    torch.cuda.set_device(0)
    cfg = workload_config(n, tol)
    P = from_dict(cfg)
    t0 = time.time()
    arrays = kuhn_mesh(n)
    S = SimulationOperators(P, mesh_arrays=arrays, keep_host=False, autotune=autotune, matrix_free_fel=True,
                            matrix_free_deriv=matrix_free_deriv)
    log(f"[bench] mesh {n}^3: {S.n_nodes} nodes, {S.n_elements} tets, {S.n_dof} DOFs, "
        f"{int(S.elements_np[:, -1].sum())} defect tets ({time.time() - t0:.1f}s)")
    t0 = time.time()
    S.csr_demag_and_deriv()
    log(f"[bench] host assembly + upload (CSR and SELL-64 forms){' + autotuning' if autotune else ''} "
        f"{time.time() - t0:.1f}s")
    for name, (sched, timings) in S.tuning.items():
        log(f"[bench] schedule {name}: {sched}" + ("" if timings is None else
            f" (best of {len(timings)}: {min(t['us'] for t in timings):.1f} us)"))
    A_el = S.csr_elasticity()[0]
    m0 = initial_m(S.n_dof)
    H = [P.Hext1, P.Hext2, P.Hext3]
    out = dict(n_dof=S.n_dof, n_nodes=S.n_nodes, n_elements=S.n_elements, modes={},
               schedules={k: v[0] for k, v in S.tuning.items()}, nnz={"A_el": A_el.nnz})
    out["nnz"].update(A_demag=S.csr_demag_and_deriv()[0].nnz, A1=S.csr_GS()[0].nnz, A2=S.csr_GS()[1].nnz)
    w64 = min(warmup, 1)
    end_fp64 = w64 + steps64
    snap_at = {end_fp64, warmup + steps, ref_steps}
    rows, nnz = A_el.shape[0], A_el.nnz
    for mode in modes:
        base, *opts = mode.split("+")
        kw = mode_kwargs(base, opts, mg_degree, mg_ratio)
        mg, mgp = kw["precond"] == "mg", kw["mg_precision"]
        if mg:
            t0 = time.time()
            S.multigrid(precision=mgp, degree=mg_degree, ratio=mg_ratio)
            log(f"[bench] multigrid hierarchies for {mode}: {time.time() - t0:.1f}s")
        if rel_delta_gs is not None:
            kw["rel_delta_gs"] = rel_delta_gs
        sim = GSPMSimulation(S, P, mode=base, rel_delta=rel_delta, **kw)
        ctx = sim.ctx
        sim.set_m(m0)
        w = w64 if base == "fp64" else warmup
        k_timed = steps64 if base == "fp64" else steps
        total = w + k_timed if base == "fp64" else max(w + k_timed, ref_steps)
        snaps = {}
        done = 0

        def one(tag):
            nonlocal done
            t = time.perf_counter()
            rep = sim.step(H, False)
            dt_ = time.perf_counter() - t
            done += 1
            if done in snap_at:
                snaps[done] = sim.get("M")
            log(f"[bench] {mode} {tag} step {done}: {dt_:.3f}s iters={rep.iters}")
            return rep, dt_

        for k in range(w):
            one("warmup")
        # time the elasticity operator's kernels only: the other solves run
        # as they do without profiling
        check(lib().cpm_prof_enable(ctx.h, 1))
        check(lib().cpm_prof_only(ctx.h, A_el.h))
        check(lib().cpm_mat_prof(A_el.h, None, None, 1))
        check(lib().cpm_mat_prof_outer(A_el.h, None, None, 1))
        torch.cuda.synchronize()
        lat, iters = [], []
        t_start = time.perf_counter()
        for k in range(k_timed):
            rep, dt_ = one("timed")
            lat.append(dt_)
            iters.append(rep.iters)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t_start
        ms, nl = C.c_double(), C.c_int64()
        check(lib().cpm_mat_prof(A_el.h, C.byref(ms), C.byref(nl), 0))
        ms2, nl2 = C.c_double(), C.c_int64()
        check(lib().cpm_mat_prof_outer(A_el.h, C.byref(ms2), C.byref(nl2), 0))
        check(lib().cpm_prof_enable(ctx.h, 0))
        while done < total:
            one("untimed")
        info = A_el.sell_info()
        half = bool(mg and mgp == 16 and S.multigrid(precision=mgp, degree=mg_degree,
                                                      ratio=mg_ratio)["A_el"].info()["half_stencil"][0])
        kname, rocname, bytes_launch = kernel_bytes(base, mg, mgp, nnz, rows, info["stencil"], half,
                                                    info["half32"], info["half64"])
        avg_s = ms.value / 1e3 / max(nl.value, 1)
        d = dict(elapsed=elapsed, steps=k_timed, lat=lat, iters=iters, snaps=snaps,
                 spmv=dict(kernel=kname, rocname=rocname, launches=nl.value, avg_us=avg_s * 1e6,
                           bytes=bytes_launch, gbs=bytes_launch / avg_s / 1e9 if nl.value else 0.0))
        if nl2.value:
            ob = outer_bytes(nnz, rows, info["half32"])
            avg2 = ms2.value / 1e3 / nl2.value
            d["outer"] = dict(kernel="k_pcg_spmv_rel<3,1,true> (A_el level 0, outer PCG product, FP32 3x3 blocks"
                                     + (", half-stored symmetric stencil" if info["half32"] else "")
                                     + ", FP64 compensated diagonal blocks, FP32 p and q)",
                              rocname="k_pcg_spmv_rel<3, 1, true>", launches=nl2.value, avg_us=avg2 * 1e6,
                              bytes=ob, gbs=ob / avg2 / 1e9)
        out["modes"][mode] = d
        del sim
    out["snap_steps"] = dict(fp64=end_fp64, timed=warmup + steps, ref=ref_steps)
    return out


PROFILES = os.path.join(ROOT, "profiles")


def profiled_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at 128^3 from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary under profiles/ (scaled by the
    copy-kernel calibration there); None for other kernels or sizes."""
    if n != 128 or not os.path.isdir(PROFILES):
        return None
    for fn in sorted(os.listdir(PROFILES)):
        if not fn.endswith(".md"):
            continue
        path = os.path.join(PROFILES, fn)
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch "):
                name, _, val = line[len("traffic_bytes_per_launch "):].strip().rpartition(" ")
                if name == kernel:
                    return dict(bytes=float(val), source=os.path.relpath(path, ROOT))
    return None


SOLVE_OPS = ("A_el", "A_demag", "A1", "A1", "A1", "A1", "A1", "A2", "A2", "A2")


def cpu_baseline(n_cpu, steps=1, tol=1e-12, n_full=None, iters_full=None, nnz_full=None):
    """The NumPy/SciPy oracle (one core: SciPy CSR matvec and NumPy
    elementwise are single-threaded) on an n_cpu^3 instance of the same
    workload, timed around the step loop only.  With the full size's CG
    iteration counts (the plain FP64 path's, which equal the oracle's to +-1)
    and operator sizes, the time per step at the full size is estimated by
    scaling the sample's time with the CG work sum_solves iterations x nnz
    (the solves are > 95% of the oracle's step)."""
    from oracle import fem, gspm, mesh as om, params
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = workload_config(n_cpu, tol)
    p = params.reduced(cfg)
    coords, tets = kuhn_mesh(n_cpu)
    tm = cfg["grid"]["tet_mesh"]
    shape = om.rotated_ellipsoid(tm["defect_parameters"], np.array(tm["rotation_matrix"]))
    flags = om.tet_defect_flags(coords, tets, shape)
    el = np.concatenate([tets, flags[:, None]], axis=1).astype(np.int32)
    gid = om.periodic_map(coords, om.tet_bounds(coords))
    ops = fem.build_operators(coords, el, gid, p)
    Avg = fem.VolumeAverage(coords, el, gid)
    st = gspm.init_state(ops, Avg, initial_m(ops["n"]))
    p["H_now"] = p["H"]
    logs = []
    t = time.perf_counter()
    for _ in range(steps):
        lg = []
        gspm.step(st, ops, Avg, p, log=lg)
        logs.append([it for _, it in lg])
    dt = time.perf_counter() - t
    out = dict(value=ops["n"] * steps / dt, unit="node-steps/s", cores=1, kind="port",
               sample=f"oracle (NumPy/SciPy FP64, CG tol {tol:g}) on the same workload scaled to {n_cpu}^3 Kuhn "
                      f"tets ({ops['n']} DOFs; defect semi-axes scaled with n), {steps} step(s) from the same "
                      f"seeded initial m, {dt:.1f}s on one host core",
               s_per_step=dt / steps)
    if iters_full and nnz_full:
        nnz_s = [ops[k].nnz for k in SOLVE_OPS]
        work_s = np.mean([sum(i * z for i, z in zip(its, nnz_s)) for its in logs])
        work_f = sum(i * nnz_full[k] for i, k in zip(iters_full, SOLVE_OPS))
        est = dt / steps * work_f / work_s
        out["estimate_full_size"] = dict(
            n=n_full, s_per_step=est, node_steps_per_s=n_full ** 3 / est,
            method=f"sample time per step x (sum over the 10 solves of CG iterations x operator nnz at {n_full}^3, "
                   f"iterations from the plain FP64 GPU path) / (the same sum in the sample, "
                   f"{work_s:.3g} vs {work_f:.3g})")
    return out


def copy_rate(mib=1024, reps=10):
    """Device-to-device copy rate of this GPU (GB/s, bytes read + written)
    over `reps` copies of `mib` MiB, timed with events: the practical
    streaming rate that the roofline kernel is also set against."""
    import torch
    src = torch.ones(mib << 17, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    del src, dst
    return 2.0 * (mib << 20) * reps / (ms / 1e3) / 1e9


def _build_instances(descs, ctx, workers=4, multigrid=False):
    """SimulationOperators of every instance description (mesh, host
    assembly, upload, SpMV schedules autotuned per size class), built on a
    few host threads; multigrid: also the
    multigrid hierarchies of the mixed mode (FP16 level operator values)
    where the mesh has them."""
    import concurrent.futures as cf

    import torch

    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh

    def one(d):
        torch.cuda.set_device(ctx.device)
        P = from_dict(d["cfg"])
        S = SimulationOperators(P, mesh_arrays=kuhn_mesh(d["n"]), ctx=ctx, keep_host=False, matrix_free_fel=True,
                                autotune=True)
        S.csr_demag_and_deriv()
        if multigrid:
            S.multigrid(precision=16)
        return P, S

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, descs))


def run_batch(args):
    """Many independent simulations of different sizes: stepped one after
    another with the single-simulation path, then all together as one
    batch (each solve one batched CG), eager and with graph replay."""
    import torch

    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.device import context
    from cupymag_amd.serving.workloads import batch_workload, initial_m as init_m

    torch.cuda.set_device(0)
    ctx = context(0)
    descs = batch_workload(args.count, seed=args.seed, nmin=args.nmin, nmax=args.nmax)
    t0 = time.time()
    use_mg = args.precond == "mg"
    built = _build_instances(descs, ctx, multigrid=use_mg)
    ndof = [S.n_dof for _, S in built]
    # members whose mesh has a multigrid hierarchy use it (n even, coarsest <= 10^3)
    pc = ["mg" if use_mg and S.multigrid(precision=16) is not None else None for _, S in built]
    log(f"[bench] {len(built)} instances, n = {min(d['n'] for d in descs)}..{max(d['n'] for d in descs)}, "
        f"{sum(ndof)} DOFs in total, {sum(p is not None for p in pc)} with multigrid, "
        f"set up in {time.time() - t0:.1f}s")
    total = float(sum(ndof))
    res = {}

    def fresh():
        sims = []
        for (P, S), d, prc in zip(built, descs, pc):
            sim = GSPMSimulation(S, P, mode="mixed", rel_delta=args.rel_delta, precond=prc)
            sim.set_m(init_m(S.n_dof, d["m_seed"]))
            sims.append(sim)
        return sims

    H = [[P.Hext1, P.Hext2, P.Hext3] for P, _ in built]
    # one after another (single-simulation path)
    sims = fresh()
    for k in range(args.warmup):
        for sim, h in zip(sims, H):
            sim.step(h, False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(args.steps):
        for sim, h in zip(sims, H):
            sim.step(h, False)
        log(f"[bench] sequential step {k} done")
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    m_seq = [s.get("M") for s in sims]
    res["sequential"] = dict(node_steps_per_s=total * args.steps / el, ms_per_step=el * 1e3 / args.steps)
    del sims
    for graphs in (False, True):
        name = "batched+graph" if graphs else "batched"
        sims = fresh()
        b = GSPMBatch(ctx, graphs=graphs)
        for s in sims:
            b.add(s)
        for k in range(args.warmup):
            b.step(H, [False] * len(sims))
        torch.cuda.synchronize()
        lat = []
        t = time.perf_counter()
        for k in range(args.steps):
            t1 = time.perf_counter()
            reps = b.step(H, [False] * len(sims))
            lat.append(time.perf_counter() - t1)
            worst = max(range(len(reps)), key=lambda j: max(reps[j].iters))
            log(f"[bench] {name} step {k}: {lat[-1]:.3f}s max iters {max(reps[worst].iters)} "
                f"(n = {descs[worst]['n']}, iterations {list(reps[worst].iters)})")
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        err = max(float((s.get("M") - m).abs().max().item()) for s, m in zip(sims, m_seq))
        res[name] = dict(node_steps_per_s=total * args.steps / el, ms_per_step=el * 1e3 / args.steps,
                         latency_ms=dict(p50=float(np.percentile(np.array(lat) * 1e3, 50)),
                                         p99=float(np.percentile(np.array(lat) * 1e3, 99))),
                         max_abs_m_diff_vs_sequential=err)
        del b, sims
    head = max(("batched", "batched+graph"), key=lambda k: res[k]["node_steps_per_s"])
    out = {
        "metric": "LLG node-steps/s", "value": res[head]["node_steps_per_s"], "unit": "node-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res[head]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 CG in f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"batch of {len(built)} independent GSPM simulations, Kuhn n^3 meshes with n "
                               f"log-uniform in [{args.nmin},{args.nmax}], sphere defects, random Hext, "
                               f"magnetoelastic on in every other one, mixed precision, dt 10 ps, alpha 0.1"
                               + (", multigrid-preconditioned elasticity/demag solves where the mesh has a "
                                  "hierarchy" if use_mg else ""),
                   "path": head, "instances": len(built), "n_dof_total": int(total),
                   "n": [d["n"] for d in descs], "precond": args.precond,
                   "members_with_multigrid": sum(p is not None for p in pc)},
        "speedup_vs_sequential": res[head]["node_steps_per_s"] / res["sequential"]["node_steps_per_s"],
        "paths": res,
    }
    print(json.dumps(out), flush=True)


def accuracy_config(seed=2025):
    """The 64^3 accuracy workload: Kuhn tets with node jitter <= 0.15 h, a
    spherical defect of radius 8 h at the centre, magnetoelastic coupling on,
    dt 5 ps, alpha 0.5, tol 1e-10."""
    from cupymag_amd.serving.workloads import instance_config
    cfg = instance_config((5e4, 0.0, 0.0), me=True, defect_shape="Sphere", defect_params=8.0, dt=5e-12, alpha=0.5)
    return cfg


def run_accuracy(args):
    """Per-field error of the fast paths against the plain FP64 path over a
    long run (default 100 steps): m, u, E and Htilde after every 10 steps."""
    import torch

    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh

    torch.cuda.set_device(0)
    n = args.n if args.n != 128 else 64
    P = from_dict(accuracy_config())
    S = SimulationOperators(P, mesh_arrays=kuhn_mesh(n, jitter=0.15, seed=3), keep_host=False, autotune=True,
                            matrix_free_fel=True)
    S.csr_demag_and_deriv()
    m0 = initial_m(S.n_dof)
    H = [P.Hext1, P.Hext2, P.Hext3]
    every = 10
    modes = [m for m in args.modes.split(",") if m] if args.modes != "fp64,fast64,mixed,fast64+mg,mixed+mg" else \
        ["fp64", "fast64", "mixed", "mixed+mg"]
    snaps, timing, iters = {}, {}, {}
    for mode in modes:
        base, *opts = mode.split("+")
        sim = GSPMSimulation(S, P, mode=base, **mode_kwargs(base, opts, args.mg_degree, args.mg_ratio))
        sim.set_m(m0)
        snaps[mode] = []
        t = time.perf_counter()
        use_init = False
        for k in range(args.steps):
            rep = sim.step(H, use_init)
            use_init = rep.sumsq < P.use_init_factor * S.n_dof * P.LLG_accuracy_factor
            if (k + 1) % every == 0:
                snaps[mode].append({f: sim.get(f).cpu() for f in ("M", "UEL", "E", "HTILDE")})
        torch.cuda.synchronize()
        timing[mode] = (time.perf_counter() - t) / args.steps
        iters[mode] = rep.iters
        log(f"[bench] accuracy {mode}: {args.steps} steps, {timing[mode] * 1e3:.1f} ms/step, last iters {rep.iters}")
        del sim
    ref = snaps["fp64"]
    errors = {}
    for mode in modes:
        if mode == "fp64":
            continue
        rows = []
        for k, (a, b) in enumerate(zip(snaps[mode], ref)):
            rel = {f: float((a[f] - b[f]).abs().max() / b[f].abs().max().clamp_min(1e-300)) for f in ("UEL", "E", "HTILDE")}
            rows.append(dict(step=(k + 1) * every, m_max_abs=float((a["M"] - b["M"]).abs().max()),
                             u_rel=rel["UEL"], E_rel=rel["E"], Htilde_rel=rel["HTILDE"]))
        errors[mode] = rows
    out = {"metric": "max per-field error vs plain FP64", "unit": "abs (m) / rel (u, E, Htilde)",
           "workload": f"{n}^3 Kuhn tets, node jitter 0.15 h, sphere defect r = 8 h, magnetoelastic on, "
                       f"dt 5 ps, alpha 0.5, tol 1e-10, {args.steps} steps",
           "n_dof": S.n_dof, "ms_per_step": {k: v * 1e3 for k, v in timing.items()}, "last_iters": iters,
           "errors": errors}
    print(json.dumps(out), flush=True)


def run_serve(args):
    """Hysteresis-sweep requests arriving as a Poisson stream, served by
    the continuous-batching server; p50/p99 per-point and per-loop latency
    and aggregate node-steps/s."""
    import torch

    from cupymag_amd.device import context
    from cupymag_amd.serving.server import SweepServer
    from cupymag_amd.serving.workloads import sweep_requests

    torch.cuda.set_device(0)
    ctx = context(0)
    reqs = sweep_requests(args.count, rate=args.rate, seed=args.seed, nmin=args.nmin, nmax=args.nmax,
                          points=args.points, steps_per_point=args.steps_per_point)
    srv = SweepServer(ctx, hbm_budget=args.hbm_budget_gb * 1e9, mode="mixed", graphs=args.graphs,
                      precond=args.precond if args.precond != "none" else None)
    results, st = srv.run(reqs, log=log)
    out = {
        "metric": "LLG node-steps/s", "value": st["node_steps_per_s"], "unit": "node-steps/s", "n_gpus": 1,
        "steps": st["batch_steps"], "warmup": 0, "ms_per_step": st["wall_s"] * 1e3 / max(st["batch_steps"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 CG in f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"hysteresis-sweep serving: {args.count} requests, Poisson arrivals "
                               f"{args.rate}/s, Kuhn n^3 with n log-uniform in [{args.nmin},{args.nmax}], "
                               f"random rotated-ellipsoid defect, magnetoelastic on, {args.points} field points "
                               f"in [-1e5, 1e5] A/m x {args.steps_per_point} steps, HBM budget "
                               f"{args.hbm_budget_gb} GB, mixed precision"
                               + (", multigrid where the mesh has a hierarchy" if args.precond == "mg" else ""),
                   "requests": args.count, "n": [r["n"] for r in reqs], "precond": args.precond},
        "latency_s": {"per_point": st["point_latency_s"], "per_loop": st["loop_latency_s"],
                      "queue_delay": st["queue_delay_s"]},
        "serving": st,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("gspm", "batch", "serve", "accuracy"), default="gspm",
                    help="gspm: one 128^3 simulation (the headline); batch: many independent simulations; "
                         "serve: the hysteresis-sweep server; accuracy: per-field error of the fast paths "
                         "against plain FP64 over a long run (64^3 jittered, --steps 100)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed steps of the headline (mixed) path")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--steps-fp64", type=int, default=2, help="timed steps of the plain FP64 path")
    ap.add_argument("--tol", type=float, default=1e-12, help="CG tolerance of every path (headline workload)")
    ap.add_argument("--ref-steps", type=int, default=REF_STEPS,
                    help="the fast paths step (untimed) to this many steps; errors against fast64 are taken there")
    ap.add_argument("--modes", default="fp64,fast64,mixed,fast64+mg,mixed+mg,mixed+mg+xwarm",
                    help="comma list of fp64, fast64, mixed; '+warm' appended to a fast mode starts every CG "
                         "from the previous step's solution, '+xwarm' from the linear extrapolation of the last "
                         "two steps' solutions, '+mg' preconditions the elasticity, demag and "
                         "Gauss-Seidel (A1) solves with the geometric multigrid, '+mged' only the elasticity and "
                         "demag solves; the mixed mode's V-cycle stores the level operators' values in FP16, "
                         "'+mg32' in FP32")
    ap.add_argument("--rel-delta", type=float, default=0.3)
    ap.add_argument("--mg-degree", type=int, default=2, help="multigrid: Chebyshev smoother degree")
    ap.add_argument("--mg-ratio", type=float, default=10.0, help="multigrid: Chebyshev eigenvalue ratio")
    ap.add_argument("--rel-delta-gs", type=float, default=None,
                    help="reliable-update factor of the Gauss-Seidel solves (default 1e-4)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default SpMV launch schedules")
    ap.add_argument("--csr-deriv", action="store_true",
                    help="strain, Htilde and the demag right-hand side from the Fx/Fy/Fz CSR matrices instead of "
                         "the matrix-free element terms (a 27-point stencil on the regular 128^3 lattice)")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--cpu-n", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this size (calibrates rocprof byte counters)")
    ap.add_argument("--count", type=int, default=64, help="batch / serve: instances or requests")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--nmin", type=int, default=16)
    ap.add_argument("--nmax", type=int, default=96)
    ap.add_argument("--rate", type=float, default=20.0, help="serve: request arrivals per second")
    ap.add_argument("--points", type=int, default=41, help="serve: field points per sweep")
    ap.add_argument("--steps-per-point", type=int, default=200)
    ap.add_argument("--hbm-budget-gb", type=float, default=240.0)
    ap.add_argument("--graphs", action="store_true", help="serve: graph replay of the CG iterations")
    ap.add_argument("--precond", choices=("none", "mg"), default="none",
                    help="batch / serve: multigrid-preconditioned elasticity and demag solves for the members "
                         "whose mesh has a hierarchy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the GSPM step of one simulation is not sharded across GPUs "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib:
        import torch
        src = torch.ones(args.copy_probe_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "accuracy":
        return run_accuracy(args)
    modes = [m for m in args.modes.split(",") if m]
    steps64 = args.steps_fp64
    g = run_gpu(args.n, args.steps, args.warmup, steps64, modes, args.tol, ref_steps=args.ref_steps,
                rel_delta=args.rel_delta, autotune=not args.no_autotune, rel_delta_gs=args.rel_delta_gs,
                mg_degree=args.mg_degree, mg_ratio=args.mg_ratio, matrix_free_deriv=not args.csr_deriv)
    iters_full = g["modes"]["fp64"]["iters"][-1] if "fp64" in g["modes"] else None
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n, args.cpu_steps, args.tol, args.n, iters_full, g["nnz"])
    st = g["snap_steps"]
    ref = g["modes"].get("fast64")
    per_mode = {}
    for mode, r in g["modes"].items():
        d = dict(node_steps_per_s=g["n_dof"] * r["steps"] / r["elapsed"], ms_per_step=r["elapsed"] * 1e3 / r["steps"],
                 cg_iters_last_step=r["iters"][-1], kernel=r["spmv"]["kernel"], kernel_GBs=r["spmv"]["gbs"],
                 kernel_frac_of_hbm_peak=r["spmv"]["gbs"] / HBM_PEAK_GBS, kernel_avg_us=r["spmv"]["avg_us"])
        if ref is not None and mode != "fast64":
            for key, k in (("err_vs_fast64_end_of_timed_run", st["timed"]), ("err_vs_fast64_after_ref_steps", st["ref"])):
                if k in r["snaps"] and k in ref["snaps"]:
                    d[key] = float((r["snaps"][k] - ref["snaps"][k]).abs().max().item())
        if "outer" in r:
            d["outer_product_GBs"] = r["outer"]["gbs"]
        per_mode[mode] = d
    if "fp64" in per_mode and ref is not None:
        k = st["fp64"]
        per_mode["fast64"]["err_vs_fp64_after_fp64_steps"] = float(
            (ref["snaps"][k] - g["modes"]["fp64"]["snaps"][k]).abs().max().item())
    if "fp64" in per_mode:
        for d in per_mode.values():
            d["speedup_vs_fp64"] = d["node_steps_per_s"] / per_mode["fp64"]["node_steps_per_s"]
    # headline: the fastest mixed-precision path whose magnetization stays
    # within 1e-7 of the FP64 path after the workload's steps
    mixed = [m for m in modes if m.split("+")[0] == "mixed"]
    ok = [m for m in mixed if per_mode[m].get("err_vs_fast64_after_ref_steps", 1.0) <= 1e-7] if ref else mixed
    head = max(ok or mixed or modes, key=lambda m: per_mode[m]["node_steps_per_s"])
    hm = g["modes"][head]
    value = g["n_dof"] * hm["steps"] / hm["elapsed"]
    lat = np.array(hm["lat"]) * 1e3
    sp = hm["spmv"]
    traffic = profiled_traffic(sp["rocname"], args.n) if sp.get("rocname") else None
    copy_gbs = copy_rate()

    def gain(a, b):
        return per_mode[a]["node_steps_per_s"] / per_mode[b]["node_steps_per_s"] if a in per_mode and b in per_mode \
            else None

    roof_outer = None
    if "outer" in hm:
        o = hm["outer"]
        tr = profiled_traffic(o["rocname"], args.n)
        roof_outer = {"bound": "hbm", "kernel": o["kernel"], "achieved": o["gbs"], "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": o["gbs"] / HBM_PEAK_GBS, "traffic": tr["bytes"] if tr else None,
                      "traffic_source": tr["source"] if tr else None, "avg_launch_us": o["avg_us"],
                      "bytes_per_launch": o["bytes"], "launches": o["launches"]}
    dtypes = {"fp64": "f64", "fast64": "f64", "mixed": "f32 CG in f64 refinement",
              "mixed+mg": "f16 level-operator values and f32 vectors in the V-cycle, f32 outer products "
                          "and search directions, f64 iterate, residual and residual replacements",
              "fast64+mg": "f64"}
    dtypes["mixed+mg+xwarm"] = dtypes["mixed+mg"]
    out = {
        "metric": "LLG node-steps/s",
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": hm["steps"],
        "warmup": args.warmup,
        "ms_per_step": hm["elapsed"] * 1e3 / hm["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes.get(head, head),
        "data": "synthetic",
        "config": {"workload": f"GSPM LLG step, {args.n}^3 Kuhn tet mesh, rotated-ellipsoid defect, "
                               f"magnetoelastic on, rot111, dt 10 ps, alpha 0.1, CG tol {args.tol:g}",
                   "path": head, "n_dof": g["n_dof"], "nodes": g["n_nodes"], "elements": g["n_elements"],
                   "spmv_schedules": g["schedules"]},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "accuracy": {"reference": "fast64: FP64 CG with the reference's recurrence and stopping rule "
                                  "(max |m - m_plain_fp64| after the plain path's steps in modes.fast64)",
                     "max_abs_m_err_end_of_timed_run": per_mode[head].get("err_vs_fast64_end_of_timed_run"),
                     "at_step": st["timed"],
                     "max_abs_m_err_after_ref_steps": per_mode[head].get("err_vs_fast64_after_ref_steps"),
                     "ref_steps": st["ref"], "bound": 1e-7},
        "speedup_vs_fp64": gain(head, "fp64"),
        "gains": {"precision_mixed_vs_fast64": gain("mixed", "fast64"),
                  "precision_mixed+mg_vs_fast64+mg": gain("mixed+mg", "fast64+mg"),
                  "multigrid_mixed+mg_vs_mixed": gain("mixed+mg", "mixed"),
                  "sell_and_3rhs_fast64_vs_fp64": gain("fast64", "fp64"),
                  "note": "the precision gains compare paths with the same solver; the multigrid preconditioner "
                          "is an algorithm the reference does not have"},
        "modes": per_mode,
        "roofline": {"bound": "hbm", "kernel": sp["kernel"], "achieved": sp["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": sp["gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "avg_launch_us": sp["avg_us"], "bytes_per_launch": sp["bytes"], "launches": sp["launches"],
                     "copy_GBs": copy_gbs, "frac_of_copy": sp["gbs"] / copy_gbs},
        "roofline_outer_product": roof_outer,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
