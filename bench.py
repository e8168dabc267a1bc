"""Benchmark of the GSPM LLG step on one MI355X.

Workload (the configuration the headline metric is quoted on): a 128^3 Kuhn
tetrahedral mesh (6 tets per cube, 12.6M tets, 2.1M DOFs), a rotated
ellipsoid defect with semi-axes (20h, 12h, 8h) and a seeded rotation,
magnetoelastic coupling on (c11 = 2.4e11, c12 = 1.4e11, c44 = 1.2e11 Pa,
lambda100 = 2.6e-4, lambda111 = -3e-5), crystal x axis along [111],
dt = 10 ps, alpha = 0.1, CG tolerance 1e-10.  Inputs are synthetic and
seeded; every operator is resident in HBM before the timed region.

Prints ONE JSON line: node-steps/s (DOFs x steps / wall time of the timed
steps), the CG SpMV roofline of the dominant operator measured with HIP
events, and a CPU baseline (the NumPy/SciPy oracle on a bounded sample of
the same workload).

    python bench.py [--steps K] [--warmup W] [--n 128] [--cpu-n 20]
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


def workload_config(n, seed=2024):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    s = n / 128.0
    cfg = {
        "physics": {"magnetic": {"ms": 1.2e6, "exchange_stiffness": 1.5e-11, "anisotropy_k1": -1e4},
                    "elastic": {"c11": 2.4e11, "c12": 1.4e11, "c44": 1.2e11},
                    "magnetostriction": {"lambda100": 2.6e-4, "lambda111": -3e-5},
                    "demagnetization": {"n_x": 1 / 3, "n_y": 1 / 3, "n_z": 1 / 3},
                    "external_field": {"hext1": 5e4, "dhext1": 0.0},
                    "external_stress": {},
                    "rotation_111": True, "magnetoelastic_coupling": True,
                    "characteristic_length": 2e-9},
        "grid": {"type": "Tet", "tet_mesh": {"mesh_file": None, "defect_shape": "RotatedEllipsoid",
                                             "defect_parameters": [20.0 * s, 12.0 * s, 8.0 * s],
                                             "rotation_matrix": q.tolist()}},
        "simulation": {"precision": "DP",
                       "llg": {"timestep": 1e-11, "alpha_damping": 0.1, "llg_accuracy_factor": 1e-30},
                       "solver": {"tolerance": 1e-10, "max_iterations": 20000, "initial_factor": 100.0},
                       "stop_condition": {"direction": [1.0, 0.0, 0.0], "value": -0.95}},
        "output": {"write_magnetization": False},
    }
    return cfg


def initial_m(ndof, seed=7, max_deg=5.0):
    rng = np.random.default_rng(seed)
    th = np.deg2rad(max_deg) * rng.random(ndof)
    ph = 2 * np.pi * rng.random(ndof)
    return np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], axis=1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(n, steps, warmup):
    import ctypes as C

    import torch

    from cupymag_amd._lib import check, lib
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh

    torch.cuda.set_device(0)
    cfg = workload_config(n)
    P = from_dict(cfg)
    t0 = time.time()
    arrays = kuhn_mesh(n)
    S = SimulationOperators(P, mesh_arrays=arrays, keep_host=False)
    log(f"[bench] mesh {n}^3: {S.n_nodes} nodes, {S.n_elements} tets, {S.n_dof} DOFs, "
        f"{int(S.elements_np[:, -1].sum())} defect tets ({time.time() - t0:.1f}s)")
    t0 = time.time()
    sim = GSPMSimulation(S, P)
    log(f"[bench] host assembly + upload + weights {time.time() - t0:.1f}s")
    A_el = S.csr_elasticity()[0]
    ctx = sim.ctx
    sim.set_m(initial_m(S.n_dof))
    H = [P.Hext1, P.Hext2, P.Hext3]
    for k in range(warmup):
        t = time.time()
        rep = sim.step(H, False)
        log(f"[bench] warmup step {k}: {time.time() - t:.3f}s iters={rep.iters}")
    check(lib().cpm_prof_enable(ctx.h, 1))
    check(lib().cpm_mat_prof(A_el.h, None, None, 1))
    torch.cuda.synchronize()
    lat = []
    iters = []
    t_start = time.perf_counter()
    for k in range(steps):
        t = time.perf_counter()
        rep = sim.step(H, False)
        lat.append(time.perf_counter() - t)
        iters.append(rep.iters)
        log(f"[bench] step {k}: {lat[-1]:.3f}s iters={rep.iters}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms = C.c_double()
    nl = C.c_int64()
    check(lib().cpm_mat_prof(A_el.h, C.byref(ms), C.byref(nl), 0))
    check(lib().cpm_prof_enable(ctx.h, 0))
    nrows, nnz = A_el.shape[0], A_el.nnz
    # compulsory bytes of one CG SpMV launch: values + column indices + row
    # pointers + p read once + q written once
    bytes_launch = 12.0 * nnz + 4.0 * (nrows + 1) + 16.0 * nrows
    avg_s = ms.value / 1e3 / max(nl.value, 1)
    return dict(n_dof=S.n_dof, n_nodes=S.n_nodes, n_elements=S.n_elements, elapsed=elapsed, lat=lat, iters=iters,
                spmv_launches=nl.value, spmv_avg_us=avg_s * 1e6, spmv_bytes=bytes_launch,
                spmv_gbs=bytes_launch / avg_s / 1e9, A_el_nnz=nnz)


def cpu_baseline(n_cpu, steps=1):
    """The NumPy/SciPy oracle (one core: SciPy CSR matvec and NumPy
    elementwise are single-threaded) on an n_cpu^3 instance of the same
    workload, timed around the step loop only."""
    from oracle import fem, gspm, mesh as om, params
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = workload_config(n_cpu)
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
    t = time.perf_counter()
    for _ in range(steps):
        gspm.step(st, ops, Avg, p)
    dt = time.perf_counter() - t
    return dict(value=ops["n"] * steps / dt, unit="node-steps/s", cores=1, kind="port",
                sample=f"oracle (NumPy/SciPy FP64) on the same workload at {n_cpu}^3 Kuhn tets "
                       f"({ops['n']} DOFs), {steps} step(s), {dt:.1f}s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--cpu-n", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this size (calibrates rocprof byte counters)")
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
    g = run_gpu(args.n, args.steps, args.warmup)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n, args.cpu_steps)
    value = g["n_dof"] * args.steps / g["elapsed"]
    lat = np.array(g["lat"]) * 1e3
    out = {
        "metric": "LLG node-steps/s",
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["elapsed"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"GSPM LLG step, {args.n}^3 Kuhn tet mesh, rotated-ellipsoid defect, "
                               f"magnetoelastic on, rot111, dt 10 ps, alpha 0.1, tol 1e-10",
                   "path": "plain FP64 (reference sequence of solves)",
                   "n_dof": g["n_dof"], "nodes": g["n_nodes"], "elements": g["n_elements"]},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "cg_iters_last_step": g["iters"][-1],
        "roofline": {"bound": "hbm", "kernel": "k_cg_spmv (A_el, CSR FP64)",
                     "achieved": g["spmv_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": g["spmv_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_us": g["spmv_avg_us"], "bytes_per_launch": g["spmv_bytes"],
                     "launches": g["spmv_launches"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
