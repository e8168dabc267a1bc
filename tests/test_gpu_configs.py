"""The five workloads the path is specified on, each at its own size on the
GPU.

* 32^3 hex grid, 4^3-cell cube defect, magnetoelastic coupling off, 200
  steps: tests/test_gpu_long.py (against the oracle every 25 steps).
* 64^3 jittered Kuhn mesh, sphere defect r = 8h, magnetoelastic on: the
  plain FP64 path (and fast64) against the oracle's golden file
  (tests/golden/cfg2_fp64_oracle.npz, written by make_cfg2_fixture.py)
  over three steps -- m per node 1e-10, u / E / Htilde 1e-10 of their
  largest value at 4096 sampled DOFs, sums over blocks of 256 DOFs to
  256 x those bounds, CG iterations +-1; then 100 steps of the mixed and
  mixed+multigrid paths against plain FP64, every field within 1e-7.
* 128^3 Kuhn mesh, rotated ellipsoid, rot111, dt 10 ps, alpha 0.1: 50
  steps of mixed+multigrid, cold-started and started from the extrapolated
  last two solutions, against fast64 (FP64 CG with the reference's
  recurrence, itself checked against plain FP64 after two steps), m within
  1e-7, at CG tolerance 1e-12 (see workloads.headline_config for why).
* 64 simulations, Kuhn n^3 with n log-uniform in [16, 96], autotuned SpMV
  schedules per size class: two batched steps equal the same steps taken
  one simulation at a time, bit for bit; |m| = 1 off the defect, 0 on it.
* Hysteresis-sweep serving with n log-uniform in [20, 128]: every request's
  results are the same alone and in the mix.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import CFG2_BLOCK, cfg2_case, cfg2_summary  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg2_fp64_oracle.npz")


def _ops(cfg, arrays, **kw):
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    P = from_dict(cfg)
    S = SimulationOperators(P, mesh_arrays=arrays, keep_host=False, **kw)
    S.csr_demag_and_deriv()
    return P, S


def _fields(sim, n):
    return {"m": sim.get("M").cpu().numpy(), "u": sim.get("UEL").cpu().numpy().reshape(3, n).T,
            "E": sim.get("E").cpu().numpy(), "Ht": sim.get("HTILDE").cpu().numpy().T}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", ["fp64", "fast64"])
def test_cfg2_fp64_paths_match_the_oracle_golden_file(mode):
    from cupymag_amd.core.gspm import GSPMSimulation
    fx = np.load(GOLDEN, allow_pickle=False)
    cfg, arrays, m0f = cfg2_case()
    assert json.loads(str(fx["cfg"])) == json.loads(json.dumps(cfg, sort_keys=True))
    P, S = _ops(cfg, arrays, matrix_free_fel=(mode != "fp64"))
    n = S.n_dof
    assert n == int(fx["n_dof"]) == 64 ** 3
    sim = GSPMSimulation(S, P, mode=mode)
    sim.set_m(m0f(n))
    H = [P.Hext1, P.Hext2, P.Hext3]
    for k in range(int(fx["steps"])):
        rep = sim.step(H, False)
        want_it = fx["iters"][k]
        assert all(abs(a - b) <= 1 for a, b in zip(rep.iters, want_it)), (k, rep.iters, want_it)
        assert abs(rep.sumsq - fx["sumsq"][k]) <= 1e-9 * fx["sumsq"][k]
        assert np.abs(rep.avg_m - fx["avg_m"][k]).max() <= 1e-10
        if f"s{k + 1}_idx" not in fx:
            continue
        got = cfg2_summary(n, _fields(sim, n))
        np.testing.assert_array_equal(got["idx"], fx[f"s{k + 1}_idx"])
        for f in ("m", "u", "E", "Ht"):
            ref_s, ref_b = fx[f"s{k + 1}_{f}_samples"], fx[f"s{k + 1}_{f}_blocks"]
            scale = 1.0 if f == "m" else float(np.abs(ref_s).max())
            err_s = np.abs(got[f"{f}_samples"] - ref_s).max()
            err_b = np.abs(got[f"{f}_blocks"] - ref_b).max()
            assert err_s <= 1e-10 * scale, (k, f, err_s, scale)
            assert err_b <= CFG2_BLOCK * 1e-10 * scale, (k, f, err_b, scale)


@pytest.mark.timeout(600)
def test_cfg2_mixed_paths_stay_within_1e7_of_plain_fp64_over_100_steps():
    from cupymag_amd.core.gspm import GSPMSimulation
    cfg, arrays, m0f = cfg2_case()
    P, S = _ops(cfg, arrays, matrix_free_fel=True)
    n = S.n_dof
    m0 = m0f(n)
    H = [P.Hext1, P.Hext2, P.Hext3]
    sims = {"fp64": GSPMSimulation(S, P, mode="fp64"), "mixed": GSPMSimulation(S, P, mode="mixed"),
            "mixed+mg": GSPMSimulation(S, P, mode="mixed", precond="mg")}
    for s in sims.values():
        s.set_m(m0)
    worst = {k: 0.0 for k in sims if k != "fp64"}
    for step in range(1, 101):
        for s in sims.values():
            s.step(H, False)
        if step % 20:
            continue
        print(f"[cfg2] step {step}", flush=True)
        ref = _fields(sims["fp64"], n)
        for name in worst:
            got = _fields(sims[name], n)
            for f in ("m", "u", "E", "Ht"):
                scale = 1.0 if f == "m" else float(np.abs(ref[f]).max())
                err = float(np.abs(got[f] - ref[f]).max()) / scale
                worst[name] = max(worst[name], err)
                assert err <= 1e-7, (step, name, f, err)
    assert worst["mixed"] > 0.0 and worst["mixed+mg"] > 0.0


@pytest.mark.timeout(900)
def test_cfg3_mixed_multigrid_within_1e7_of_fp64_after_50_steps():
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.serving.workloads import headline_config, initial_m
    cfg = headline_config(128, tol=1e-12)
    P, S = _ops(cfg, kuhn_mesh(128), matrix_free_fel=True, autotune=True)
    n = S.n_dof
    m0 = initial_m(n, 7)
    H = [P.Hext1, P.Hext2, P.Hext3]
    # fast64 follows plain FP64 (the reference's arithmetic) to rounding
    a, b = GSPMSimulation(S, P, mode="fp64"), GSPMSimulation(S, P, mode="fast64")
    for s in (a, b):
        s.set_m(m0)
    for _ in range(2):
        ra, rb = a.step(H, False), b.step(H, False)
        assert all(abs(x - y) <= 1 for x, y in zip(ra.iters, rb.iters))
    assert float((a.get("M") - b.get("M")).abs().max()) <= 1e-12
    del a
    ref = GSPMSimulation(S, P, mode="fast64")
    fast = {"mixed+mg": GSPMSimulation(S, P, mode="mixed", precond="mg"),
            "mixed+mg+xwarm": GSPMSimulation(S, P, mode="mixed", precond="mg", warm_start="extrapolate")}
    for s in (ref, *fast.values()):
        s.set_m(m0)
    m_prev = torch.as_tensor(m0, device="cuda")
    for step in range(1, 51):
        ref.step(H, False)
        for name, sim in fast.items():
            rep = sim.step(H, False)
            assert max(rep.iters[:7]) < 60, (name, rep.iters)   # multigrid solves stay short at this tolerance
        if step % 10 == 0:
            for name, sim in fast.items():
                err = float((sim.get("M") - ref.get("M")).abs().max())
                print(f"[cfg3] step {step}: max |m_{name} - m_fast64| = {err:.3e}", flush=True)
                assert err <= 1e-7, (name, step, err)
    moved = float((ref.get("M") - m_prev).abs().max())
    assert moved > 0.5    # the state really evolves (m reverses in places)


@pytest.mark.timeout(900)
def test_cfg4_batch_of_64_equals_one_at_a_time():
    import concurrent.futures as cf

    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.device import context
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.serving.workloads import batch_workload, initial_m
    ctx = context(0)
    descs = batch_workload(64, seed=4)
    ns = [d["n"] for d in descs]
    assert min(ns) >= 16 and max(ns) <= 96 and len(set(ns)) > 20

    def build(d):
        torch.cuda.set_device(ctx.device)
        P = from_dict(d["cfg"])
        S = SimulationOperators(P, mesh_arrays=kuhn_mesh(d["n"]), ctx=ctx, keep_host=False, matrix_free_fel=True,
                                autotune=True, nthreads=4)
        S.csr_demag_and_deriv()
        S.multigrid(precision=16)
        return P, S

    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        built = []
        for k, b in enumerate(ex.map(build, descs)):
            built.append(b)
            if k % 8 == 7:
                print(f"[cfg4] {k + 1} instances built", flush=True)
    # every operator carries the schedule of its size class
    from cupymag_amd.autotune import get_schedule, size_class
    by_class = {}
    for _, S in built:
        for name, M in S._upload().items():
            if hasattr(M, "sell_info"):
                by_class.setdefault(size_class(M), set()).add(tuple(sorted(get_schedule(M).items())))
    assert all(len(v) == 1 for v in by_class.values())

    def fresh():
        out = []
        for (P, S), d in zip(built, descs):
            sim = GSPMSimulation(S, P, mode="mixed", precond="mg")
            sim.set_m(initial_m(S.n_dof, d["m_seed"]))
            out.append(sim)
        return out

    H = [[P.Hext1, P.Hext2, P.Hext3] for P, _ in built]
    seq = fresh()
    for k in range(2):
        for s, h in zip(seq, H):
            s.step(h, False)
        print(f"[cfg4] sequential step {k + 1} done", flush=True)
    bat = fresh()
    b = GSPMBatch(ctx)
    for s in bat:
        b.add(s)
    for _ in range(2):
        reps = b.step(H, [False] * len(bat))
        assert not any(r.failed for r in reps)
    for (P, S), s1, s2 in zip(built, seq, bat):
        m1, m2 = s1.get("M"), s2.get("M")
        assert torch.equal(m1, m2)
        nrm = m2.norm(dim=1).cpu().numpy()
        dd = np.zeros(S.n_dof, dtype=bool)
        dd[np.asarray(S.DefDOF, dtype=np.int64)] = True
        assert np.all(np.abs(nrm[~dd] - 1.0) < 1e-12) and np.all(nrm[dd] == 0.0)


@pytest.mark.timeout(900)
def test_cfg5_requests_up_to_128_cubed_same_alone_and_in_the_mix():
    from cupymag_amd.serving.server import SweepServer
    from cupymag_amd.serving.workloads import sweep_requests
    # 20 steps per request: the first large request (prepared in ~16 s) is
    # still stepping when the small ones, prepared after it, join
    reqs = sweep_requests(4, rate=20.0, seed=7, nmin=20, nmax=128, points=2, steps_per_point=10)
    ns = sorted(r["n"] for r in reqs)
    assert ns == [20, 26, 106, 125], ns
    for k, r in enumerate(reqs):
        r["arrival"] = 0.01 * k
    say = lambda *a: print(*a, flush=True)  # noqa: E731
    mixed, stats = SweepServer(mode="mixed", precond="mg").run(reqs, log=say)
    assert stats["requests"] == len(reqs) and stats["failed"] == 0
    assert stats["max_members"] >= 2
    for r in reqs:
        alone, _ = SweepServer(mode="mixed", precond="mg").run([dict(r, arrival=0.0)], log=say)
        a, b = mixed[r["rid"]], alone[r["rid"]]
        np.testing.assert_array_equal(a["m"], b["m"])
        for (h1, m1, s1), (h2, m2, s2) in zip(a["points"], b["points"]):
            np.testing.assert_array_equal(m1, m2)
            assert s1 == s2
