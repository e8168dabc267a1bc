"""SpMV autotuning: tuned schedules change launch shapes only, never the
results beyond rounding (products 1e-14 relative, CG solutions 1e-10 with
iterations within +-1 of the oracle, GSPM steps to the fast modes' own
tolerances); schedules are reused per size class."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m  # noqa: E402


def _ops(n=8, me=True, autotune=True, seed=3):
    from cupymag_amd.autotune import clear_cache
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    clear_cache()
    cfg = config("Tet", n=n, me=me, rot=True, dt=5e-12, defect_params=1.6)
    arr = kuhn_mesh(n, jitter=0.15, seed=seed)
    P = from_dict(cfg)
    S = SimulationOperators(P, mesh_arrays=arr, keep_host=True, autotune=autotune)
    S.csr_demag_and_deriv()  # uploads (and tunes) every operator
    return cfg, arr, P, S


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_tuned_schedules_keep_products_and_solves():
    from cupymag_amd.autotune import get_schedule
    from cupymag_amd.solvers.linear_solvers import solve_cg
    from oracle import gspm as og
    cfg, arr, P, S = _ops()
    assert set(S.tuning) >= {"A_demag", "F3", "A1", "A2", "F1", "A_el", "F_el"}
    sched, timings = S.tuning["A_el"]
    assert timings and {t["format"] for t in timings} == {"csr", "sell32", "sell64"}
    assert all(t["us"] > 0 for t in timings)
    best_csr = min((t for t in timings if t["format"] == "csr"), key=lambda t: t["us"])
    assert (sched["lanes"], sched["grid_cap"]) == (best_csr["lanes"], best_csr["grid_cap"])
    rng = np.random.default_rng(0)
    A_el, F_el = S.csr_elasticity()
    A1, A2, F1 = S.csr_GS()
    for name, A in (("A_el", A_el), ("F_el", F_el), ("A1", A1), ("F1", F1)):
        x = rng.standard_normal(A.shape[1])
        y = (A @ torch.as_tensor(x, device="cuda")).cpu().numpy()
        assert _rel(y, A.to_scipy() @ x) < 1e-14
        assert get_schedule(A) == S.tuning[name][0]
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    n = ops["n"]
    b = rng.standard_normal(3 * n)
    b[[0, n, 2 * n]] = 0.0
    x = solve_cg(A_el, torch.as_tensor(b, device="cuda"), tol=1e-10, maxiter=5000)
    xr, _, _ = og.cg(ops["A_el"], b, None, 1e-10, 5000)
    assert _rel(x.cpu().numpy(), xr) < 1e-10


def test_sell_product_format():
    from cupymag_amd.autotune import get_schedule, set_schedule
    _, _, _, S = _ops(n=6, me=True, autotune=False)
    A1 = S.csr_GS()[0]
    A_el = S.csr_elasticity()[0]
    rng = np.random.default_rng(1)
    x = rng.standard_normal(A1.shape[1])
    ref = A1.to_scipy() @ x
    sch = get_schedule(A1)
    set_schedule(A1, dict(sch, spmv_sell=1))
    y = (A1 @ torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert _rel(y, ref) < 1e-14
    set_schedule(A1, dict(sch, spmv_sell=0, lanes=64, grid_cap=1))
    y = (A1 @ torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert _rel(y, ref) < 1e-14
    with pytest.raises(ValueError, match="scalar SELL"):
        set_schedule(A_el, dict(get_schedule(A_el), spmv_sell=1))
    with pytest.raises(ValueError, match="power of 2"):
        set_schedule(A1, dict(sch, lanes=3))
    with pytest.raises(ValueError, match="grid cap"):
        set_schedule(A1, dict(sch, grid_cap=1 << 20))


def test_schedule_reused_per_size_class():
    from cupymag_amd.autotune import size_class, tune
    _, _, _, S1 = _ops(n=8, me=False, autotune=True, seed=3)
    # A2 has A1's pattern (same defect surgery): A1 is timed, A2 reuses it
    assert S1.tuning["A1"][1] is not None and S1.tuning["A2"][1] is None
    assert S1.tuning["A2"][0] == S1.tuning["A1"][0]
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    P = from_dict(config("Tet", n=8, me=False, rot=True, defect_params=1.6))
    S2 = SimulationOperators(P, mesh_arrays=kuhn_mesh(8, jitter=0.15, seed=4), keep_host=True)
    A1 = S2.csr_GS()[0]
    assert size_class(A1) == size_class(S1.csr_GS()[0])
    sched, timings = tune(A1)
    assert timings is None and sched == S1.tuning["A1"][0]


@pytest.mark.parametrize("mode", ["fast64", "mixed"])
def test_steps_with_tuned_operators_match_oracle(mode):
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    cfg, arr, P, S = _ops(n=6)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    m0 = tilted_m(ops["n"], seed=5)
    sim = GSPMSimulation(S, P, mode=mode)
    sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    H = [P.Hext1, P.Hext2, P.Hext3]
    for k in range(2):
        log = []
        ui = st["use_init"]
        og.step(st, ops, Avg, p, log)
        rep = sim.step(H, ui)
        err = np.max(np.abs(sim.get("M").cpu().numpy() - st["m"]))
        if mode == "fast64":
            assert err < 1e-10
            ref_it = [it for _, it in log]
            assert all(abs(a - b) <= 1 for a, b in zip(rep.iters, ref_it)), (rep.iters, ref_it)
        else:
            assert err < 1e-7
