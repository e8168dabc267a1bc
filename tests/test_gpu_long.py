"""Long-trajectory parity on the reference's CPU-runnable case: a 32^3 hex
grid with a central 4^3-cell non-magnetic cube, Ms 1.2e6 A/m, A 1.5e-11 J/m,
K1 -1e4 J/m^3, demag factors 1/3, Hext (5e4, 0, 0) A/m, no magnetoelastic
coupling, m = +x with seeded tilts <= 5 degrees, 200 steps at dt 1 ps,
alpha 0.5, CG tol 1e-10.  The plain FP64 and fast64 paths stay within 1e-10
of the oracle over all 200 steps, the mixed path within 1e-7; CG iteration
counts of the FP64 paths within +-1 at every checked step."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m  # noqa: E402


def test_hex32_200_steps_match_oracle():
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.core.parameters import from_dict
    from oracle import gspm as og
    cfg = config("Hex", n=32, me=False, ndef=4, dt=1e-12, alpha=0.5, tol=1e-10, hext=(5e4, 0.0, 0.0))
    P = from_dict(cfg)
    S = SimulationOperators(P, keep_host=False)
    p, ops, Avg, _ = oracle_setup(cfg)
    m0 = tilted_m(ops["n"], seed=2024)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    sims = {mode: GSPMSimulation(S, P, mode=mode) for mode in ("fp64", "fast64", "mixed")}
    for sim in sims.values():
        sim.set_m(m0)
    H = [P.Hext1, P.Hext2, P.Hext3]
    tol = {"fp64": 1e-10, "fast64": 1e-10, "mixed": 1e-7}
    worst = {mode: 0.0 for mode in sims}
    for k in range(200):
        ui = st["use_init"]
        log = []
        og.step(st, ops, Avg, p, log)
        for mode, sim in sims.items():
            rep = sim.step(H, ui)
            if (k + 1) % 25 == 0:
                err = float(np.max(np.abs(sim.get("M").cpu().numpy() - st["m"])))
                worst[mode] = max(worst[mode], err)
                assert err < tol[mode], (mode, k + 1, err)
                if mode != "mixed":
                    ref_it = [it for _, it in log]
                    got = rep.iters[1:]
                    assert all(abs(a - b) <= 1 for a, b in zip(got, ref_it)), (mode, k + 1, got, ref_it)
        if (k + 1) % 50 == 0:
            print(f"step {k + 1}: max |m - m_oracle| " + ", ".join(f"{m} {e:.2e}" for m, e in worst.items()),
                  flush=True)
    assert np.isclose(np.linalg.norm(st["avg_m"]), np.linalg.norm(sims["fp64"].avg_m), rtol=1e-12)
