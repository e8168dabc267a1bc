"""LLG / hysteresis driver (cupymag/core/Micromagnetics.py:34-351).

Same control flow as the reference: set up the operators, initialise m,
then step the GSPM integrator until the average magnetization along
``stop_direction`` drops below ``stop_value``; on LLG convergence
(``sumsq < LLG_accuracy`` after ``llg_min_step`` steps) write a field point
to hysteresis.txt, a VTU snapshot and last_m, and step the external field
by -dHext.  Each LLG step is one call into the device engine; the loop's
control (use_init, LLG convergence, the field step, the stop condition) is
evaluated on the device after each step (GSPMSimulation.set_control), and
the host reads its outcome from the step report.
"""
import os
import time

import numpy as np

from .assemble_simulation_operators import SimulationOperators
from .gspm import GSPMSimulation
from .parameters import load_parameters
from ..utils.magnetization_io import initialize_m, write_array


def _fields(sim, me):
    Ht = sim.get("HTILDE")
    m = sim.get("M")
    E = sim.get("E")
    return {"Htilde1": Ht[0], "Htilde2": Ht[1], "Htilde3": Ht[2], "m1": m[:, 0], "m2": m[:, 1], "m3": m[:, 2],
            "E11": E[:, 0], "E22": E[:, 1], "E33": E[:, 2], "E12": E[:, 3], "E23": E[:, 4], "E13": E[:, 5]}


ENGINES = ("auto", "fp64", "fast64", "mixed")


def engine_for(P, SimOp, engine="auto"):
    """GSPMSimulation options of an engine.  "auto" follows the config's
    ``simulation.precision`` (precision_select.py:37-53): DP -> fast64 (FP64
    CG with the reference's recurrence and stopping rule on the SELL
    operators: the plain FP64 path's results to rounding, about twice its
    speed); SP -> mixed (FP32 CG refined in FP64, magnetization within 1e-7
    of FP64, where the reference computes everything in FP32), with the
    multigrid-preconditioned solves when the mesh has a hierarchy (a
    structured Kuhn mesh of more than 8 cells per dimension).  "fp64": CSR
    operators, one solve per right-hand side, the reference's sequence."""
    if engine not in ENGINES:
        raise ValueError(f"unknown engine {engine!r} (one of {', '.join(ENGINES)})")
    if engine == "auto":
        engine = "mixed" if P.precision == "SP" else "fast64"
    kw = {"mode": engine}
    if engine == "mixed" and SimOp.multigrid(precision=16) is not None:
        kw["precond"] = "mg"
    return kw


def main(config_path=None, max_steps=None, mesh_arrays=None, verbose=True, engine="auto"):
    """Run the simulation of ``config_path``; ``max_steps`` caps the total
    number of LLG steps (the reference loop has no cap); ``engine``: see
    engine_for."""
    P = load_parameters(config_path)
    say = print if verbose else (lambda *a, **k: None)
    out_dir = str(P.output_dir)
    if not os.path.exists(out_dir):
        os.makedirs(out_dir)
        say(f"Directory {out_dir} did not exist previously, create directory {out_dir}.")
    t0 = time.time()
    SimOp = SimulationOperators(P, mesh_arrays=mesh_arrays)
    say(f"Initialized {P.grid_type} FEM mesh with #nodes={SimOp.n_nodes}, #elements={SimOp.n_elements}, "
        f"#defect elements={int(SimOp.elements_np[:, -1].sum())}.")
    kw = engine_for(P, SimOp, engine)
    sim = GSPMSimulation(SimOp, P, **kw)
    say(f"Engine: {kw['mode']}{' with multigrid-preconditioned solves' if kw.get('precond') else ''} "
        f"(simulation.precision {P.precision}).")
    Avg = sim.Avg
    nDOF = SimOp.n_dof
    m0 = initialize_m(P.restart_m, P.restart, P.initial_magnetization, nDOF, SimOp.DefDOF)
    avg_m = sim.set_m(m0)
    nstep = count = 0
    sumsq = np.inf
    H = [P.Hext1, P.Hext2, P.Hext3]
    dH = [P.dHext1, P.dHext2, P.dHext3]
    # the loop's control (use_init, LLG convergence, the field step, the stop
    # condition) runs on the device after every step; the host reads its
    # outcome from the step report
    sim.set_control(H0=H, dH=dH)
    hyst = open(os.path.join(out_dir, "hysteresis.txt"), "w", buffering=1)
    hyst.write("Hext1\tHext2\tHext3\tavg_m1\tavg_m2\tavg_m3\n")
    t1 = time.time()
    say(f"FEM system set up done using {t1 - t0:.2f} s, begin LLG iterations.")
    total = 0
    while avg_m @ P.stop_direction > P.stop_value:
        if max_steps is not None and total >= max_steps:
            break
        nstep += 1
        total += 1
        rep = sim.step()
        sumsq = rep.sumsq
        if nstep % 200 == 0:
            say(f"m diff L2 norm sq: {sumsq:.10f} at LLG step {nstep}")
        avg_m = rep.avg_m
        Hs = rep.H  # the field of this step
        if rep.point_done or nstep % P.save_frequency == 0:
            count += 1
            Avg.write_to_paraview(_fields(sim, P.ME), os.path.join(out_dir, f"field_{Hs[0] * P.ms:.0f}_{count}.vtu"),
                                  alpha=P.vtu_alpha)
            hyst.write(f"{Hs[0] * P.ms:.1f}\t{Hs[1] * P.ms:.1f}\t{Hs[2] * P.ms:.1f}\t{avg_m[0]}\t{avg_m[1]}\t{avg_m[2]}\n")
            if P.write_m:
                write_array(sim.get("M"), os.path.join(out_dir, "last_m.h5"))
            if nstep >= 5000:
                say(f"LLG not converging at external field {Hs[0] * P.ms:.0f} A/m, at LLG step {nstep}.")
        H = list(rep.H_next)
        if rep.point_done:
            say(f"LLG converge at LLG step {nstep}.")
            say(f"Current average magnetization is m1 = {avg_m[0]}, m2 = {avg_m[1]}, m3 = {avg_m[2]}.")
            nstep = count = 0
            if avg_m @ P.stop_direction > P.stop_value:
                say(f"Now calculating LLG with Hext1 = {H[0] * P.ms:.0f} A/m.")
    hyst.write(f"{H[0] * P.ms:.1f}\t{H[1] * P.ms:.1f}\t{H[2] * P.ms:.1f}\t{avg_m[0]}\t{avg_m[1]}\t{avg_m[2]}\n")
    hyst.close()
    say(f"Final state for average magnetization: m1 = {avg_m[0]}, m2 = {avg_m[1]}, m3 = {avg_m[2]}.")
    count += 1
    Avg.write_to_paraview(_fields(sim, P.ME), os.path.join(out_dir, f"field_{H[0] * P.ms:.0f}_{count}.vtu"),
                          alpha=P.vtu_alpha)
    if P.write_m:
        write_array(sim.get("M"), os.path.join(out_dir, "last_m.h5"))
    t2 = time.time()
    say(f"The code uses {t2 - t0:.2f} s in total, {t1 - t0:.2f} s for system set up, and {t2 - t1:.2f} for LLG "
        f"calculation.")
    return sim, avg_m


if __name__ == "__main__":
    main()
