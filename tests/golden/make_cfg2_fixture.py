"""Writes tests/golden/cfg2_fp64_oracle.npz: the CPU oracle (FP64
NumPy/SciPy restatement of the reference step) on the 64^3 magnetoelastic
workload of tests.cases.cfg2_case, three steps from the seeded initial m.
Kept per step: the CG iteration count of every solve, sumsq and avg_m;
after steps 1 and 3: m, u, E and Htilde at seeded sample DOFs and their
sums over blocks of consecutive DOFs (tests.cases.cfg2_summary).

    python tests/golden/make_cfg2_fixture.py      (about ten minutes, ~30 GB RAM)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import gspm  # noqa: E402
from tests.cases import cfg2_case, cfg2_summary, oracle_setup  # noqa: E402

STEPS = 3
OUT = os.path.join(ROOT, "tests", "golden", "cfg2_fp64_oracle.npz")


def main():
    cfg, arrays, m0f = cfg2_case()
    t = time.time()
    p, ops, Avg, _ = oracle_setup(cfg, arrays)
    n = ops["n"]
    print(f"oracle operators for {n} DOFs: {time.time() - t:.0f}s", flush=True)
    st = gspm.init_state(ops, Avg, m0f(n))
    p["H_now"] = p["H"]
    order = ("elasticity", "demag", "Gauss-Seidel g1n", "Gauss-Seidel g2n", "Gauss-Seidel g3n",
             "Gauss-Seidel g1star", "Gauss-Seidel g2star", "Gauss-Seidel m1starstar",
             "Gauss-Seidel m2starstar", "Gauss-Seidel m3starstar")
    rec = {"iters": [], "sumsq": [], "avg_m": []}
    for k in range(STEPS):
        t = time.time()
        log = []
        gspm.step(st, ops, Avg, p, log=log)
        its = dict(log)
        rec["iters"].append([its[s] for s in order])
        rec["sumsq"].append(st["sumsq"])
        rec["avg_m"].append(st["avg_m"])
        print(f"step {k + 1}: {time.time() - t:.0f}s iterations {rec['iters'][-1]}", flush=True)
        if k + 1 in (1, STEPS):
            fields = {"m": st["m"], "u": st["u"].reshape(3, n).T, "E": st["E"], "Ht": np.stack(st["Ht"], axis=1)}
            for key, v in cfg2_summary(n, fields).items():
                rec[f"s{k + 1}_{key}"] = v
    np.savez_compressed(OUT, n_dof=n, steps=STEPS, cfg=json.dumps(cfg, sort_keys=True),
                        **{k: np.asarray(v) for k, v in rec.items()})
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
