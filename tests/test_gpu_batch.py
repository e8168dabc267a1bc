"""Batches of simulations (cpm_batch): many meshes of different sizes
stepped together, each solve one batched CG over all members.

Tolerances: a fast64 member matches the FP64 oracle as the single fast64
path does (m within 1e-10, CG iterations within +-1); a mixed member is
within 1e-7 of it.  A member's results do not depend on the batch: alone,
in a mix, joining late or staying after others left, m, u and the CG
iteration counts are bit-identical; with graph replay of the CG iterations
they are bit-identical to eager launches.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m  # noqa: E402

MEMBERS = [
    dict(grid="Tet", n=6, me=True, rot=True, dt=5e-12),
    dict(grid="Hex", n=6, me=True, rot=False, sigma11=1e8),
    dict(grid="Tet", n=8, me=False, rot=True, dt=1e-11, alpha=0.1),
    dict(grid="Hex", n=5, me=False, rot=False, hext=(2e4, 3e4, -1e4)),
]


def _member(case, oracle=False):
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    case = dict(case)
    grid, n = case.pop("grid"), case.pop("n")
    cfg = config(grid, n=n, defect_params=1.6, **case)
    arr = kuhn_mesh(n, jitter=0.15, seed=3) if grid == "Tet" else None
    P = from_dict(cfg)
    S = SimulationOperators(P, mesh_arrays=arr, keep_host=True)
    orc = oracle_setup(cfg, arr) if oracle else None
    return P, S, orc


def _H(P):
    return [P.Hext1, P.Hext2, P.Hext3]


def _state(sim):
    out = [sim.get("M").cpu().numpy()]
    if sim.P.ME:
        out.append(sim.get("UEL").cpu().numpy())
    return out


@pytest.mark.parametrize("mode", ["fast64", "mixed"])
def test_batch_matches_oracle(mode):
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    built = [_member(c, oracle=True) for c in MEMBERS]
    sims, states = [], []
    batch = None
    for k, (P, S, (p, ops, Avg, _)) in enumerate(built):
        sim = GSPMSimulation(S, P, mode=mode)
        m0 = tilted_m(ops["n"], seed=20 + k)
        sim.set_m(m0)
        st = og.init_state(ops, Avg, m0)
        p["H_now"] = p["H"]
        if batch is None:
            batch = GSPMBatch(sim.ctx)
        batch.add(sim)
        sims.append(sim)
        states.append(st)
    for step in range(3):
        ui = [st["use_init"] for st in states]
        logs = []
        for (P, S, (p, ops, Avg, _)), st in zip(built, states):
            log = []
            og.step(st, ops, Avg, p, log)
            logs.append(log)
        reps = batch.step([_H(P) for P, _, _ in built], ui)
        for k, (sim, st, rep, log) in enumerate(zip(sims, states, reps, logs)):
            err = np.max(np.abs(sim.get("M").cpu().numpy() - st["m"]))
            me = built[k][2][0]["ME"]
            if mode == "fast64":
                assert err < 1e-10, (k, step, err)
                ref_it = [it for _, it in log]
                got = [it for s, it in enumerate(rep.iters) if me or s > 0]
                assert all(abs(a - b) <= 1 for a, b in zip(got, ref_it)), (k, got, ref_it)
            else:
                assert err < 1e-7, (k, step, err)
            assert abs(rep.sumsq - st["sumsq"]) <= 1e-6 * max(st["sumsq"], 1e-30) + 1e-20


@pytest.mark.parametrize("mode", ["fast64", "mixed", "mixed+xwarm"])
def test_batch_member_results_independent_of_the_batch(mode):
    """Solo runs vs one continuously changing batch: member 2 joins after
    the first step, member 0 leaves after the second, member 1 stays to the
    end; every member's state after each of its steps equals its solo run
    bit for bit (also with initial guesses extrapolated from each member's
    own last two solutions)."""
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    built = [_member(c) for c in MEMBERS[:3]]
    nsteps = [2, 4, 3]
    base, _, opt = mode.partition("+")

    def fresh(k):
        P, S, _ = built[k]
        sim = GSPMSimulation(S, P, mode=base, warm_start="extrapolate" if opt == "xwarm" else False)
        sim.set_m(tilted_m(S.n_dof, seed=30 + k))
        return sim

    solo = []
    for k in range(3):
        sim = fresh(k)
        b = GSPMBatch(sim.ctx)
        b.add(sim)
        hist = []
        for s in range(nsteps[k]):
            rep = b.step([_H(built[k][0])], [False])[0]
            hist.append((_state(sim), rep.iters, rep.sumsq))
        solo.append(hist)
        b.remove(sim)
        del b, sim

    sims = [fresh(k) for k in range(3)]
    b = GSPMBatch(sims[0].ctx)
    done = [0, 0, 0]
    schedule = [[0, 1], [0, 1, 2], [1, 2], [1, 2]]  # members stepped in each batch step
    member = []
    for active in schedule:
        want = [k for k in active if done[k] < nsteps[k]]
        for k in list(member):
            if k not in want:
                b.remove(sims[k])
                member.remove(k)
        for k in want:
            if k not in member:
                b.add(sims[k])
                member.append(k)
        reps = b.step([_H(built[k][0]) for k in member], [False] * len(member))
        for k, rep in zip(member, reps):
            st, iters, sumsq = solo[k][done[k]]
            got = _state(sims[k])
            for a, r in zip(got, st):
                np.testing.assert_array_equal(a, r)
            assert rep.iters == iters
            assert rep.sumsq == sumsq
            done[k] += 1
    assert done == nsteps


def test_batch_graph_replay_matches_eager():
    """Captured chunks of CG iterations replayed many times per solve (with
    the step's eager kernels and host reads between replays) give the same
    bits as launching every kernel."""
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    built = [_member(c) for c in MEMBERS[:2]]
    runs = []
    for graphs in (False, True):
        sims = []
        b = None
        for k, (P, S, _) in enumerate(built):
            sim = GSPMSimulation(S, P, mode="mixed")
            sim.set_m(tilted_m(S.n_dof, seed=40 + k))
            if b is None:
                b = GSPMBatch(sim.ctx, graphs=graphs)
            b.add(sim)
            sims.append(sim)
        hist = []
        for s in range(3):
            reps = b.step([_H(P) for P, _, _ in built], [False, s > 0])
            hist.append(([_state(x) for x in sims], [r.iters for r in reps]))
        runs.append(hist)
    for (st0, it0), (st1, it1) in zip(*runs):
        assert it0 == it1
        for a, c in zip(st0, st1):
            for x, y in zip(a, c):
                np.testing.assert_array_equal(x, y)


def test_batch_rejects_bad_members():
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    P, S, _ = _member(MEMBERS[3])
    plain = GSPMSimulation(S, P, mode="fp64")
    b = GSPMBatch(plain.ctx)
    with pytest.raises(ValueError, match="SELL paths"):
        b.add(plain)
    fast = GSPMSimulation(S, P, mode="fast64")
    mixed = GSPMSimulation(S, P, mode="mixed")
    b.add(fast)
    with pytest.raises(ValueError, match="same mode"):
        b.add(mixed)
    with pytest.raises(ValueError, match="already belongs"):
        b.add(fast)
    fast.set_m(tilted_m(S.n_dof, seed=1))
    with pytest.raises(ValueError, match="belongs to a batch"):
        fast.step(_H(P), False)
    b.remove(fast)
    fast.step(_H(P), False)
