"""Batched multigrid-preconditioned CG (cpm_cg_mg_batch, and the batched
elasticity/demag solves of cpm_batch_step for members created with
precond="mg").

* Hierarchies of different depths (10^3 -> 5^3, 20^3 -> 10^3 -> 5^3,
  13^3 -> 7^3 non-nested; the scalar ones also with their dense level at
  <= 16 cells, 20^3 -> 10^3), scalar (demag, Gauss-Seidel A1) and 3x3-block
  operators, FP32 and FP64 V-cycles, one column or three (A1): every
  system's solution, iteration counts and residuals equal the one-system
  solve (cpm_cg_mg) bit for bit, with and without initial guesses, for a
  zero right-hand side too; maxiter reached raises.
* GSPM batches mixing multigrid members of different depths, a member
  without the preconditioner and two members on one set of operators (one
  hierarchy, solved in turn): each member's m, u, iteration counts and
  sumsq equal its run alone in a batch bit for bit, and fast64 members stay
  within 1e-8 of the FP64 oracle (the single multigrid path's tolerance).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m  # noqa: E402


def _ops(n, me=True, jitter=0.1, seed=3):
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = config("Tet", n=n, me=me, rot=True, dt=5e-12, defect_params=2.5)
    arr = kuhn_mesh(n, jitter=jitter, seed=seed)
    P = from_dict(cfg)
    return cfg, arr, P, SimulationOperators(P, mesh_arrays=arr, keep_host=True)


@pytest.fixture(scope="module")
def meshes():
    return [_ops(n) for n in (10, 20, 13)]


@pytest.mark.parametrize("name,a1", [("A_demag", "dense"), ("A_el", "dense"), ("A1", "dense"), ("A1", "auto"),
                                     ("A1", "mix")])
@pytest.mark.parametrize("prec", [16, 32, 64])
@pytest.mark.parametrize("coarsest_scalar", [8, 16])
def test_solve_batch_equals_single_solves(meshes, name, a1, prec, coarsest_scalar):
    """The scalar hierarchies at the elasticity's depths (20^3 -> 10^3 ->
    5^3), or with coarsest_scalar = 16 stopping at <= 16 cells (20^3 ->
    10^3); A1 also with its first coarse level solved by Chebyshev steps
    (a1_coarse "auto" on these meshes), alone and mixed with dense-coarsest
    hierarchies in one batch."""
    from cupymag_amd.solvers.multigrid import solve_batch
    if name == "A_el" and coarsest_scalar == 8:
        pytest.skip("the elasticity hierarchy has one depth rule")
    if a1 != "dense" and coarsest_scalar == 16:
        pytest.skip("a smoothed coarsest level does not depend on coarsest_scalar")
    kinds = {"dense": ["dense"] * 3, "auto": ["auto"] * 3, "mix": ["auto", "dense", "auto"]}[a1]
    mgs = [S.multigrid(precision=prec, coarsest_scalar=coarsest_scalar, a1_coarse=k)[name]
           for (_, _, _, S), k in zip(meshes, kinds)]
    if a1 == "dense":
        want = [2, 3, 2] if (name == "A_el" or coarsest_scalar == 8) else [2, 2, 2]
    else:
        want = [2 if k == "auto" else l for k, l in zip(kinds, [2, 3, 2])]
        assert [m.info()["coarsest"] != "dense" for m in mgs] == [k == "auto" for k in kinds]
    assert [m.info()["levels"] for m in mgs] == want
    bs = 3 if name == "A_el" else 1
    rng = np.random.default_rng(11)
    rhs, x0s = [], []
    for k, m in enumerate(mgs):
        n = m.levels[0].shape[0]
        b = rng.standard_normal(n)
        b[:bs] = 0.0
        rhs.append(torch.as_tensor(b, device="cuda"))
        x0s.append(torch.as_tensor(rng.standard_normal(n) * 0.1, device="cuda") if k == 1 else None)
    got = solve_batch(mgs, rhs, x0s, tol=1e-10, maxiter=500)
    for m, b, x0, (x, it, res) in zip(mgs, rhs, x0s, got):
        xs, its, ress = m.solve(b, x0=x0, tol=1e-10, maxiter=500, use_init=x0 is not None)
        assert it == its and 0 < it < 60
        assert res == ress
        np.testing.assert_array_equal(x.cpu().numpy(), xs.cpu().numpy())
    # a zero right-hand side among others: x = 0 after no iteration
    rhs[0] = torch.zeros_like(rhs[0])
    got = solve_batch(mgs, rhs, [None, x0s[1], None], tol=1e-10, maxiter=500)
    assert got[0][1] == 0 and float(got[0][0].abs().max()) == 0.0
    for m, b, x0, (x, it, res) in list(zip(mgs, rhs, x0s, got))[1:]:
        xs, its, _ = m.solve(b, x0=x0, tol=1e-10, maxiter=500, use_init=x0 is not None)
        assert it == its
        np.testing.assert_array_equal(x.cpu().numpy(), xs.cpu().numpy())
    with pytest.raises(RuntimeError, match="did not converge"):
        solve_batch(mgs, rhs, tol=1e-14, maxiter=2)


@pytest.mark.parametrize("prec", [16, 32, 64])
def test_solve_batch_three_columns_equals_single_solves(meshes, prec):
    """The Gauss-Seidel operator A1 with three interleaved columns (one of
    them zero in one system, an initial guess in another)."""
    from cupymag_amd.solvers.multigrid import solve_batch
    mgs = [S.multigrid(precision=prec)["A1"] for _, _, _, S in meshes]
    rng = np.random.default_rng(12)
    rhs, x0s = [], []
    for k, m in enumerate(mgs):
        B = rng.standard_normal((m.levels[0].shape[0], 3))
        if k == 0:
            B[:, 2] = 0.0
        rhs.append(torch.as_tensor(B.ravel(), device="cuda"))
        x0s.append(torch.as_tensor(0.1 * rng.standard_normal(B.size), device="cuda") if k == 2 else None)
    got = solve_batch(mgs, rhs, x0s, tol=1e-10, maxiter=500, nrhs=3)
    for m, b, x0, (x, its, res) in zip(mgs, rhs, x0s, got):
        xs, its1, res1 = m.solve(b, x0=x0, tol=1e-10, maxiter=500, use_init=x0 is not None, nrhs=3)
        assert its == its1 and res == res1
        np.testing.assert_array_equal(x.cpu().numpy(), xs.cpu().numpy())
    assert got[0][1][2] == 0 and float(got[0][0].reshape(-1, 3)[:, 2].abs().max()) == 0.0
    assert all(0 < it < 40 for it in got[1][1])


def test_solve_batch_rejects_shared_or_mismatched_hierarchies(meshes):
    from cupymag_amd.solvers.multigrid import solve_batch
    S0, S1 = meshes[0][3], meshes[1][3]
    m0 = S0.multigrid(precision=32)["A_demag"]
    b = torch.zeros(m0.levels[0].shape[0], dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="share one hierarchy"):
        solve_batch([m0, m0], [b, b])
    m1 = S1.multigrid(precision=64)["A_demag"]
    b1 = torch.zeros(m1.levels[0].shape[0], dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="differ"):
        solve_batch([m0, m1], [b, b1])
    me = meshes[0][3].multigrid(precision=32)["A_el"]
    be = torch.zeros(me.levels[0].shape[0] * 3, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="3 for scalar"):
        solve_batch([me], [be], nrhs=3)


def _H(P):
    return [P.Hext1, P.Hext2, P.Hext3]


def _state(sim):
    out = [sim.get("M").cpu().numpy()]
    if sim.P.ME:
        out.append(sim.get("UEL").cpu().numpy())
    return out


@pytest.mark.parametrize("mode", ["fast64", "mixed"])
def test_batch_step_with_multigrid_members(meshes, mode):
    """Members: 10^3 (ME, multigrid), 20^3 (no ME, multigrid, 3 levels),
    13^3 (ME, multigrid, non-nested), 13^3 again on the same operators (one
    hierarchy, solved in a second round), 8^3 without the preconditioner."""
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    _, _, P20, S20 = _ops(20, me=False, seed=4)
    _, _, P8, S8 = _ops(8)
    specs = [(meshes[0][2], meshes[0][3], "mg"), (P20, S20, "mg"), (meshes[2][2], meshes[2][3], "mg"),
             (meshes[2][2], meshes[2][3], "mg"), (P8, S8, None)]
    nsteps = 3

    def fresh(k):
        P, S, pc = specs[k]
        sim = GSPMSimulation(S, P, mode=mode, precond=pc)
        sim.set_m(tilted_m(S.n_dof, seed=50 + k))
        return sim

    solo = []
    for k in range(len(specs)):
        sim = fresh(k)
        b = GSPMBatch(sim.ctx)
        b.add(sim)
        hist = []
        for s in range(nsteps):
            rep = b.step([_H(specs[k][0])], [s > 0])[0]
            hist.append((_state(sim), rep.iters, rep.sumsq))
        solo.append(hist)
        b.remove(sim)
        del b, sim

    sims = [fresh(k) for k in range(len(specs))]
    b = GSPMBatch(sims[0].ctx)
    for sim in sims:
        b.add(sim)
    for s in range(nsteps):
        reps = b.step([_H(P) for P, _, _ in specs], [s > 0] * len(specs))
        for k, rep in enumerate(reps):
            st, iters, sumsq = solo[k][s]
            for a, r in zip(_state(sims[k]), st):
                np.testing.assert_array_equal(a, r)
            assert rep.iters == iters and rep.sumsq == sumsq
            if specs[k][2] == "mg":
                assert 0 < rep.iters[1] < 40
    if mode != "fast64":
        return
    # oracle check of the 13^3 ME multigrid member (from its seed, same steps)
    cfg, arr, P, S = meshes[2]
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    st = og.init_state(ops, Avg, tilted_m(ops["n"], seed=52))
    p["H_now"] = p["H"]
    for s in range(nsteps):
        st["use_init"] = s > 0
        og.step(st, ops, Avg, p, [])
    assert np.max(np.abs(sims[2].get("M").cpu().numpy() - st["m"])) < 1e-8


def test_hierarchies_built_concurrently_match_sequential():
    """Hierarchies built from several host threads on one context (as the
    batch and serving preparation do) equal those built one at a time: the
    power-iteration estimates of every level, and a solve's iterations and
    solution.  (The estimate once summed partials from the context's shared
    buffer, which the interleaved builds overwrote.)"""
    import concurrent.futures as cf

    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.device import context
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    ctx = context(0)
    sizes = [10, 12, 13, 16, 18, 11]

    def build(n):
        torch.cuda.set_device(ctx.device)
        cfg = config("Tet", n=n, me=True, rot=True, dt=5e-12, defect_params=2.5)
        S = SimulationOperators(from_dict(cfg), mesh_arrays=kuhn_mesh(n, jitter=0.1, seed=n), ctx=ctx)
        return S, S.multigrid(precision=16)

    seq = [build(n) for n in sizes]
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        par = list(ex.map(build, sizes))
    rng = np.random.default_rng(3)
    for (_, hs), (_, hp) in zip(seq, par):
        for name in ("A_demag", "A_el", "A1"):
            assert hs[name].info()["lmax"] == hp[name].info()["lmax"], name
        b = rng.standard_normal(hs["A_el"].levels[0].shape[0])
        b[:3] = 0.0
        xs, its, _ = hs["A_el"].solve(torch.as_tensor(b, device="cuda"), tol=1e-10, maxiter=500)
        xp, itp, _ = hp["A_el"].solve(torch.as_tensor(b, device="cuda"), tol=1e-10, maxiter=500)
        assert its == itp and its < 40
        np.testing.assert_array_equal(xs.cpu().numpy(), xp.cpu().numpy())


def test_set_m_from_concurrent_threads_on_one_context():
    """Simulations sharing one context take their initial m from several
    host threads (as the serving preparation does): each returns its own
    average magnetization, equal to a one-at-a-time set_m, because every
    simulation copies its scalars through its own pinned buffer (a shared
    context buffer once let one thread read another's average)."""
    import concurrent.futures as cf

    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.gspm import GSPMSimulation
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.device import context
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    ctx = context(0)
    cfg = config("Tet", n=8, me=True, rot=False, dt=5e-12, defect_params=2.0)
    P = from_dict(cfg)
    S = SimulationOperators(P, mesh_arrays=kuhn_mesh(8, jitter=0.1, seed=2), ctx=ctx)
    sims = [GSPMSimulation(S, P, mode="mixed") for _ in range(8)]
    m0 = [tilted_m(S.n_dof, seed=k, max_deg=10.0 + 8.0 * k) for k in range(8)]
    want = [sim.set_m(m).copy() for sim, m in zip(sims, m0)]
    assert len({tuple(w) for w in want}) == 8

    def go(k):
        torch.cuda.set_device(ctx.device)
        return [sims[k].set_m(m0[k]).copy() for _ in range(25)]

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        got = list(ex.map(go, range(8)))
    for k in range(8):
        for g in got[k]:
            np.testing.assert_array_equal(g, want[k])
