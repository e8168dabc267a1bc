"""Multigrid-preconditioned CG on the GPU (cpm_mg_* / cpm_cg_mg).

* One V-cycle in FP64 equals a NumPy/SciPy V-cycle built from the same
  level operators, the host transfer matrix (whose stencil the Galerkin
  identity pins in test_multigrid_host.py) and the same Chebyshev interval,
  to 1e-11 relative; the FP32 V-cycle to 1e-5; the V-cycle with FP16
  operator values (precision 16) to 3e-3.  For the floating hierarchies of
  the anchored operators the reference is F^T V(F r) + e_0 r_0 on the
  periodic stiffness (level 0 recovered from the anchored operator by
  tests.cases.unanchor, the coarsest solve a pseudo-inverse).
* The preconditioned CG meets tol 1e-10 (CuPy's rule) and its solution
  matches a sparse direct solve to 1e-8 relative, for the scalar (demag)
  and 3x3-block (elasticity) operators, on regular and jittered meshes, in
  a handful of iterations; b = 0 gives x = 0; maxiter reached raises.
* GSPM steps with the multigrid solves stay within 1e-8 (fast64) and 1e-7
  (mixed) of the FP64 oracle, the other solves keep their iteration counts.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m, unanchor  # noqa: E402


def _setup(n=16, me=True, jitter=0.0):
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = config("Tet", n=n, me=me, rot=True, dt=5e-12, defect_params=2.5)
    arr = kuhn_mesh(n, jitter=jitter, seed=3)
    P = from_dict(cfg)
    return cfg, arr, P, SimulationOperators(P, mesh_arrays=arr, keep_host=True)


def _interleaved(M, bs):
    A = M.to_scipy()
    n = A.shape[0] // bs
    perm = (np.arange(bs)[None, :] * n + np.arange(n)[:, None]).ravel()
    return A[perm][:, perm].tocsr()


def _host_vcycle(mats, dims, lmax, cinv, bs, degree, ratio, b, anchor=True, coarse_smooth=None):
    """The V-cycle in SciPy: cinv solves the coarsest level, or (cinv None)
    coarse_smooth = (degree, ratio) Chebyshev steps on it from zero; degree:
    one for every level, or a list per level."""
    degs = list(degree) if isinstance(degree, (list, tuple)) else [degree] * len(mats)
    from cupymag_amd.mesh.coarsen import prolongation

    def dinv_of(A):
        n = A.shape[0] // bs
        D = np.zeros((n, bs, bs))
        Ad = A.toarray() if A.shape[0] <= 3000 else None
        for i in range(n):
            blk = (Ad[i * bs:(i + 1) * bs, i * bs:(i + 1) * bs] if Ad is not None
                   else A[i * bs:(i + 1) * bs, i * bs:(i + 1) * bs].toarray())
            D[i] = np.linalg.inv(blk)
        return lambda v: np.einsum("nij,nj->ni", D, v.reshape(n, bs)).ravel()

    Dinv = [dinv_of(A) for A in (mats if cinv is None else mats[:-1])]

    def coef(l, deg, rat):
        hi = 1.1 * lmax[l]
        lo = hi / rat
        th, de = 0.5 * (hi + lo), 0.5 * (hi - lo)
        sg = th / de
        rho = 1.0 / sg
        c1, c2 = [0.0], [0.0]
        for k in range(1, deg):
            rn = 1.0 / (2 * sg - rho)
            c1.append(rn * rho)
            c2.append(2 * rn / de)
            rho = rn
        return 1.0 / th, c1, c2

    def smooth(l, b, x, r, deg=None, rat=ratio):
        deg = degs[l] if deg is None else deg
        A = mats[l]
        c0, c1, c2 = coef(l, deg, rat)
        d = c0 * Dinv[l](r)
        x = x + d
        for k in range(1, deg):
            r = r - A @ d
            d = c1[k] * d + c2[k] * Dinv[l](r)
            x = x + d
        return x, r, d

    def V(l, b):
        if l == len(mats) - 1:
            if cinv is None:
                return smooth(l, b, np.zeros_like(b), b.copy(), *coarse_smooth)[0]
            return cinv @ b
        A = mats[l]
        x, r, d = smooth(l, b, np.zeros_like(b), b.copy())
        r = r - A @ d
        Pm = prolongation(dims[l], bs=bs, anchor=anchor)
        ec = V(l + 1, Pm.T @ r)
        x = x + Pm @ ec
        x, _, _ = smooth(l, b, x, b - A @ x)
        return x

    return V(0, b)


@pytest.mark.parametrize("floating", [True, False])
@pytest.mark.parametrize("n", [16, 18, 9])
@pytest.mark.parametrize("name,bs", [("A_demag", 1), ("A_el", 3), ("A1", 1)])
def test_vcycle_matches_host_reference(name, bs, n, floating):
    """16 -> 8 (nested), 18 -> 9 -> 5 (nested, then non-nested), 9 -> 5
    (non-nested at the top): the device transfers equal the host P.  A1
    (no anchor row) keeps node 0 in the transfers.  Floating hierarchies of
    the anchored operators: the V-cycle on the periodic stiffness (level 0
    recovered from the anchored operator on the device, here by unanchor),
    node 0 in the transfers, the pseudo-inverse on the coarsest level,
    applied as F^T V(F r) + e_0 r_0 with F = I - e_0 1^T; otherwise node 0
    is left out of the transfers."""
    from cupymag_amd.mesh.coarsen import hierarchy_sizes
    from cupymag_amd.solvers.multigrid import scalar_levels
    if floating and name == "A1":
        pytest.skip("A1 has no anchor: one kind of hierarchy")
    _, _, _, S = _setup(n=n)
    sizes = hierarchy_sizes(n) if bs == 3 else scalar_levels(hierarchy_sizes(n))
    # A1 both ways: down to the dense solve, and its default on these small
    # meshes, the first coarse level solved by Chebyshev steps
    for prec, tol, a1 in ((64, 1e-11, "dense"), (32, 1e-5, "dense"), (16, 3e-3, "dense"), (64, 1e-11, "auto"),
                          (16, 3e-3, "auto")):
        if a1 == "auto" and name != "A1":
            continue
        mg = S.multigrid(precision=prec, floating=floating, a1_coarse=a1)[name]
        info = mg.info()
        if a1 == "auto":
            from cupymag_amd.solvers.multigrid import A1_CHEB, A1_CHEB_RATIO
            assert info["levels"] == 2 and info["coarsest"] == ("chebyshev", A1_CHEB, A1_CHEB_RATIO), info
            mats = [_interleaved(L, bs) for L in mg.levels]
            r = np.random.default_rng(5).standard_normal(mats[0].shape[0])
            z = mg.apply(torch.as_tensor(r, device="cuda")).cpu().numpy()
            zr = _host_vcycle(mats, [(m, m, m) for m in sizes[:2]], info["lmax"], None, bs, mg.degree, mg.ratio, r,
                              anchor=False, coarse_smooth=info["coarsest"][1:])
            assert np.abs(z - zr).max() / np.abs(zr).max() < tol, prec
            continue
        assert info["levels"] == len(sizes) and info["dims"][-1] == (sizes[-1],) * 3
        assert info["coarsest"] == "dense"
        assert all(l > 0 for l in info["lmax"][:-1])
        # FP16 forms of the (slot-symmetric) periodic Kuhn levels keep half
        # of the stencil; the other precisions have no FP16 form
        if prec == 16 and (floating or name == "A1"):
            assert all(info["half_stencil"][:-1]), info["half_stencil"]
        elif prec != 16:
            assert not any(info["half_stencil"])
        mats = [_interleaved(L, bs) for L in mg.levels]
        dims = [(m, m, m) for m in sizes]
        # the elasticity hierarchy smooths its coarser levels with a higher degree
        degs = [mg.degree] + [info["coarse_degree"]] * (len(mats) - 1)
        assert info["coarse_degree"] == (3 if name == "A_el" else mg.degree)
        from cupymag_amd.solvers.multigrid import _interleave_perm
        Ac = mg.levels[-1].to_scipy().toarray()
        perm = _interleave_perm(Ac.shape[0] // bs, bs)
        anchored = name != "A1"
        fl = floating and anchored
        rng = np.random.default_rng(5)
        r = rng.standard_normal(mats[0].shape[0])
        if anchored and not fl:
            r[:bs] = 0.0  # anchor node
        z = mg.apply(torch.as_tensor(r, device="cuda")).cpu().numpy()
        if fl:
            # singular periodic levels: row 0 of level 1 is no identity row
            assert np.count_nonzero(mats[1][0].toarray()) > bs
            mats[0] = unanchor(mats[0], bs)
            cinv = np.linalg.pinv(Ac[np.ix_(perm, perm)])
            Fr = r.copy()
            for d in range(bs):
                Fr[d] = r[d] - r[d::bs].sum()
            y = _host_vcycle(mats, dims, info["lmax"], cinv, bs, degs, mg.ratio, Fr, anchor=False)
            zr = y.copy()
            for d in range(bs):
                zr[d::bs] -= y[d]
            zr[:bs] = r[:bs]
        else:
            cinv = np.linalg.inv(Ac[np.ix_(perm, perm)])
            zr = _host_vcycle(mats, dims, info["lmax"], cinv, bs, degs, mg.ratio, r, anchor=anchored)
        err = np.abs(z - zr).max() / np.abs(zr).max()
        assert err < tol, (prec, err)
        if fl:
            np.testing.assert_array_equal(z[:bs], r[:bs])
        elif anchored:
            assert np.all(z[:bs] == 0.0)
        elif prec == 64:
            # with node 0 left out of the transfers the V-cycle differs (so
            # the check above tells the two apart)
            zw = _host_vcycle(mats, dims, info["lmax"], cinv, bs, degs, mg.ratio, r, anchor=True)
            assert np.abs(zw - zr).max() > 1e-6 * np.abs(zr).max()


@pytest.mark.parametrize("n,jitter", [(16, 0.0), (16, 0.15), (21, 0.15)])
@pytest.mark.parametrize("name,bs", [("A_demag", 1), ("A_el", 3), ("A1", 1)])
def test_mg_pcg_matches_direct_solve(name, bs, n, jitter):
    _, _, _, S = _setup(n=n, jitter=jitter)
    for prec in (16, 32, 64):
        mg = S.multigrid(precision=prec)[name]
        A = _interleaved(mg.levels[0], bs)
        rng = np.random.default_rng(6)
        b = rng.standard_normal(A.shape[0])
        if name != "A1":
            b[:bs] = 0.0
        x, it, res = mg.solve(torch.as_tensor(b, device="cuda"), tol=1e-10, maxiter=500)
        x = x.cpu().numpy()
        assert np.linalg.norm(b - A @ x) <= 1.0001e-10 * np.linalg.norm(b)
        xr = spla.spsolve(A.tocsc(), b)
        assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-8
        assert 0 < it < 40, it
        if name == "A1":
            # three interleaved columns (the g1n..g3n solve), one zero column
            B = rng.standard_normal((A.shape[0], 3))
            B[:, 1] = 0.0
            X, its, ress = mg.solve(torch.as_tensor(B.ravel(), device="cuda"), tol=1e-10, maxiter=500, nrhs=3)
            X = X.cpu().numpy().reshape(-1, 3)
            assert its[1] == 0 and np.all(X[:, 1] == 0.0) and 0 < its[0] < 40 and 0 < its[2] < 40, its
            for j in (0, 2):
                assert np.linalg.norm(B[:, j] - A @ X[:, j]) <= 1.0001e-10 * np.linalg.norm(B[:, j])
                xr = spla.spsolve(A.tocsc(), B[:, j])
                assert np.abs(X[:, j] - xr).max() / np.abs(xr).max() < 1e-8
    z, it0, _ = mg.solve(torch.zeros(A.shape[0], dtype=torch.float64, device="cuda"), tol=1e-10)
    assert it0 == 0 and float(z.abs().max()) == 0.0
    with pytest.raises(RuntimeError, match="did not converge"):
        mg.solve(torch.as_tensor(b, device="cuda"), tol=1e-14, maxiter=2)


@pytest.mark.parametrize("mg_gs", [False, True])
@pytest.mark.parametrize("mode", ["fast64", "mixed"])
def test_gspm_steps_with_multigrid_match_oracle(mode, mg_gs):
    """mg_gs: the Gauss-Seidel solves on A1 (slots 2-6) run multigrid CG
    too; the A2 solves (slots 7-9) keep their iteration counts."""
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    cfg, arr, P, S = _setup(n=16)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    m0 = tilted_m(ops["n"], seed=8)
    sim = GSPMSimulation(S, P, mode=mode, precond="mg", mg_gs=mg_gs)
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
        assert err < (1e-8 if mode == "fast64" else 1e-7), (k, err)
        ref_it = [it for _, it in log]
        assert rep.iters[0] < 40 and rep.iters[1] < 40 and ref_it[0] > 3 * rep.iters[0]
        plain = slice(7, 10) if mg_gs else slice(2, 10)
        if mode == "fast64":
            assert all(abs(a - b) <= 1 for a, b in zip(rep.iters[plain], ref_it[plain])), (rep.iters, ref_it)
        if mg_gs:
            assert all(0 < it < 30 for it in rep.iters[2:7]), rep.iters
        assert np.linalg.norm(sim.get("UEL").cpu().numpy() - st["u"]) <= 1e-6 * np.linalg.norm(st["u"])


@pytest.mark.parametrize("mode,warm", [("fast64", False), ("mixed", False), ("mixed", "extrapolate")])
def test_demag_solve_beside_the_elasticity_solve_changes_nothing(mode, warm):
    """The demag solve on the context's second stream, beside the
    elasticity solve (overlap), gives every field, CG iteration count and
    step scalar of the two solves one after the other, bit for bit, over
    four steps; so does a failed step: with maxiter too small both report
    the elasticity solve's failure first."""
    from cupymag_amd.core.gspm import GSPMSimulation
    cfg, arr, P, S = _setup(n=16)
    m0 = tilted_m(S.n_dof, seed=21)
    H = [P.Hext1, P.Hext2, P.Hext3]
    sims = [GSPMSimulation(S, P, mode=mode, precond="mg", warm_start=warm, overlap=ov) for ov in (True, False)]
    for s in sims:
        s.set_m(m0)
    for k in range(4):
        reps = [s.step(H, False) for s in sims]
        assert reps[0].iters == reps[1].iters and reps[0].sumsq == reps[1].sumsq, (k, reps[0].iters, reps[1].iters)
        assert reps[0].iters[1] > 0
        for f in ("M", "U", "UEL", "E", "HTILDE"):
            assert torch.equal(sims[0].get(f), sims[1].get(f)), (k, f)
    saved = P.maxiter
    P.maxiter = 2
    try:
        errs = []
        for ov in (True, False):
            sim = GSPMSimulation(S, P, mode=mode, precond="mg", overlap=ov)
            sim.set_m(m0)
            with pytest.raises(RuntimeError, match="elasticity") as ei:
                sim.step(H, False)
            errs.append(str(ei.value))
        assert errs[0] == errs[1]
    finally:
        P.maxiter = saved


def test_timing_of_one_operator_leaves_the_others_untimed():
    """cpm_prof_only restricts the kernel timing to one operator: with it
    set to the elasticity operator, the mixed multigrid step times that
    operator's level-0 sweeps and outer products, and no launch of the
    Gauss-Seidel or demag operators; the timing changes no result bit."""
    import ctypes as C
    from cupymag_amd._lib import check, lib
    from cupymag_amd.core.gspm import GSPMSimulation
    cfg, arr, P, S = _setup(n=16)
    m0 = tilted_m(S.n_dof, seed=5)
    H = [P.Hext1, P.Hext2, P.Hext3]
    A_el, A1, Ad = S.csr_elasticity()[0], S.csr_GS()[0], S.csr_demag_and_deriv()[0]
    sims = [GSPMSimulation(S, P, mode="mixed", precond="mg", warm_start="extrapolate") for _ in range(2)]
    ctx = sims[0].ctx
    for s in sims:
        s.set_m(m0)
    for A in (A_el, A1, Ad):
        check(lib().cpm_mat_prof(A.h, None, None, 1))
        check(lib().cpm_mat_prof_outer(A.h, None, None, 1))

    def counts(A):
        n1, n2 = C.c_int64(), C.c_int64()
        check(lib().cpm_mat_prof(A.h, None, C.byref(n1), 0))
        check(lib().cpm_mat_prof_outer(A.h, None, C.byref(n2), 0))
        return n1.value, n2.value

    try:
        for k in range(3):
            check(lib().cpm_prof_enable(ctx.h, 1))
            check(lib().cpm_prof_only(ctx.h, A_el.h))
            r0 = sims[0].step(H, False)
            check(lib().cpm_prof_enable(ctx.h, 0))
            r1 = sims[1].step(H, False)
            assert r0.iters == r1.iters and r0.sumsq == r1.sumsq, (k, r0.iters, r1.iters)
            assert torch.equal(sims[0].get("M"), sims[1].get("M"))
    finally:
        check(lib().cpm_prof_enable(ctx.h, 0))
        check(lib().cpm_prof_only(ctx.h, None))
    n_el = counts(A_el)
    assert n_el[0] > 0 and n_el[1] > 0, n_el
    assert counts(A1) == (0, 0) and counts(Ad) == (0, 0)


def test_multigrid_rejections():
    from cupymag_amd.core.batch import GSPMBatch
    from cupymag_amd.core.gspm import GSPMSimulation
    _, _, P, S = _setup(n=16, me=False)
    with pytest.raises(ValueError, match="fast-mode"):
        GSPMSimulation(S, P, mode="fp64", precond="mg")
    sim = GSPMSimulation(S, P, mode="mixed", precond="mg")
    b = GSPMBatch(sim.ctx)
    b.add(sim)  # batched multigrid solves (test_gpu_batch_mg.py)
    with pytest.raises(ValueError, match="same mode"):
        b.add(GSPMSimulation(S, P, mode="fast64", precond="mg"))
    _, _, P7, S7 = _setup(n=8, me=False)
    with pytest.raises(ValueError, match="multigrid needs"):
        GSPMSimulation(S7, P7, mode="mixed", precond="mg")
    # a coarsest level solved by Chebyshev steps is for operators without an
    # anchor (A1): the anchored demag operator is refused
    from cupymag_amd.solvers.multigrid import KuhnMultigrid
    demag = S.multigrid(precision=32, floating=False)["A_demag"]
    with pytest.raises(ValueError, match="dense coarsest"):
        KuhnMultigrid(demag.levels, None, 32, coarse_smooth=(6, 20.0))
    with pytest.raises(ValueError, match="a1_coarse"):
        S.multigrid(precision=32, a1_coarse="sometimes")
    # coarse-level degrees: 2..8 under a finest degree >= 2
    a1 = S.multigrid(precision=32)["A1"]
    for bad in (1, 9):
        with pytest.raises(ValueError, match="degree must be"):
            KuhnMultigrid(a1.levels, None, 32, coarse_smooth=(6, 20.0), coarse_degree=bad)


@pytest.mark.parametrize("prec", [16, 64])
def test_graph_replay_matches_direct_launches(prec):
    """Iterations replayed from captured hipGraph chunks equal launching
    every kernel, bit for bit: several solves in a row per hierarchy (each
    replaying the cached chunk many times, with the solve's own launches and
    host reads of the state in between), one and three columns, and three
    GSPM steps with every solve on the multigrid."""
    from cupymag_amd.core.gspm import GSPMSimulation
    cfg, arr, P, S = _setup(n=12, jitter=0.1)
    hier = S.multigrid(precision=prec)
    rng = np.random.default_rng(21)
    cases = [("A_el", 1), ("A_demag", 1), ("A1", 1), ("A1", 3)]
    rhs = {c: [rng.standard_normal(hier[c[0]].levels[0].shape[0] * c[1]) for _ in range(3)] for c in cases}
    for c in cases:
        if c[0] != "A1":
            for b in rhs[c]:
                b[:3 if c[0] == "A_el" else 1] = 0.0
    out = {}
    for graphs in (False, True):
        for m in hier.values():
            m.set_graphs(graphs)
        res = []
        for (name, k) in cases:
            for b in rhs[(name, k)]:
                x, it, r = hier[name].solve(torch.as_tensor(b, device="cuda"), tol=1e-10, maxiter=500, nrhs=k)
                res.append((x.cpu().numpy(), it, r))
        out[graphs] = res
    for (x0, i0, r0), (x1, i1, r1) in zip(out[False], out[True]):
        assert i0 == i1 and r0 == r1
        np.testing.assert_array_equal(x0, x1)
    runs = []
    for graphs in (False, True):
        sim = GSPMSimulation(S, P, mode="mixed" if prec == 16 else "fast64", precond="mg")
        for m in sim.mg.values():
            m.set_graphs(graphs)
        sim.set_m(tilted_m(S.n_dof, seed=4))
        hist = []
        for k in range(3):
            rep = sim.step([P.Hext1, P.Hext2, P.Hext3], k > 0)
            hist.append((sim.get("M").cpu().numpy(), sim.get("UEL").cpu().numpy(), rep.iters))
        runs.append(hist)
    for (m0, u0, i0), (m1, u1, i1) in zip(*runs):
        assert i0 == i1
        np.testing.assert_array_equal(m0, m1)
        np.testing.assert_array_equal(u0, u1)


@pytest.mark.parametrize("name,bs,nrhs", [("A_el", 3, 1), ("A_demag", 1, 1), ("A1", 1, 3)])
def test_reliable_update_pcg_meets_the_fp64_tolerance(name, bs, nrhs):
    """FP32 outer products with FP64 residual replacements (delta 1e-3): the
    FP64 residual meets tol 1e-10 (the stopping test is on it), the solution
    equals a sparse direct solve to 1e-8, within two iterations of the FP64
    outer products."""
    _, _, _, S = _setup(n=16, jitter=0.1)
    mg = S.multigrid(precision=16)[name]
    A = _interleaved(mg.levels[0], bs)
    rng = np.random.default_rng(13)
    B = rng.standard_normal((A.shape[0], nrhs))
    if name != "A1":
        B[:bs] = 0.0
    b = torch.as_tensor(B.ravel(), device="cuda")
    x64, it64, _ = mg.solve(b, tol=1e-10, maxiter=500, nrhs=nrhs)
    x32, it32, res = mg.solve(b, tol=1e-10, maxiter=500, nrhs=nrhs, delta=1e-3)
    X = x32.cpu().numpy().reshape(-1, nrhs)
    it64 = it64 if nrhs > 1 else [it64]
    it32 = it32 if nrhs > 1 else [it32]
    for j in range(nrhs):
        assert np.linalg.norm(B[:, j] - A @ X[:, j]) <= 1e-10 * np.linalg.norm(B[:, j])
        xr = spla.spsolve(A.tocsc(), B[:, j])
        assert np.abs(X[:, j] - xr).max() / np.abs(xr).max() < 1e-8
        assert abs(it32[j] - it64[j]) <= 2, (it32, it64)


@pytest.mark.parametrize("name,bs,below", [("A_el", 3, True), ("A_demag", 1, False)])
def test_reliable_update_pcg_at_tolerances_below_the_fp64_floor(name, bs, below):
    """tol 1e-14 at 64^3: below what FP64 products attain on the anchored
    elasticity operator (its true residual bottoms out near
    eps64 ||A|| ||x||, ~8e-14 of ||b||; the demag operator still reaches
    it).  The FP64-product PCG stops on its recurrence residual (the
    reference's rule) with the true residual at that floor; the mixed PCG,
    whose stopping test is on the FP64 residual, switches its products to
    FP64 and ends at the same floor as converged, within a few iterations of
    the FP64 products, never with a non-finite residual."""
    _, _, _, S = _setup(n=64, jitter=0.1)
    mg = S.multigrid(precision=16)[name]
    A = _interleaved(mg.levels[0], bs)
    B = np.random.default_rng(21).standard_normal(A.shape[0])
    B[:bs] = 0.0
    b = torch.as_tensor(B, device="cuda")
    nb = np.linalg.norm(B)
    x64, it64, _ = mg.solve(b, tol=1e-14, maxiter=300)
    floor = np.linalg.norm(B - A @ x64.cpu().numpy()) / nb
    if below:
        assert floor > 1e-14                               # the recurrence went below what x attains
    x32, it32, res = mg.solve(b, tol=1e-14, maxiter=300, delta=1e-3)
    true = np.linalg.norm(B - A @ x32.cpu().numpy()) / nb
    # at the floor b - A x is rounding noise itself: the device's and
    # SciPy's summation orders agree to a small factor, not to digits
    assert np.isfinite(res) and true / 3.0 <= res / nb <= 3.0 * true
    assert true <= max(4.0 * floor, 1e-14), (true, floor)
    assert it32 <= it64 + 12, (it32, it64)
    # at an attainable tolerance the products stay FP32 and the counts match
    _, it_a, _ = mg.solve(b, tol=1e-10, maxiter=300)
    _, it_b, _ = mg.solve(b, tol=1e-10, maxiter=300, delta=1e-3)
    assert abs(it_a - it_b) <= 2


@pytest.mark.parametrize("name,bs", [("A_demag", 1), ("A_el", 3), ("A1", 1)])
def test_vcycle_is_symmetric_positive(name, bs):
    """The V-cycle is a symmetric positive definite preconditioner: the same
    Chebyshev polynomial in one D^-1 A before and after the coarse
    correction (for the FP16 level forms, every step inverts the same
    FP32-rounded diagonal blocks), R = P^T, an exact coarsest solve.  So
    u.Mv = v.Mu up to the rounding of the V-cycle's vectors (FP64: 1e-12,
    FP32: 1e-5 of |u| |Mv|) and u.Mu > 0."""
    _, _, _, S = _setup(n=18, jitter=0.1)
    rng = np.random.default_rng(8)
    for prec, tol in ((64, 1e-12), (32, 1e-5), (16, 1e-5)):
        mg = S.multigrid(precision=prec)[name]
        N = mg.levels[0].shape[0]
        for _ in range(3):
            u, v = rng.standard_normal(N), rng.standard_normal(N)
            Mu = mg.apply(torch.as_tensor(u, device="cuda")).cpu().numpy()
            Mv = mg.apply(torch.as_tensor(v, device="cuda")).cpu().numpy()
            assert abs(u @ Mv - v @ Mu) <= tol * np.linalg.norm(u) * np.linalg.norm(Mv), (prec, u @ Mv, v @ Mu)
            assert u @ Mu > 0 and v @ Mv > 0
