"""GPU parity of the HIP engine against the CPU oracle (same seeded inputs).

Tolerances (FP64 path): SpMV 1e-14 relative; CG solutions 1e-10 relative
with iteration counts within +-1 of the oracle's CuPy-rule CG; field
operators 1e-12 relative; full GSPM steps: magnetization per node within
1e-10, displacement within 1e-10 relative, CG iteration counts within +-1.
The operators themselves come from two independent assemblies (native
host assembler vs the oracle's SciPy COO path), so these checks also cover
the assembly.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m  # noqa: E402


def _product(cfg, tet_arrays=None):
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    P = from_dict(cfg)
    return P, SimulationOperators(P, mesh_arrays=tet_arrays, keep_host=True)


def _kuhn(n, jitter=0.15, seed=3):
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    return kuhn_mesh(n, jitter=jitter, seed=seed)


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_spmv_matches_scipy():
    cfg = config("Tet", me=True, defect_params=1.5)
    arr = _kuhn(6)
    P, S = _product(cfg, arr)
    A_el, F_el = S.csr_elasticity()
    A1, A2, F1 = S.csr_GS()
    rng = np.random.default_rng(0)
    for A in (A_el, F_el, A1, F1):
        x = rng.standard_normal(A.shape[1])
        y = (A @ torch.as_tensor(x, device="cuda")).cpu().numpy()
        ref = A.to_scipy() @ x
        assert _rel(y, ref) < 1e-14


@pytest.mark.parametrize("which", ["A1", "A_demag", "A_el"])
def test_cg_matches_oracle(which):
    from cupymag_amd.solvers.linear_solvers import solve_cg
    from oracle import gspm as og
    cfg = config("Tet", me=True, defect_params=1.8)
    arr = _kuhn(6)
    P, S = _product(cfg, arr)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    A_dev = {"A1": S.csr_GS()[0], "A_demag": S.csr_demag_and_deriv()[0], "A_el": S.csr_elasticity()[0]}[which]
    A_ref = ops[which]
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A_ref.shape[0])
    if which == "A_demag":
        b[0] = 0.0
    x, it, _ = solve_cg(A_dev, torch.as_tensor(b, device="cuda"), tol=1e-10, maxiter=5000, return_info=True)
    xr, info, itr = og.cg(A_ref, b, None, 1e-10, 5000)
    assert info == 0
    assert abs(it - itr) <= 1
    assert _rel(x.cpu().numpy(), xr) < 1e-10
    # warm start from a perturbed solution
    x0 = xr + 1e-6 * rng.standard_normal(len(b))
    x2, it2, _ = solve_cg(A_dev, torch.as_tensor(b, device="cuda"), x0=torch.as_tensor(x0, device="cuda"),
                          use_init=True, tol=1e-10, return_info=True)
    xr2, _, itr2 = og.cg(A_ref, b, x0, 1e-10, 5000)
    assert abs(it2 - itr2) <= 1
    assert _rel(x2.cpu().numpy(), xr2) < 1e-10


@pytest.mark.parametrize("which", ["A1", "A_demag", "A_el"])
def test_preconditioned_cg_matches_oracle(which):
    """solve_cg(M=...) forwards the preconditioner as the reference does to
    cupyx cg (linear_solvers.py:59-62): with a Jacobi M (SciPy sparse, and
    as an uploaded device operator) solutions match the oracle's
    preconditioned CG to 1e-10, iterations +-1, cold and warm; M = I gives
    the unpreconditioned iteration count."""
    import scipy.sparse as sps

    from cupymag_amd.device import DeviceCSR
    from cupymag_amd.solvers.linear_solvers import solve_cg
    from oracle import gspm as og
    cfg = config("Tet", me=True, defect_params=1.8)
    arr = _kuhn(6)
    P, S = _product(cfg, arr)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    A_dev = {"A1": S.csr_GS()[0], "A_demag": S.csr_demag_and_deriv()[0], "A_el": S.csr_elasticity()[0]}[which]
    A_ref = ops[which]
    n = A_ref.shape[0]
    J = sps.diags(1.0 / A_ref.diagonal()).tocsr()
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n)
    if which == "A_demag":
        b[0] = 0.0
    bd = torch.as_tensor(b, device="cuda")
    xr, info, itr = og.cg(A_ref, b, None, 1e-10, 5000, M=J)
    assert info == 0
    Jd = DeviceCSR(J.indptr, J.indices, J.data, J.shape)
    for M in (J, Jd):
        x, it, _ = solve_cg(A_dev, bd, M=M, tol=1e-10, maxiter=5000, return_info=True)
        assert abs(it - itr) <= 1
        assert _rel(x.cpu().numpy(), xr) < 1e-10
    x0 = xr + 1e-6 * rng.standard_normal(n)
    x2, it2, _ = solve_cg(A_dev, bd, M=Jd, x0=torch.as_tensor(x0, device="cuda"), use_init=True, tol=1e-10,
                          return_info=True)
    xr2, _, itr2 = og.cg(A_ref, b, x0, 1e-10, 5000, M=J)
    assert abs(it2 - itr2) <= 1 and _rel(x2.cpu().numpy(), xr2) < 1e-10
    _, itI, _ = solve_cg(A_dev, bd, M=sps.identity(n, format="csr"), tol=1e-10, return_info=True)
    _, it0, _ = solve_cg(A_dev, bd, tol=1e-10, return_info=True)
    assert itI == it0
    with pytest.raises(TypeError, match="M must be a sparse or dense matrix"):
        solve_cg(A_dev, bd, M=lambda r: r)
    with pytest.raises(RuntimeError, match="did not converge"):
        solve_cg(A_dev, bd, M=Jd, tol=1e-14, maxiter=2, system="demag")


def test_cg_edge_cases():
    from cupymag_amd.solvers.linear_solvers import solve_cg
    cfg = config("Hex", n=6)
    P, S = _product(cfg)
    A1 = S.csr_GS()[0]
    n = A1.shape[0]
    # ||b|| == 0 returns b (zeros), also with a warm start
    z = torch.zeros(n, dtype=torch.float64, device="cuda")
    x, it, _ = solve_cg(A1, z, x0=torch.ones(n, dtype=torch.float64, device="cuda"), use_init=True,
                        return_info=True)
    assert it == 0 and float(x.abs().max()) == 0.0
    # non-convergence raises the reference's RuntimeError
    b = torch.as_tensor(np.random.default_rng(2).standard_normal(n), device="cuda")
    with pytest.raises(RuntimeError, match=r"Error! CG for test did not converge. info=2, residual="):
        solve_cg(A1, b, tol=1e-14, maxiter=2, system="test")


@pytest.mark.parametrize("rot", [False, True])
def test_field_operators(rot):
    from cupymag_amd.utils.final_assembly import build_E0_from_m
    from oracle import gspm as og
    from oracle.params import rot111_M, rot111_R
    cfg = config("Tet", me=True, rot=rot, defect_params=1.5)
    arr = _kuhn(5)
    P, S = _product(cfg, arr)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    n = ops["n"]
    rng = np.random.default_rng(5)
    _, Fx, Fy, Fz, Deriv = S.csr_demag_and_deriv()
    U = rng.standard_normal(n)
    H = Deriv.compute_Hd_from_U(torch.as_tensor(U, device="cuda"))
    Hr = og.hd_from_U(ops, U)
    for k in range(3):
        assert _rel(H[k].cpu().numpy(), Hr[k]) < 1e-12
    u = rng.standard_normal(3 * n)
    R = rot111_R() if rot else None
    E = Deriv.compute_E_from_u(n, torch.as_tensor(u, device="cuda"), R).cpu().numpy()
    assert _rel(E, og.strain_from_u(ops, n, u, R)) < 1e-12
    m = tilted_m(n, 1, 40.0)
    if rot:
        E0 = build_E0_from_m(p["lambda100"], p["lambda111"], m, rot111_R(), rot111_M()).cpu().numpy()
        ref = (og.build_E0(p["lambda100"], p["lambda111"], m @ rot111_R()).reshape(n, 6) @ rot111_M().T).ravel()
    else:
        E0 = build_E0_from_m(p["lambda100"], p["lambda111"], m).cpu().numpy()
        ref = og.build_E0(p["lambda100"], p["lambda111"], m)
    assert _rel(E0, ref) < 1e-13
    Avgp = S.volume_average()
    f = rng.standard_normal((n, 3))
    assert _rel(Avgp.compute_average_field_gpu(torch.as_tensor(f, device="cuda")), Avg.average(f)) < 1e-11
    g = rng.standard_normal(n)
    assert abs(Avgp.compute_average_field_gpu(g) - Avg.average(g)) < 1e-12 * max(1.0, abs(Avg.average(g)))


STEP_CASES = [
    dict(grid="Hex", n=8, me=False, rot=False),
    dict(grid="Hex", n=6, me=True, rot=False, sigma11=1e8),
    dict(grid="Tet", n=6, me=True, rot=True, dt=5e-12),
    dict(grid="Tet", n=7, me=False, rot=True, dt=1e-11, alpha=0.1),
    dict(grid="Tet", n=6, me=True, rot=False, dt=5e-12, K1=2e4, hext=(2e4, 3e4, -1e4)),
]


@pytest.mark.parametrize("case", STEP_CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_gspm_steps_match_oracle(case):
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    case = dict(case)
    grid, n = case.pop("grid"), case.pop("n")
    cfg = config(grid, n=n, defect_params=1.6, **case)
    arr = _kuhn(n) if grid == "Tet" else None
    P, S = _product(cfg, arr)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    nd = ops["n"]
    m0 = tilted_m(nd, seed=11)
    sim = GSPMSimulation(S, P)
    am = sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    assert _rel(am, st["avg_m"]) < 1e-12
    p["H_now"] = p["H"]
    H = [P.Hext1, P.Hext2, P.Hext3]
    for k in range(4):
        log = []
        ui = st["use_init"]
        og.step(st, ops, Avg, p, log)
        rep = sim.step(H, ui)
        m = sim.get("M").cpu().numpy()
        assert np.max(np.abs(m - st["m"])) < 1e-10, f"step {k}"
        assert abs(rep.sumsq - st["sumsq"]) <= 1e-9 * max(st["sumsq"], 1e-30) + 1e-22
        ref_it = [it for _, it in log]
        got = [it for it, s in zip(rep.iters, range(10)) if (p["ME"] or s > 0)]
        assert all(abs(a - b) <= 1 for a, b in zip(got, ref_it)), (got, ref_it)
        if p["ME"]:
            u = sim.get("UEL").cpu().numpy()
            assert _rel(u, st["u"]) < 1e-10
            assert _rel(sim.get("E").cpu().numpy(), st["E"]) < 1e-10
        Ht = sim.get("HTILDE").cpu().numpy()
        assert _rel(Ht, np.stack(st["Ht"])) < 1e-10
        assert _rel(rep.avg_m, st["avg_m"]) < 1e-10


def test_warm_start_path_matches_oracle():
    """use_init switches on after the first step (large initial_factor)."""
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    cfg = config("Hex", n=6, me=True)
    cfg["simulation"]["llg"]["llg_accuracy_factor"] = 1e-3
    cfg["simulation"]["solver"]["initial_factor"] = 1e6
    P, S = _product(cfg)
    p, ops, Avg, _ = oracle_setup(cfg)
    m0 = tilted_m(ops["n"], seed=4)
    sim = GSPMSimulation(S, P)
    sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    H = [P.Hext1, P.Hext2, P.Hext3]
    use_init = False
    for k in range(3):
        og.step(st, ops, Avg, p)
        rep = sim.step(H, use_init)
        use_init = rep.sumsq < P.use_init_factor * ops["n"] * P.LLG_accuracy_factor
        assert use_init == st["use_init"]
        assert np.max(np.abs(sim.get("M").cpu().numpy() - st["m"])) < 1e-10


# ---------------------------------------------------------------- fast paths

def _interleave(cols):
    return np.stack(cols, axis=1).ravel()


def test_sell_spmv_matches_scipy():
    from cupymag_amd._lib import check, lib, ptr
    cfg = config("Tet", me=True, defect_params=1.5)
    P, S = _product(cfg, _kuhn(6))
    A_el = S.csr_elasticity()[0]
    A1 = S.csr_GS()[0]
    n = A1.shape[0]
    rng = np.random.default_rng(7)
    # scalar, 3 interleaved right-hand sides
    X = rng.standard_normal((n, 3))
    y = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    check(lib().cpm_spmv_sell(A1.ctx.h, A1.h, 3, ptr(torch.as_tensor(X.ravel(), device="cuda")), ptr(y)))
    ref = A1.to_scipy() @ X
    assert _rel(y.cpu().numpy().reshape(n, 3), ref) < 1e-14
    # 3x3 blocks on node-major vectors
    u = rng.standard_normal(3 * n)
    y = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    ui = _interleave([u[:n], u[n:2 * n], u[2 * n:]])
    check(lib().cpm_spmv_sell(A_el.ctx.h, A_el.h, 1, ptr(torch.as_tensor(ui, device="cuda")), ptr(y)))
    ref = A_el.to_scipy() @ u
    assert _rel(y.cpu().numpy(), _interleave([ref[:n], ref[n:2 * n], ref[2 * n:]])) < 1e-14


@pytest.mark.parametrize("which", ["A1x3", "A_el"])
def test_sell_cg_matches_oracle(which):
    from cupymag_amd.solvers.linear_solvers import solve_cg_sell
    from oracle import gspm as og
    cfg = config("Tet", me=True, defect_params=1.8)
    arr = _kuhn(6)
    P, S = _product(cfg, arr)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    n = ops["n"]
    rng = np.random.default_rng(8)
    if which == "A1x3":
        A = S.csr_GS()[0]
        B = rng.standard_normal((n, 3))
        x, it, _ = solve_cg_sell(A, B.ravel(), tol=1e-10, nrhs=3)
        X = x.cpu().numpy().reshape(n, 3)
        for j in range(3):
            xr, _, itr = og.cg(ops["A1"], B[:, j], None, 1e-10, 5000)
            assert abs(it[j] - itr) <= 1 and _rel(X[:, j], xr) < 1e-10
        # mixed precision: FP64 residual meets the tolerance, error small
        x32, it32, res32 = solve_cg_sell(A, B.ravel(), tol=1e-10, nrhs=3, precision=32)
        X32 = x32.cpu().numpy().reshape(n, 3)
        for j in range(3):
            xr, _, _ = og.cg(ops["A1"], B[:, j], None, 1e-10, 5000)
            assert np.linalg.norm(B[:, j] - ops["A1"] @ X32[:, j]) <= 1e-10 * np.linalg.norm(B[:, j])
            assert _rel(X32[:, j], xr) < 1e-8
    else:
        A = S.csr_elasticity()[0]
        b = rng.standard_normal(3 * n)
        b[[0, n, 2 * n]] = 0.0
        bi = _interleave([b[:n], b[n:2 * n], b[2 * n:]])
        x, it, _ = solve_cg_sell(A, bi, tol=1e-10)
        xr, _, itr = og.cg(ops["A_el"], b, None, 1e-10, 5000)
        assert abs(it[0] - itr) <= 1
        assert _rel(x.cpu().numpy(), _interleave([xr[:n], xr[n:2 * n], xr[2 * n:]])) < 1e-10
        x32, _, _ = solve_cg_sell(A, bi, tol=1e-10, precision=32)
        assert _rel(x32.cpu().numpy(), _interleave([xr[:n], xr[n:2 * n], xr[2 * n:]])) < 1e-8


@pytest.mark.parametrize("mode", ["fast64", "mixed", "mixed+warm", "mixed+xwarm"])
@pytest.mark.parametrize("case", [dict(grid="Tet", n=6, me=True, rot=True, dt=5e-12),
                                  dict(grid="Hex", n=6, me=True, rot=False, sigma11=1e8)],
                         ids=["tet-me-rot", "hex-me"])
def test_fast_modes_match_oracle(mode, case):
    """fast64: same tolerances as the plain FP64 path (m 1e-10, CG iterations
    +-1).  mixed (and mixed with warm starts from the last solution or from
    the extrapolation of the last two): magnetization within 1e-7 of the
    FP64 oracle over four steps (the extrapolation first applies at the
    third)."""
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    case = dict(case)
    grid, n = case.pop("grid"), case.pop("n")
    cfg = config(grid, n=n, defect_params=1.6, **case)
    arr = _kuhn(n) if grid == "Tet" else None
    P, S = _product(cfg, arr)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    m0 = tilted_m(ops["n"], seed=12)
    base, _, opt = mode.partition("+")
    sim = GSPMSimulation(S, P, mode=base, warm_start={"warm": True, "xwarm": "extrapolate"}.get(opt, False))
    sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    H = [P.Hext1, P.Hext2, P.Hext3]
    for k in range(4):
        log = []
        ui = st["use_init"]
        og.step(st, ops, Avg, p, log)
        rep = sim.step(H, ui)
        err = np.max(np.abs(sim.get("M").cpu().numpy() - st["m"]))
        if mode == "fast64":
            assert err < 1e-10, f"step {k}: {err}"
            ref_it = [it for _, it in log]
            assert all(abs(a - b) <= 1 for a, b in zip(rep.iters, ref_it)), (rep.iters, ref_it)
            assert _rel(sim.get("UEL").cpu().numpy(), st["u"]) < 1e-10
        else:
            assert err < 1e-7, f"step {k}: {err}"


@pytest.mark.parametrize("grid,precision,engine", [("Hex", "DP", "auto"), ("Hex", "DP", "fp64"),
                                                   ("Hex", "SP", "auto"), ("Tet", "SP", "auto")])
def test_cli_hysteresis_loop_matches_oracle(tmp_path, grid, precision, engine):
    """`cupymag config.yaml` end to end: field points of hysteresis.txt (LLG
    convergence -> Hext -= dHext) match the oracle's loop, VTU snapshots and
    last_m are written.  simulation.precision picks the engine: DP runs
    fast64 (FP64 CG, the oracle's iterates to rounding; --engine fp64 the
    CSR path), SP the mixed-precision engine (within 1e-7), on a Kuhn .nas
    mesh of 10^3 cells with the multigrid solves (there the loop ends on
    the step cap rather than on the stop condition)."""
    import yaml

    from cupymag_amd.cli import main
    from cupymag_amd.mesh.kuhn import kuhn_mesh, write_nas
    from oracle import gspm as og
    arrays = None
    if grid == "Hex":
        cfg = config("Hex", n=6, me=False, ndef=2, hext=(-2e4, 0.0, 0.0))
    else:
        arrays = kuhn_mesh(10, jitter=0.1, seed=6)
        write_nas(tmp_path / "mesh.nas", *arrays)
        cfg = config("Tet", n=10, me=False, hext=(2e4, 0.0, 0.0), defect_params=2.0)
        cfg["grid"]["tet_mesh"]["mesh_file"] = "mesh.nas"
    cfg["simulation"]["precision"] = precision
    cfg["physics"]["external_field"]["dhext1"] = 2e4
    cfg["simulation"]["llg"].update(llg_accuracy_factor=1e-6, timestep=5e-12)
    cfg["simulation"]["stop_condition"] = {"component": 0, "value": -0.5}
    cfg["output"] = {"directory": "out", "write_magnetization": True, "save_frequency": 100000,
                     "vtu_blend_alpha": 0.5}
    path = tmp_path / "cfg.yaml"
    path.write_text(yaml.safe_dump(cfg))
    assert main([str(path), "--max-steps", "400", "--engine", engine]) == 0
    out = tmp_path / "out"
    lines = (out / "hysteresis.txt").read_text().splitlines()
    assert lines[0] == "Hext1\tHext2\tHext3\tavg_m1\tavg_m2\tavg_m3"
    assert list(out.glob("field_*.vtu"))
    from cupymag_amd.utils.magnetization_io import read_array
    ndof = 6 ** 3 if grid == "Hex" else 10 ** 3
    last = read_array(out / "last_m.h5", ndof)  # an HDF5 file (h5py, or the built-in writer)
    nrm = np.linalg.norm(last, axis=1)
    assert np.all((np.abs(nrm - 1) < 1e-12) | (nrm == 0))
    # oracle: same control flow, in FP64
    cfg_o = dict(cfg, simulation=dict(cfg["simulation"], precision="DP"))
    p, ops, Avg, _ = oracle_setup(cfg_o, arrays)
    st = og.init_state(ops, Avg, og.uniform_m(ops["n"]))
    pts = og.run(st, ops, Avg, p, 400)
    ms = cfg["physics"]["magnetic"]["ms"]
    got = [list(map(float, l.split("\t"))) for l in lines[1:]]
    # one line per LLG convergence, then the final state
    assert len(pts) >= 1 and len(got) == len(pts) + 1
    atol = 1e-9 if precision == "DP" else 1e-7
    for (H, am), g in zip(pts, got):
        assert abs(H[0] * ms - g[0]) < 0.1
        assert np.allclose(am, g[3:], atol=atol)
    # the last line is the final state: at the next field point, where the
    # loop stopped on the stop condition (m1 < -0.5) or on the step cap
    assert abs(got[-1][0] - (pts[-1][0][0] - p["dH"][0]) * ms) < 0.1
    if st["avg_m"][0] <= -0.5:
        assert got[-1][3] <= -0.5
    else:
        assert np.allclose(got[-1][3:], st["avg_m"], atol=atol)


@pytest.mark.parametrize("grid", ["Tet", "Hex"])
def test_stencil_and_indexed_sell_agree(grid):
    """The index-free SELL form (structured periodic grid; on the Kuhn
    mesh with the neighbour offsets decoded at compile time, on the hex grid
    with the 26 offsets read per slot) and the indexed form give the same
    products; a mesh with permuted node numbering falls back to the indexed
    form.  On the Kuhn mesh the FP32 form keeps half of the (slot-symmetric)
    stencil: the mixed-precision CG on it and on the indexed form (every
    slot in FP32) take about the same iterations (within 25%) to the same
    solution."""
    from cupymag_amd._lib import check, lib, ptr
    from cupymag_amd.device import DeviceCSR
    cfg = config(grid, n=6, me=True, defect_params=1.5)
    P, S = _product(cfg, _kuhn(6) if grid == "Tet" else None)
    assert S.grid_dims == (6, 6, 6)
    A_el = S.csr_elasticity()[0]
    A1 = S.csr_GS()[0]
    assert A_el.sell_info()["stencil"] and A1.sell_info()["stencil"]
    assert A_el.sell_info()["kuhn"] == A1.sell_info()["kuhn"] == (grid == "Tet")
    assert A_el.sell_info()["half32"] == A1.sell_info()["half32"] == (grid == "Tet")
    assert A_el.sell_info()["half64"] == A1.sell_info()["half64"] == (grid == "Tet")
    rng = np.random.default_rng(3)
    for A, bs, K in ((A_el, 3, 1), (A1, 1, 3)):
        ip, ix, v = A.to_arrays()
        B = DeviceCSR(ip, ix, v, A.shape).add_sell(bs)
        assert not B.sell_info()["stencil"]
        x = torch.as_tensor(rng.standard_normal(A.shape[0] * (K if bs == 1 else 1)), device="cuda")
        y1 = torch.empty_like(x)
        y2 = torch.empty_like(x)
        check(lib().cpm_spmv_sell(A.ctx.h, A.h, K, ptr(x), ptr(y1)))
        check(lib().cpm_spmv_sell(B.ctx.h, B.h, K, ptr(x), ptr(y2)))
        assert _rel(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-14
        sols = []
        for M in (A, B):
            xs = torch.zeros_like(x)
            it, res = (C.c_int32 * K)(), (C.c_double * K)()
            check(lib().cpm_cg_sell(M.ctx.h, M.h, K, ptr(x), ptr(xs), 0, 1e-10, 5000, 32, 0.3, it, res))
            sols.append((list(it), xs.cpu().numpy()))
        (it_a, x_a), (it_b, x_b) = sols
        # the two forms sum in different orders, and the FP32 recurrence of
        # the elasticity solve is sensitive to that rounding (Hex: 71 vs 78)
        assert all(abs(p - q) <= max(2, 0.25 * q) for p, q in zip(it_a, it_b)), (it_a, it_b)
        assert _rel(x_a, x_b) < 1e-8


def test_unstructured_numbering_uses_indexed_sell():
    from oracle import gspm as og
    from cupymag_amd.core.gspm import GSPMSimulation
    coords, tets = _kuhn(5)
    perm = np.random.default_rng(1).permutation(len(coords))
    inv = np.argsort(perm)
    arr = (coords[perm], inv[tets].astype(np.int32))
    cfg = config("Tet", n=5, me=True, rot=True, defect_params=1.4, dt=5e-12)
    P, S = _product(cfg, arr)
    assert S.grid_dims is None
    assert not S.csr_elasticity()[0].sell_info()["stencil"]
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    m0 = tilted_m(ops["n"], seed=2)
    sim = GSPMSimulation(S, P, mode="fast64")
    sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    og.step(st, ops, Avg, p)
    sim.step([P.Hext1, P.Hext2, P.Hext3], False)
    assert np.max(np.abs(sim.get("M").cpu().numpy() - st["m"])) < 1e-10


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_device_control_equals_host_control(mode):
    """The hysteresis control on the device (set_control) drives the same
    trajectory as the host applying the reference's rules
    (Micromagnetics.py:299-335) around step(H, use_init): the same field
    points at the same steps, and m bit-identical after them."""
    from cupymag_amd.core.gspm import GSPMSimulation
    cfg = config("Hex", n=6, me=False, ndef=2, hext=(-2e4, 0.0, 0.0))
    cfg["physics"]["external_field"]["dhext1"] = 2e4
    cfg["simulation"]["llg"].update(llg_accuracy_factor=1e-6, timestep=5e-12)
    P, S = _product(cfg)
    m0 = tilted_m(S.n_dof, seed=12)
    a, b = GSPMSimulation(S, P, mode=mode), GSPMSimulation(S, P, mode=mode)
    a.set_m(m0)
    b.set_m(m0)
    b.set_control(H0=[P.Hext1, P.Hext2, P.Hext3], dH=[P.dHext1, P.dHext2, P.dHext3], stop_value=-2.0)
    acc = S.n_dof * P.LLG_accuracy_factor
    H = np.array([P.Hext1, P.Hext2, P.Hext3])
    use_init, nstep, pts_a, pts_b = False, 0, [], []
    for k in range(300):
        nstep += 1
        ra = a.step(H, use_init)
        rb = b.step()
        np.testing.assert_array_equal(rb.H, H)
        use_init = ra.sumsq < P.use_init_factor * acc
        assert rb.use_init == use_init and rb.sumsq == ra.sumsq
        if ra.sumsq < acc and nstep > 100:
            pts_a.append(k)
            H = H - np.array([P.dHext1, P.dHext2, P.dHext3])
            nstep = 0
        if rb.point_done:
            pts_b.append(k)
        assert rb.nstep == nstep
    assert pts_a == pts_b and len(pts_a) >= 1
    assert torch.equal(a.get("M"), b.get("M"))
