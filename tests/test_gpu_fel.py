"""Matrix-free F_el and Fx, Fy, Fz (cpm_fel_*): the product F_el E0 and the
derivative products (Fk x, the strain of compute_E_from_u with and without
the [111] rotation, Htilde of compute_Hd_from_U) computed from the element
geometry equal the assembled CSR operators' products to 1e-13 relative
(regular and jittered Kuhn meshes, cubic and [111]-rotated stiffness), and
GSPM steps using them match the oracle to the plain FP64 path's
tolerances."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config, oracle_setup, tilted_m  # noqa: E402


def _ops(n, jitter, rot, matrix_free, perm=None, me=True, deriv=False):
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = config("Tet", n=n, me=me, rot=rot, dt=5e-12, defect_params=1.6)
    coords, tets = kuhn_mesh(n, jitter=jitter, seed=7)
    if perm is not None:
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        coords, tets = coords[perm], inv[tets].astype(np.int32)
    P = from_dict(cfg)
    S = SimulationOperators(P, mesh_arrays=(coords, tets), keep_host=True, matrix_free_fel=matrix_free,
                            matrix_free_deriv=deriv)
    return cfg, (coords, tets), P, S


@pytest.mark.parametrize("n,jitter,rot", [(8, 0.0, False), (10, 0.0, True), (8, 0.15, True), (12, 0.15, False)])
def test_matrix_free_product_equals_csr(n, jitter, rot):
    """Regular meshes (the stencil form of the product) and jittered ones
    (the element walk), cubic and [111]-rotated C."""
    from cupymag_amd.physics.fel import MatrixFreeFel
    _, _, _, Sc = _ops(n, jitter, rot, False)
    _, _, _, Sm = _ops(n, jitter, rot, True)
    Fc = Sc.csr_elasticity()[1]
    Fm = Sm.csr_elasticity()[1]
    assert isinstance(Fm, MatrixFreeFel) and Fm.shape == Fc.shape
    rng = np.random.default_rng(3)
    E0 = rng.standard_normal(Fc.shape[1])
    yc = (Fc @ torch.as_tensor(E0, device="cuda")).cpu().numpy()
    ym = (Fm @ torch.as_tensor(E0, device="cuda")).cpu().numpy()
    ref = Fc.to_scipy() @ E0
    assert np.abs(yc - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.abs(ym - ref).max() <= 1e-13 * np.abs(ref).max()
    N = Fc.shape[0] // 3
    assert ym[0] == 0.0 and ym[N] == 0.0 and ym[2 * N] == 0.0


def test_unordered_mesh_keeps_csr():
    from cupymag_amd.device import DeviceCSR
    perm = np.random.default_rng(0).permutation(9 ** 3)
    _, _, _, S = _ops(8, 0.0, False, True, perm=perm)
    assert isinstance(S.csr_elasticity()[1], DeviceCSR)


@pytest.mark.parametrize("rot,jitter", [(False, 0.15), (True, 0.15), (True, 0.0)])
def test_steps_with_matrix_free_fel_match_oracle(rot, jitter):
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    cfg, arr, P, S = _ops(6, jitter, rot, True)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    m0 = tilted_m(ops["n"], seed=9)
    sim = GSPMSimulation(S, P, mode="fp64")
    sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    H = [P.Hext1, P.Hext2, P.Hext3]
    for k in range(3):
        log = []
        ui = st["use_init"]
        og.step(st, ops, Avg, p, log)
        rep = sim.step(H, ui)
        assert np.max(np.abs(sim.get("M").cpu().numpy() - st["m"])) < 1e-10
        u = sim.get("UEL").cpu().numpy()
        assert np.abs(u - st["u"]).max() <= 1e-10 * np.abs(st["u"]).max()
        ref_it = [it for _, it in log]
        assert all(abs(a - b) <= 1 for a, b in zip(rep.iters, ref_it)), (rep.iters, ref_it)


@pytest.mark.parametrize("n,jitter", [(8, 0.0), (11, 0.0), (8, 0.15), (12, 0.15)])
def test_matrix_free_derivatives_equal_csr(n, jitter):
    """Regular meshes (the 27-point stencil form) and jittered ones (the
    element walk)."""
    from cupymag_amd.physics.fel import FelDerivView
    from oracle.params import rot111_R
    _, _, _, Sc = _ops(n, jitter, False, False)
    _, _, _, Sm = _ops(n, jitter, False, True, deriv=True)
    _, Fxc, Fyc, Fzc, Dc = Sc.csr_demag_and_deriv()
    _, Fxm, Fym, Fzm, Dm = Sm.csr_demag_and_deriv()
    assert isinstance(Fxm, FelDerivView) and Sm._upload()["F3"] is None
    N = Sc.n_dof
    rng = np.random.default_rng(5)
    x = torch.as_tensor(rng.standard_normal(N), device="cuda")
    for Fc, Fm in ((Fxc, Fxm), (Fyc, Fym), (Fzc, Fzm)):
        ref = Fc.to_scipy() @ x.cpu().numpy()
        got = (Fm @ x).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    u = torch.as_tensor(rng.standard_normal(3 * N), device="cuda")
    for R in (None, rot111_R()):
        ref = Dc.compute_E_from_u(N, u, R).cpu().numpy()
        got = Dm.compute_E_from_u(N, u, R).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    U = torch.as_tensor(rng.standard_normal(N), device="cuda")
    for hc, hm in zip(Dc.compute_Hd_from_U(U), Dm.compute_Hd_from_U(U)):
        hc, hm = hc.cpu().numpy(), hm.cpu().numpy()
        assert np.abs(hm - hc).max() <= 1e-13 * np.abs(hc).max()


@pytest.mark.parametrize("me,rot,mode,jitter", [(True, False, "fp64", 0.15), (True, True, "fp64", 0.15),
                                                (False, False, "fp64", 0.15), (True, True, "mixed", 0.15),
                                                (True, True, "fp64", 0.0)])
def test_steps_with_matrix_free_derivatives_match_oracle(me, rot, mode, jitter):
    """Strain, Htilde and the demag right-hand side from the element walk
    (no Fx/Fy/Fz CSR on the device): steps match the oracle (FP64: m 1e-10,
    u 1e-10, iterations +-1; mixed: m 1e-7)."""
    from cupymag_amd.core.gspm import GSPMSimulation
    from oracle import gspm as og
    cfg, arr, P, S = _ops(6, jitter, rot, True, me=me, deriv=True)
    assert S._upload()["F3"] is None
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    m0 = tilted_m(ops["n"], seed=10)
    sim = GSPMSimulation(S, P, mode=mode)
    sim.set_m(m0)
    st = og.init_state(ops, Avg, m0)
    p["H_now"] = p["H"]
    H = [P.Hext1, P.Hext2, P.Hext3]
    tol = 1e-10 if mode == "fp64" else 1e-7
    for k in range(3):
        log = []
        ui = st["use_init"]
        og.step(st, ops, Avg, p, log)
        rep = sim.step(H, ui)
        assert np.max(np.abs(sim.get("M").cpu().numpy() - st["m"])) < tol
        Ht = sim.get("HTILDE").cpu().numpy()
        assert np.abs(Ht - np.stack(st["Ht"])).max() <= tol * np.abs(np.stack(st["Ht"])).max()
        if me:
            u = sim.get("UEL").cpu().numpy()
            assert np.abs(u - st["u"]).max() <= tol * np.abs(st["u"]).max()
        if mode == "fp64":
            ref_it = [it for _, it in log]
            got = [it for it, s in zip(rep.iters, range(10)) if (me or s > 0)]
            assert all(abs(a - b) <= 1 for a, b in zip(got, ref_it)), (got, ref_it)
