"""Pins of the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests or golden vectors, so each oracle function is
checked here against something the mathematics fixes independently of the
restated formula (see the docstrings of oracle/*.py).
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from oracle import fem, gspm, mesh as om, params
from tests.cases import config, oracle_setup, tilted_m

MU0, GAMMA = 1.25663706127e-6, 1.76e11


def test_reduced_constants():
    cfg = config("Hex", me=True)
    p = params.reduced(cfg)
    ms = 1.2e6
    psi0 = MU0 * ms * ms
    assert p["psi0"] == pytest.approx(1.80956e6, rel=1e-5)
    assert p["K1"] == pytest.approx(-1e4 / psi0, rel=1e-14)
    assert p["astar"] == pytest.approx(2 * 1.5e-11 / psi0 / (2e-9) ** 2, rel=1e-14)
    assert p["astar"] == pytest.approx(4.14466, rel=1e-5)
    # dt = gamma mu0 Ms dt_SI / (1 + alpha^2): 1 ps, alpha 0.5
    assert p["dt"] == pytest.approx(GAMMA * MU0 * ms * 1e-12 / 1.25, rel=1e-14)
    assert p["dt"] == pytest.approx(0.212321, rel=1e-5)
    assert p["alpha1"] == pytest.approx(0.88, rel=1e-5)
    assert p["alpha2"] == pytest.approx(0.44, rel=1e-5)
    assert p["H"][0] == pytest.approx(5e4 / ms, rel=1e-14)
    c1, c2, c3 = p["elasC_raw"]
    assert (c1, c2, c3) == pytest.approx((-8.2263e-3, -21.5522, 11.9366), rel=1e-4)
    # 10 ps, alpha 0.1
    p3 = params.reduced(config("Tet", dt=1e-11, alpha=0.1))
    assert (p3["dt"], p3["alpha1"], p3["alpha2"]) == pytest.approx((2.62774, 10.891, 1.0891), rel=1e-4)
    # ME off zeroes the coupling constants (Micromagnetics.py:40-43)
    p0 = params.reduced(config("Hex", me=False))
    assert (p0["elasC1"], p0["elasC2"], p0["elasC3"]) == (0.0, 0.0, 0.0)


def test_reduced_constants_validation():
    cfg = config("Hex")
    cfg["physics"]["magnetic"]["ms"] = 0.0
    with pytest.raises(ValueError, match="Saturation magnetization must be positive"):
        params.reduced(cfg)
    cfg = config("Hex")
    cfg["simulation"]["stop_condition"] = {"component": 2, "value": 0.5}
    p = params.reduced(cfg)
    assert list(p["stop_direction"]) == [0.0, 0.0, 1.0] and p["stop_value"] == -0.95


def test_rotation_matrices():
    R, M = params.rot111_R(), params.rot111_M()
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15) and np.linalg.det(R) == pytest.approx(1.0)
    # R maps [111] onto the x axis (first row is [111]/sqrt3)
    assert np.allclose(R @ (np.ones(3) / np.sqrt(3)), [1, 0, 0])
    # M is the tensor-shear Voigt form, in the order (11,22,33,23,13,12), of
    # eps -> R eps R^T (M^-T the engineering-shear form).  The reference
    # applies it to vectors ordered (11,22,33,12,23,13); that is reproduced
    # as is.
    rng = np.random.default_rng(0)
    for _ in range(3):
        a = rng.standard_normal((3, 3))
        e = a + a.T
        v = lambda t, f: np.array([t[0, 0], t[1, 1], t[2, 2], f * t[1, 2], f * t[0, 2], f * t[0, 1]])
        assert np.allclose(M @ v(e, 1), v(R @ e @ R.T, 1), atol=1e-13)
        assert np.allclose(np.linalg.inv(M).T @ v(e, 2), v(R @ e @ R.T, 2), atol=1e-13)


def _tet_closed(x):
    """P1 tet: volume and barycentric gradients from the 4x4 vertex system."""
    A = np.hstack([np.ones((4, 1)), x])
    G = np.linalg.inv(A)[1:, :].T          # rows: grad N_a
    return abs(np.linalg.det(A)) / 6.0, G


def test_tet_element_matrices_closed_form():
    rng = np.random.default_rng(1)
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], dtype=float) + 0.1 * rng.standard_normal((4, 3))
    el = np.array([[0, 1, 2, 3, 0]], dtype=np.int32)
    K, M, F = fem.element_matrices(x, el)
    V, G = _tet_closed(x)
    assert np.allclose(K[0], V * G @ G.T, rtol=1e-12, atol=1e-14)
    a7, b7 = 0.5854102, 0.1381966
    diag, off = (a7 ** 2 + 3 * b7 ** 2) * V / 4, (2 * a7 * b7 + 2 * b7 ** 2) * V / 4
    Mref = np.full((4, 4), off) + np.eye(4) * (diag - off)
    assert np.allclose(M[0], Mref, rtol=1e-12)
    # the 7-digit table is within 1e-7 of the exact P1 mass V/20 (1 + delta)
    assert np.allclose(M[0], V / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=2e-7)
    # F_k[a, b] = integral N_a dN_b/dx_k = (V/4) G[b, k] (quadrature weights sum to V)
    for k in range(3):
        assert np.allclose(F[k, 0], np.outer(np.full(4, V / 4 * (a7 + 3 * b7)), G[:, k]), rtol=1e-12, atol=1e-14)


def test_hex_unit_cube_closed_form():
    coords, el, _ = om.hex_grid(1, 1, 1, 0, 0, 0)
    K, M, F = fem.element_matrices(coords, el)
    node = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    m1 = lambda a, b: 1 / 3 if a == b else 1 / 6
    k1 = lambda a, b: 1.0 if a == b else -1.0
    Mref = np.zeros((8, 8))
    Kref = np.zeros((8, 8))
    for a, b in itertools.product(range(8), range(8)):
        ms = [m1(node[a, d], node[b, d]) for d in range(3)]
        Mref[a, b] = np.prod(ms)
        Kref[a, b] = sum(k1(node[a, d], node[b, d]) * np.prod([ms[e] for e in range(3) if e != d]) for d in range(3))
    assert np.allclose(M[0], Mref, atol=1e-10)
    assert np.allclose(K[0], Kref, atol=1e-9)


def test_isotropic_elasticity_closed_form():
    lam, mu = 1.3, 0.7
    C = fem.voigt_C(lam + 2 * mu, lam, mu)
    rng = np.random.default_rng(2)
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], dtype=float) + 0.1 * rng.standard_normal((4, 3))
    el = np.array([[0, 1, 2, 3, 0]], dtype=np.int32)
    Ke, Fe = fem.elastic_element_matrices(x, el, C)
    V, G = _tet_closed(x)
    ref = np.zeros((12, 12))
    for a, b in itertools.product(range(4), range(4)):
        blk = lam * np.outer(G[a], G[b]) + mu * np.outer(G[b], G[a]) + mu * (G[a] @ G[b]) * np.eye(3)
        ref[3 * a:3 * a + 3, 3 * b:3 * b + 3] = V * blk
    assert np.allclose(Ke[0], ref, rtol=1e-11, atol=1e-13)
    # F_e applied to a uniform eigenstrain eps0 = the nodal forces of the
    # uniform stress C eps0: sum_b F_e[:, 6b:6b+6] eps0 = V B^T C eps0
    eps0 = rng.standard_normal(6)
    f = Fe[0].reshape(12, 4, 6).sum(axis=1) @ eps0
    sig = C @ eps0
    S = np.array([[sig[0], sig[3], sig[5]], [sig[3], sig[1], sig[4]], [sig[5], sig[4], sig[2]]])
    assert np.allclose(f.reshape(4, 3), V * G @ S * (0.5854102 + 3 * 0.1381966), rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("grid", ["Hex", "Tet"])
def test_assembled_operator_invariants(grid):
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = config(grid, n=4, me=True, ndef=0, defect_params=0.0)
    arr = kuhn_mesh(4, jitter=0.1, seed=1) if grid == "Tet" else None
    p, ops, Avg, (coords, el, gid) = oracle_setup(cfg, arr)
    n = ops["n"]
    one = np.ones(n)
    K = fem.to_csr(*fem.scalar_coo(el, gid, fem.element_matrices(coords, el)[0]), (n, n))
    assert np.abs(K @ one).max() < 1e-12            # periodic Laplacian kills constants
    for k in ("Fx", "Fy", "Fz"):
        assert np.abs(ops[k] @ one).max() < 1e-12   # derivative of a constant
    assert ops["F1"].sum() == pytest.approx(64.0, rel=1e-6)   # total volume 4^3 (7-digit table for tets)
    # rigid translations are in the kernel of the unanchored elastic operator
    C = fem.voigt_C(p["c11"], p["c12"], p["c44"])
    Ke, _ = fem.elastic_element_matrices(coords, el, C)
    npe = el.shape[1] - 1
    g = gid[el[:, :npe]].astype(np.int64)
    rowd = (g[:, :, None] + n * np.arange(3)[None, None, :]).reshape(len(g), -1)
    A = sp.coo_matrix((Ke.ravel(), (np.repeat(rowd[:, :, None], 3 * npe, 2).ravel(),
                                    np.repeat(rowd[:, None, :], 3 * npe, 1).ravel())), shape=(3 * n, 3 * n)).tocsr()
    for d in range(3):
        t = np.zeros(3 * n)
        t[d * n:(d + 1) * n] = 1.0
        assert np.abs(A @ t).max() < 1e-10 * abs(A).max()


def test_periodic_map_brute_force():
    for shape in ((2, 2, 2), (3, 2, 4)):
        coords, el, b = om.hex_grid(*shape, 0, 0, 0)
        gid = om.periodic_map(coords, b)
        L = np.array(shape, dtype=float)
        nn = len(coords)
        same = np.zeros((nn, nn), dtype=bool)
        for i in range(nn):
            d = np.mod(coords - coords[i], L[None, :])
            same[i] = np.all(np.isclose(d, 0) | np.isclose(d, L), axis=1)
        assert (same == (gid[:, None] == gid[None, :])).all()
        assert gid.max() + 1 == np.prod(shape)
        # ids are handed out in first-appearance order
        _, first = np.unique(gid, return_index=True)
        assert (np.diff(first) > 0).all()


def test_nas_round_trip(tmp_path):
    from cupymag_amd.mesh.kuhn import kuhn_mesh, write_nas
    coords, tets = kuhn_mesh(3, jitter=0.1, seed=5)
    path = tmp_path / "m.nas"
    write_nas(path, coords, tets)
    # shuffle the GRID card order and renumber ids: the reader must map ids
    lines = path.read_text().splitlines()
    grids = [l for l in lines if l.startswith("GRID")]
    rest = [l for l in lines if not l.startswith("GRID") and not l.startswith("ENDDATA")]
    perm = np.random.default_rng(0).permutation(len(grids))
    path.write_text("\n".join(rest[:2] + [grids[i] for i in perm] + rest[2:] + ["ENDDATA", "GRID,999,,0,0,0"]) + "\n")
    xyz, idx = om.read_nas(path)
    assert np.array_equal(xyz[idx], coords[tets])


def _spd(n, seed, eig=None):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = eig if eig is not None else rng.uniform(0.5, 20.0, n)
    return sp.csr_matrix(Q @ np.diag(d) @ Q.T), Q


def test_cg_against_direct_solve():
    A, _ = _spd(40, 0)
    b = np.random.default_rng(1).standard_normal(40)
    x, info, it = gspm.cg(A, b, tol=1e-12, maxiter=1000)
    assert info == 0 and it >= 1
    assert np.allclose(x, spsolve(A.tocsc(), b), rtol=1e-10)


def test_cg_iteration_counts():
    # k distinct eigenvalues -> CG terminates in k iterations
    eig = np.repeat([1.0, 3.0, 7.0], 10)
    A, Q = _spd(30, 2, eig)
    b = np.random.default_rng(3).standard_normal(30)
    _, info, it = gspm.cg(A, b, tol=1e-10, maxiter=100)
    assert info == 0 and it == 3
    # an eigenvector right-hand side converges in exactly one iteration
    _, _, it1 = gspm.cg(A, Q[:, 0], tol=1e-10)
    assert it1 == 1
    # zero rhs returns b without iterating; maxiter reached -> info = maxiter
    x0, info0, it0 = gspm.cg(A, np.zeros(30), tol=1e-10)
    assert it0 == 0 and info0 == 0 and not x0.any()
    A2, _ = _spd(30, 4)
    _, info2, it2 = gspm.cg(A2, b, tol=1e-14, maxiter=2)
    assert info2 == 2 and it2 == 2
    with pytest.raises(gspm.CGError, match="Error! CG for demag did not converge. info=2"):
        gspm.solve_cg(A2, b, tol=1e-14, maxiter=2, system="demag")


def test_preconditioned_cg_closed_forms():
    """cupyx cg with M (what solve_cg forwards, linear_solvers.py:59-62):
    M = I takes the unpreconditioned iterates exactly; M = A^-1 converges
    in one iteration; for M = (A restricted to k eigenvalue clusters)^-1
    ... PCG on M A with k distinct eigenvalues ends in k iterations; a
    Jacobi M gives the direct solve."""
    A, Q = _spd(40, 5)
    b = np.random.default_rng(6).standard_normal(40)
    x0_, _, it0 = gspm.cg(A, b, tol=1e-12, maxiter=200)
    x1, _, it1 = gspm.cg(A, b, tol=1e-12, maxiter=200, M=sp.identity(40, format="csr"))
    assert it0 == it1 and np.array_equal(x0_, x1)
    Ainv = np.linalg.inv(A.toarray())
    x2, info2, it2 = gspm.cg(A, b, tol=1e-10, maxiter=50, M=Ainv)
    assert info2 == 0 and it2 == 1 and np.allclose(x2, spsolve(A.tocsc(), b), rtol=1e-10)
    # M A has the eigenvalues 1, 2, 5 when M = Q diag(s_i / d_i) Q^T
    d = np.linalg.eigvalsh(A.toarray())
    Qd = np.linalg.eigh(A.toarray())[1]
    s_ = np.repeat([1.0, 2.0, 5.0], [10, 15, 15])
    Mk = Qd @ np.diag(s_ / d) @ Qd.T
    _, info3, it3 = gspm.cg(A, b, tol=1e-10, maxiter=50, M=Mk)
    assert info3 == 0 and it3 == 3
    J = sp.diags(1.0 / A.diagonal()).tocsr()
    x4, info4, _ = gspm.cg(A, b, tol=1e-12, maxiter=500, M=J)
    assert info4 == 0 and np.allclose(x4, spsolve(A.tocsc(), b), rtol=1e-9)


def test_build_E0_tensor_definition():
    m = tilted_m(5, 1, 60.0)
    l100, l111 = 2.6e-4, -3e-5
    E0 = gspm.build_E0(l100, l111, m).reshape(5, 6)
    for i in range(5):
        lam = np.full((3, 3), l111)
        np.fill_diagonal(lam, l100)
        e = 1.5 * lam * (np.outer(m[i], m[i]) - np.eye(3) / 3)
        assert np.allclose(E0[i], [e[0, 0], e[1, 1], e[2, 2], 2 * e[0, 1], 2 * e[1, 2], 2 * e[0, 2]], atol=1e-18)


def test_strain_of_affine_displacement():
    """On a non-periodic mesh (identity node->DOF map) the weak-form
    derivative of x_j is the lumped mass: E(u = G x) = sym(G) * integral N_a."""
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    coords, tets = kuhn_mesh(3, jitter=0.1, seed=2)
    el = np.concatenate([tets, np.zeros((len(tets), 1), np.int32)], axis=1)
    gid = np.arange(len(coords), dtype=np.int32)
    p = params.reduced(config("Tet"))
    ops = fem.build_operators(coords, el, gid, p)
    n = ops["n"]
    G = np.random.default_rng(3).standard_normal((3, 3))
    u = (coords @ G.T).T.ravel()          # u_j = sum_k G[j,k] x_k, stored [ux; uy; uz]
    lump = np.asarray(ops["F1"].sum(axis=1)).ravel() / (0.5854102 + 3 * 0.1381966)
    E = gspm.strain_from_u(ops, n, u)
    ref = [G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]]
    assert np.allclose(E, np.outer(lump, ref), rtol=1e-10, atol=1e-12)


def test_volume_average_invariants_and_brute_force():
    # (on a 2^3 grid every Jacobian taken from coords[gid] is degenerate and
    # the reference's volume clamp 1e-6 returns 0; 4^3 has a nonzero volume)
    coords, el, b = om.hex_grid(2, 2, 2, 0, 0, 0)
    gid = om.periodic_map(coords, b)
    assert fem.VolumeAverage(coords, el, gid).average(np.full(8, 0.37)) == 0.0
    coords, el, b = om.hex_grid(4, 4, 4, 0, 0, 0)
    gid = om.periodic_map(coords, b)
    Avg = fem.VolumeAverage(coords, el, gid)
    assert Avg.average(np.full(64, 0.37)) == pytest.approx(0.37, rel=1e-13)
    f = np.random.default_rng(0).standard_normal((64, 2))
    # dense re-evaluation with explicit loops, Jacobians from coords[gid[el]]
    pts, w = fem.hex_points()
    num = np.zeros(2)
    vol = 0.0
    for e in range(len(el)):
        xe = coords[gid[el[e, :8]]]
        for g, (r, s, t) in enumerate(pts):
            J = xe.T @ fem.hex_dN(r, s, t)
            dv = abs(np.linalg.det(J)) * w[g]
            num += dv * (fem.hex_N(r, s, t) @ f[gid[el[e, :8]]])
            vol += dv
    assert np.allclose(Avg.average(f), num / vol, rtol=1e-12)


def _macrospin_step(p, m, E):
    """One GSPM step of a uniform magnetization from the energy density: for
    uniform m the FEM solves reduce to (M + a K)^-1 M c = c and Htilde = 0,
    so g = f dt + m with f = -grad_m e(m) + Hbar/2 + H, where e is the cubic
    anisotropy + magnetoelastic energy density (gradient by central
    differences of e, not by the closed-form expression)."""
    K1, c1, c2, c3, l100 = p["K1"], p["elasC1"], p["elasC2"], p["elasC3"], p["lambda100"]
    R = params.rot111_R() if p["rot111"] else np.eye(3)

    def energy(mv):
        a = mv @ R
        s = a * a
        return (K1 * (s[0] * s[1] + s[1] * s[2] + s[2] * s[0]) + c1 * (s[0] * s[1] + s[1] * s[2] + s[2] * s[0])
                + c2 * sum((E[i] - l100) * s[i] for i in range(3))
                + c3 * (E[3] * a[0] * a[1] + E[4] * a[1] * a[2] + E[5] * a[0] * a[2]))

    def field(mv, Hbar):
        h = 1e-6
        grad = np.array([(energy(mv + h * np.eye(3)[i]) - energy(mv - h * np.eye(3)[i])) / (2 * h)
                         for i in range(3)])
        return -grad + 0.5 * Hbar + np.asarray(p["H_now"])

    Hbar = -np.asarray(p["N"]) * m
    dt = p["dt"]
    f = field(m, Hbar)
    g = f * dt + m
    m1s = m[0] + g[1] * m[2] - g[2] * m[1]
    g1s = f[0] * dt + m1s
    m2s = m[1] + g[2] * m1s - g1s * m[2]
    g2s = f[1] * dt + m2s
    m3s = m[2] + g1s * m2s - g2s * m1s
    ms_ = np.array([m1s, m2s, m3s])
    mss = field(ms_, Hbar) * dt * p["alpha_damp"] + ms_
    return mss / np.linalg.norm(mss)


@pytest.mark.parametrize("me,rot", [(False, False), (True, False), (True, True), (False, True)])
def test_step_macrospin_reduction(me, rot):
    cfg = config("Hex", n=3, me=me, rot=rot, ndef=0, hext=(3e4, -2e4, 1e4), K1=4e4, sigma11=5e7)
    p, ops, Avg, _ = oracle_setup(cfg)
    p["H_now"] = p["H"]
    m = np.array([0.6, 0.64, 0.48])
    st = gspm.init_state(ops, Avg, np.tile(m, (ops["n"], 1)))
    for _ in range(3):
        am = st["avg_m"]
        if me:
            # uniform eigenstrain: divergence-free stress, u = 0, E = Ebar (in the crystal frame)
            a = am @ params.rot111_R() if rot else am
            Es = p["Ebar_sigma"]
            Eb = np.array([1.5 * p["lambda100"] * (a[i] ** 2 - 1 / 3) + Es[i] for i in range(3)]
                          + [3 * p["lambda111"] * a[0] * a[1] + Es[3], 3 * p["lambda111"] * a[1] * a[2] + Es[4],
                             3 * p["lambda111"] * a[0] * a[2] + Es[5]])
            if rot:
                Mv = params.rot111_M()
                Eb = np.linalg.inv(Mv).T.T @ (Mv @ Eb)  # (M Eb) @ M^-T as a row vector
            E = Eb
        else:
            E = np.zeros(6)
        ref = _macrospin_step(p, am, E)
        gspm.step(st, ops, Avg, p)
        assert np.allclose(st["m"], ref[None, :], atol=2e-8)
        m_unit = np.linalg.norm(st["m"], axis=1)
        assert np.allclose(m_unit, 1.0, atol=1e-14)


def test_step_invariants_with_defect():
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = config("Tet", n=5, me=True, rot=True, defect_params=1.6, dt=5e-12)
    arr = kuhn_mesh(5, jitter=0.15, seed=3)
    p, ops, Avg, _ = oracle_setup(cfg, arr)
    p["H_now"] = p["H"]
    st = gspm.init_state(ops, Avg, tilted_m(ops["n"], 3, 20.0))
    dd = ops["DefDOF"]
    assert len(dd) > 0
    for _ in range(2):
        gspm.step(st, ops, Avg, p)
        assert not st["m"][dd].any()
        off = np.setdiff1d(np.arange(ops["n"]), dd)
        assert np.allclose(np.linalg.norm(st["m"][off], axis=1), 1.0, atol=1e-14)


def _random_rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_rotated_ellipsoid_axes_are_the_rotation_columns():
    """RotatedEllipsoid (defect_shapes.py:113-134): the semi-axis k lies
    along column k of the rotation (x^T R diag(1/a^2) R^T x < 1).  Points at
    0.999 and 1.001 of each semi-axis along +-R[:, k] fall inside and
    outside; with the transposed rotation (axes along the rows) they do not,
    so the check separates R from R^T.  Both the oracle and the product's
    shape function are checked."""
    from cupymag_amd.utils.defect_shapes import get_defect_shape_function
    abc = np.array([5.0, 3.0, 2.0])
    for seed in (0, 1, 2):
        R = _random_rotation(seed)
        assert np.abs(R - R.T).max() > 0.1
        fns = [om.rotated_ellipsoid(abc, R), get_defect_shape_function("RotatedEllipsoid", (abc, R))]
        pts, want = [], []
        for k in range(3):
            for sgn in (1.0, -1.0):
                pts += [sgn * 0.999 * abc[k] * R[:, k], sgn * 1.001 * abc[k] * R[:, k]]
                want += [True, False]
        pts = np.array(pts)
        for fn in fns:
            assert list(fn(pts)) == want
        assert all(bool(fns[1](p)) == w for p, w in zip(pts, want))     # the product's one-point form
        assert list(om.rotated_ellipsoid(abc, R.T)(pts)) != want


def test_strain_of_affine_displacement_rotated_frame():
    """compute_E_from_u with R (compute_derivatives.py:94-123): for
    u = G x the weak-form derivatives are D_i u_j = G[j, i] * lump, so the
    R-mixed strain is the Voigt form of G R (engineering shears:
    E12 = (GR)01 + (GR)10 ...) times the lumped mass.  With R^T in place of
    R the result is the Voigt form of G R^T, which differs for the [111]
    rotation and for a random one."""
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    coords, tets = kuhn_mesh(3, jitter=0.1, seed=2)
    el = np.concatenate([tets, np.zeros((len(tets), 1), np.int32)], axis=1)
    gid = np.arange(len(coords), dtype=np.int32)
    p = params.reduced(config("Tet"))
    ops = fem.build_operators(coords, el, gid, p)
    n = ops["n"]
    G = np.random.default_rng(3).standard_normal((3, 3))
    u = (coords @ G.T).T.ravel()
    lump = np.asarray(ops["F1"].sum(axis=1)).ravel() / (0.5854102 + 3 * 0.1381966)
    for R in (params.rot111_R(), _random_rotation(7)):
        E = gspm.strain_from_u(ops, n, u, R=R)
        GR = G @ R
        ref = [GR[0, 0], GR[1, 1], GR[2, 2], GR[0, 1] + GR[1, 0], GR[1, 2] + GR[2, 1], GR[0, 2] + GR[2, 0]]
        assert np.allclose(E, np.outer(lump, ref), rtol=1e-10, atol=1e-12)
        GRt = G @ R.T
        wrong = [GRt[0, 0], GRt[1, 1], GRt[2, 2], GRt[0, 1] + GRt[1, 0], GRt[1, 2] + GRt[2, 1], GRt[0, 2] + GRt[2, 0]]
        assert not np.allclose(E, np.outer(lump, wrong), rtol=1e-6, atol=1e-9)


def _dense_scatter(Ke, blocks):
    """Dense global matrix from element matrices: entry (a, b) of element e
    adds to (blocks[e][a], blocks[e][b])."""
    N = int(max(b.max() for b in blocks)) + 1
    D = np.zeros((N, N))
    for e in range(Ke.shape[0]):
        idx = blocks[e]
        for a in range(len(idx)):
            for b in range(len(idx)):
                D[idx[a], idx[b]] += Ke[e, a, b]
    return D


@pytest.mark.parametrize("grid", ["Hex", "Tet"])
def test_operator_surgery_dense(grid):
    """Anchors and defect surgery, densely on a small periodic mesh with a
    defect (final_assembly.py:59-158, assemble_demag.py:307-341,
    assemble_elasticity.py:461-527): every operator equals its element
    matrices scattered through the DOF map, then A_demag has row/column 0
    replaced by the identity, A_el rows/columns 0, n, 2n, A1 and A2 the
    rows/columns of the defect DOFs; F1 loses the defect rows and F_el the
    anchor rows; Fx, Fy, Fz are untouched.  The identity diagonals must be
    stored (CG never reads decoupled rows, so a missing one would go
    unnoticed in a solve)."""
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    cfg = config(grid, n=4, me=True, ndef=2, defect_params=0.9)
    if grid == "Hex":
        p, ops, _, (coords, el, gid) = oracle_setup(cfg)
    else:
        p, ops, _, (coords, el, gid) = oracle_setup(cfg, kuhn_mesh(3, jitter=0.1, seed=4))
    n = ops["n"]
    npe = el.shape[1] - 1
    g = gid[el[:, :npe]].astype(np.int64)
    dd = ops["DefDOF"]
    # defect DOFs: every DOF of a node of a flagged element
    want_dd = sorted({int(gid[v]) for e in range(len(el)) if el[e, -1] == 1 for v in el[e, :npe]})
    assert list(dd) == want_dd and len(dd) > 0
    K, M, F = fem.element_matrices(coords, el)
    Kd, Md = _dense_scatter(K, g), _dense_scatter(M, g)

    def identity_on(D, rows):
        D = D.copy()
        D[rows, :] = 0.0
        D[:, rows] = 0.0
        D[rows, rows] = 1.0
        return D

    def check(A, want):
        A = A.tocsr()
        assert np.abs(A.toarray() - want).max() <= 1e-13 * np.abs(want).max()
        return A

    A = check(ops["A_demag"], identity_on(Kd, [0]))
    assert A[0, 0] == 1.0 and A.indptr[1] - A.indptr[0] >= 1
    for name, a in (("A1", p["alpha1"]), ("A2", p["alpha2"])):
        A = check(ops[name], identity_on(Md + a * Kd, dd))
        assert all(A[d, d] == 1.0 for d in dd)
    F1 = Md.copy()
    F1[dd, :] = 0.0
    check(ops["F1"], F1)
    for k, name in enumerate(("Fx", "Fy", "Fz")):
        check(ops[name], _dense_scatter(F[k], g))
    C = fem.voigt_C(p["c11"], p["c12"], p["c44"], p["rot111"], params.rot111_M())
    Ke, Fe = fem.elastic_element_matrices(coords, el, C)
    # element DOF (3 a + d) is the global DOF g_a + d n (assemble_elasticity.py:389-398)
    blk = np.stack([g[:, a] + d * n for a in range(npe) for d in range(3)], axis=1)
    perm = np.array([3 * a + d for a in range(npe) for d in range(3)])
    Kel = _dense_scatter(Ke[:, perm][:, :, perm], blk)
    A = check(ops["A_el"], identity_on(Kel, [0, n, 2 * n]))
    assert all(A[r, r] == 1.0 for r in (0, n, 2 * n))
    Fw = np.zeros((3 * n, 6 * n))
    for e in range(len(el)):
        for a in range(npe):
            for d in range(3):
                for b in range(npe):
                    Fw[g[e, a] + d * n, 6 * g[e, b]:6 * g[e, b] + 6] += Fe[e, 3 * a + d, 6 * b:6 * b + 6]
    Fw[[0, n, 2 * n], :] = 0.0
    check(ops["F_el"], Fw)


def test_hex_defect_box_closed_form():
    """gridHex.py:99-139: a cell is defective iff its centre lies in the
    half-open box [c - Nd/2, c + Nd/2); extract_defect_dofs
    (final_assembly.py:27-57) gives the DOFs of those cells' nodes.
    32^3 with a centred 4^3 box: cells 14..17 per axis (64 cells) and nodes
    14..18 (125 DOFs).  A box whose faces pass through cell centres takes
    the lower face's cells and not the upper's; a box at the periodic face
    shares DOFs with the opposite face."""
    coords, el, b = om.hex_grid(32, 32, 32, 4, 4, 4)
    gid = om.periodic_map(coords, b)
    assert int(el[:, -1].sum()) == 64
    assert len(om.defect_dofs(el, gid)) == 125
    cells = np.argwhere(el[:, -1].reshape(32, 32, 32) == 1)
    assert cells.min() == 14 and cells.max() == 17
    # faces through cell centres: [2.5, 5.5) holds centres 2.5, 3.5, 4.5
    coords, el, b = om.hex_grid(8, 8, 8, 3, 3, 3, defect_center=(4.0, 4.0, 4.0))
    assert int(el[:, -1].sum()) == 27
    cells = np.argwhere(el[:, -1].reshape(8, 8, 8) == 1)
    assert cells.min() == 2 and cells.max() == 4
    # a box reaching the upper face x = 8: cells 6, 7 in x; nodes x = 6, 7, 8 and 8 wraps to 0
    coords, el, b = om.hex_grid(8, 8, 8, 3, 3, 3, defect_center=(8.0, 4.0, 4.0))
    gid = om.periodic_map(coords, b)
    assert int(el[:, -1].sum()) == 2 * 3 * 3
    dd = om.defect_dofs(el, gid)
    assert len(dd) == 3 * 4 * 4
    x_of_dof = {int(gid[i]): coords[i, 0] for i in range(len(coords)) if coords[i, 0] < 8}
    assert sorted({x_of_dof[int(d)] for d in dd}) == [0.0, 6.0, 7.0]
