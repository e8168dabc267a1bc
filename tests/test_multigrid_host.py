"""Host side of the multigrid hierarchy: Kuhn-mesh detection, level sizes,
and the transfer operator pinned by the Galerkin identity of nested P1
spaces: for an undistorted Kuhn mesh, P^T K_fine P equals the operator
assembled on the coarse mesh (scalar stiffness and elasticity, operators
from the oracle's element matrices), which fails for a wrong stencil,
weight or wrap-around; the non-nested transfers against a brute-force
search of the coarse tetrahedra; and the periodic stiffness of the floating
levels (the host assembly without anchors against the oracle's stiffness,
its recovery from the anchored operator)."""
import numpy as np
import scipy.sparse as sp

from cupymag_amd.mesh.coarsen import coarse_kuhn, hierarchy_sizes, kuhn_cells, prolongation
from cupymag_amd.mesh.kuhn import kuhn_mesh
from cupymag_amd.mesh.periodic import structured_dims
from oracle import fem
from oracle import mesh as om


def _gid(coords):
    return om.periodic_map(coords, om.tet_bounds(coords))


def _elements(tets):
    el = np.zeros((tets.shape[0], 5), dtype=np.int32)
    el[:, :4] = tets
    return el


def _K(coords, tets):
    el = _elements(tets)
    gid = _gid(coords)
    n = int(gid.max()) + 1
    assert structured_dims(coords, gid) is not None
    K, _, _ = fem.element_matrices(coords, el, want=("K",))
    r, c, v = fem.scalar_coo(el, gid, K)
    return sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()


def _Kel(coords, tets, Cv):
    """elasticity stiffness, node-major interleaved (3i + d)"""
    el = _elements(tets)
    gid = _gid(coords)
    n = int(gid.max()) + 1
    Ke, _ = fem.elastic_element_matrices(coords, el, Cv)
    g = gid[tets]                                   # (ne, 4)
    # element DOF 3a + d (the oracle's B matrix order) -> global 3 gid_a + d
    dof = (3 * g[:, :, None] + np.arange(3)[None, None, :]).reshape(len(tets), 12)
    Kn = Ke
    rows = np.repeat(dof[:, :, None], 12, axis=2).ravel()
    cols = np.repeat(dof[:, None, :], 12, axis=1).ravel()
    return sp.coo_matrix((Kn.ravel(), (rows, cols)), shape=(3 * n, 3 * n)).tocsr()


def test_kuhn_detection_and_sizes():
    coords, tets = kuhn_mesh(8)
    assert kuhn_cells(coords, _elements(tets)) == 8
    cj, tj = kuhn_mesh(8, jitter=0.15, seed=1)
    assert kuhn_cells(cj, _elements(tj)) == 8
    perm = np.random.default_rng(0).permutation(len(coords))
    assert kuhn_cells(coords[perm], _elements(tets)) is None
    assert hierarchy_sizes(128) == [128, 64, 32, 16, 8]
    assert hierarchy_sizes(96) == [96, 48, 24, 12, 6]
    assert hierarchy_sizes(90) == [90, 45, 23, 12, 6]
    assert hierarchy_sizes(87) == [87, 44, 22, 11, 6]
    assert hierarchy_sizes(9) == [9, 5] and hierarchy_sizes(8) == [8]
    # scalar hierarchies with their dense level at <= 16 cells (two levels at least)
    from cupymag_amd.solvers.multigrid import scalar_levels
    assert scalar_levels(hierarchy_sizes(128)) == [128, 64, 32, 16, 8]
    assert scalar_levels(hierarchy_sizes(128), 16) == [128, 64, 32, 16]
    assert scalar_levels(hierarchy_sizes(96), 16) == [96, 48, 24, 12]
    assert scalar_levels(hierarchy_sizes(16), 16) == [16, 8]
    assert scalar_levels(hierarchy_sizes(18), 16) == [18, 9]
    assert scalar_levels(hierarchy_sizes(9), 16) == [9, 5]
    C, T = coarse_kuhn(coords, 8)
    assert np.array_equal(T, kuhn_mesh(4)[1])
    np.testing.assert_array_equal(C, kuhn_mesh(4, h=2.0)[0])
    # odd: the regular mesh of (n+1)/2 cells over the same box
    c9, _ = kuhn_mesh(9, h=1.5, jitter=0.15, seed=4)
    C, T = coarse_kuhn(c9, 9)
    assert np.array_equal(T, kuhn_mesh(5)[1])
    np.testing.assert_allclose(C, kuhn_mesh(5, h=1.5 * 9 / 5)[0], rtol=0, atol=1e-12)


def test_prolongation_basics():
    P = prolongation((8, 8, 8), anchor=False)
    np.testing.assert_allclose(P @ np.ones(64), 1.0)
    Pa = prolongation((8, 8, 8), anchor=True)
    assert Pa[0].nnz == 0 and Pa[:, 0].nnz == 0
    assert (P - Pa)[1:, 1:].nnz == 0


def test_galerkin_identity_scalar():
    cf, tf = kuhn_mesh(8)
    cc, tc = coarse_kuhn(cf, 8)
    Kf, Kc = _K(cf, tf), _K(cc, tc)
    P = prolongation((8, 8, 8), anchor=False)
    G = (P.T @ Kf @ P).toarray()
    assert np.abs(G - Kc.toarray()).max() < 1e-12 * np.abs(Kc.toarray()).max()
    # the identity is specific: a jittered fine mesh breaks it
    cj, _ = kuhn_mesh(8, jitter=0.15, seed=2)
    Gj = (P.T @ _K(cj, tf) @ P).toarray()
    assert np.abs(Gj - _K(*coarse_kuhn(cj, 8)).toarray()).max() > 1e-6


def test_galerkin_identity_elasticity():
    Cv = fem.voigt_C(2.4e11, 1.4e11, 1.2e11) / 1e11
    cf, tf = kuhn_mesh(8)
    cc, tc = coarse_kuhn(cf, 8)
    Kf, Kc = _Kel(cf, tf, Cv), _Kel(cc, tc, Cv)
    P3 = prolongation((8, 8, 8), bs=3, anchor=False)
    G = (P3.T @ Kf @ P3).toarray()
    assert np.abs(G - Kc.toarray()).max() < 1e-12 * np.abs(Kc.toarray()).max()


def _kuhn_eval(uc, nc, x):
    """Brute force: the P1 function of the periodic coarse Kuhn mesh with
    nodal values uc (nc cells per dimension, kuhn_mesh tetrahedra) at point
    x (coarse cell units): find a tetrahedron of the containing cube whose
    barycentric coordinates are all >= 0 and interpolate."""
    nc = np.asarray(nc)
    cell = np.minimum(np.floor(x).astype(int), nc - 1)
    grid, tets = kuhn_mesh(1)
    for t in tets:
        V = grid[t]                                  # 4 x 3 local vertex offsets
        A = np.vstack([V.T, np.ones(4)])
        lam = np.linalg.solve(A, np.append(x - cell, 1.0))
        if lam.min() >= -1e-12:
            val = 0.0
            for a in range(4):
                v = (cell + V[a].astype(int)) % nc
                val += lam[a] * uc[(v[0] * nc[1] + v[1]) * nc[2] + v[2]]
            return val
    raise AssertionError("point in no tetrahedron")


def test_non_nested_prolongation_interpolates_coarse_p1_functions():
    """P of a non-nested pair (9 -> 5, and a box of 9 x 7 x 5 -> 5 x 4 x 3
    cells) applied to a random coarse nodal vector gives the coarse P1
    function at every fine node's nominal position, found by brute-force
    search of the coarse Kuhn tetrahedra; for even sizes the same code is the
    nested P (pinned by the Galerkin identities above)."""
    rng = np.random.default_rng(8)
    for nf in ((9, 9, 9), (9, 7, 5)):
        nf = np.asarray(nf)
        nc = (nf + 1) // 2
        P = prolongation(tuple(nf), anchor=False)
        uc = rng.standard_normal(int(np.prod(nc)))
        uf = P @ uc
        for f in rng.choice(int(np.prod(nf)), size=150, replace=False):
            i, j, k = f // (nf[1] * nf[2]), (f // nf[2]) % nf[1], f % nf[2]
            x = np.array([i, j, k]) * nc / nf
            assert abs(uf[f] - _kuhn_eval(uc, nc, x)) < 1e-12
        np.testing.assert_allclose(P @ np.ones(len(uc)), 1.0, rtol=0, atol=1e-14)
        Pa = prolongation(tuple(nf), anchor=True)
        assert Pa[0].nnz == 0 and Pa[:, 0].nnz == 0 and (P - Pa)[1:, 1:].nnz == 0


def test_unanchored_assembly_and_anchor_recovery():
    """The floating levels' operators: the host assembler without anchors
    gives the oracle's periodic stiffness (scalar) and the anchored
    operators differ from it only in node 0's rows and columns; recovering
    K from the anchored operator (tests.cases.unanchor, the rule the device
    applies to level 0) gives it back, and K annihilates the translations."""
    from cupymag_amd.mesh.periodic import periodic_node_map
    from cupymag_amd.physics.host_ops import HostOps, voigt_C
    from tests.cases import unanchor
    coords, tets = kuhn_mesh(6, jitter=0.15, seed=2)
    el = _elements(tets)
    gid = periodic_node_map(coords, coords.min(axis=0), coords.max(axis=0))
    n = int(gid.max()) + 1
    Cv = voigt_C(2.4e11, 1.4e11, 1.2e11, True) / 1e11
    ops = {}
    for anchors in (True, False):
        H = HostOps(coords, el, gid, n, None, 2.0, 1.0, Cv, 2, anchors=anchors)
        ops[anchors] = {k: H.scipy(k) for k in ("A_demag", "A_el", "A1")}
        H.free()
    K = _K(coords, tets)
    Ku = ops[False]["A_demag"]
    assert abs(Ku - K).max() <= 1e-12 * abs(K).max()
    np.testing.assert_array_equal(ops[True]["A1"].toarray(), ops[False]["A1"].toarray())
    for name, bs in (("A_demag", 1), ("A_el", 3)):
        perm = (np.arange(bs)[None, :] * n + np.arange(n)[:, None]).ravel()   # node-major interleaved
        As = ops[True][name][perm][:, perm].tocsr()   # the pattern keeps the anchor's explicit zeros
        Aa = As.toarray()
        Au = ops[False][name][perm][:, perm].toarray()
        keep = np.arange(bs, bs * n)
        np.testing.assert_array_equal(Aa[np.ix_(keep, keep)], Au[np.ix_(keep, keep)])
        np.testing.assert_array_equal(Aa[:bs, :bs], np.eye(bs))
        assert np.all(Aa[:bs, bs:] == 0.0) and np.all(Aa[bs:, :bs] == 0.0)
        Kr = unanchor(As, bs).toarray()
        assert np.abs(Kr - Au).max() <= 1e-12 * np.abs(Au).max(), name
        for d in range(bs):
            t = np.zeros(bs * n)
            t[d::bs] = 1.0
            assert np.abs(Au @ t).max() <= 1e-12 * np.abs(Au).max()


def test_floating_preconditioner_is_the_anchored_inverse():
    """The floating hierarchy's preconditioner z = F^T V(F r) + e_0 r_0,
    F = I - e_0 1^T per component, is the inverse of the anchored operator
    when V is exact (V = K^+, the pseudo-inverse of the periodic stiffness,
    as the coarsest solve computes it): checked densely on a jittered 4^3
    mesh for the scalar and the elasticity operator, with the pseudo-inverse
    from solvers.multigrid._pinv_periodic (pinned against numpy's pinv)."""
    from cupymag_amd.mesh.periodic import periodic_node_map
    from cupymag_amd.physics.host_ops import HostOps, voigt_C
    from cupymag_amd.solvers.multigrid import _interleave_perm, _pinv_periodic
    coords, tets = kuhn_mesh(4, jitter=0.1, seed=5)
    el = _elements(tets)
    gid = periodic_node_map(coords, coords.min(axis=0), coords.max(axis=0))
    n = int(gid.max()) + 1
    Cv = voigt_C(2.4e11, 1.4e11, 1.2e11, True) / 1e11
    ops = {}
    for anchors in (True, False):
        H = HostOps(coords, el, gid, n, None, 2.0, 1.0, Cv, 2, anchors=anchors)
        ops[anchors] = {k: H.scipy(k).toarray() for k in ("A_demag", "A_el")}
        H.free()
    for name, bs in (("A_demag", 1), ("A_el", 3)):
        perm = _interleave_perm(n, bs)
        A = ops[True][name][np.ix_(perm, perm)]
        K = ops[False][name][np.ix_(perm, perm)]
        Kp = _pinv_periodic(K, bs)
        assert np.abs(Kp - np.linalg.pinv(K)).max() <= 1e-9 * np.abs(Kp).max()
        N = bs * n
        F = np.eye(N)
        for d in range(bs):
            F[d, d::bs] -= 1.0          # row of node 0, component d: r_0 - sum of that component
        E0 = np.zeros((N, N))
        E0[:bs, :bs] = np.eye(bs)
        P = F.T @ Kp @ F + E0
        Ainv = np.linalg.inv(A)
        assert np.abs(P - Ainv).max() <= 1e-8 * np.abs(Ainv).max(), name
        # and the plain shift without F (the V-cycle on r as it is) is not
        assert np.abs(F.T @ Kp + E0 - Ainv).max() > 1e-3 * np.abs(Ainv).max()


def _transfer_host(tmp_path_factory):
    """tests/transfer_host.hip built for the host (the device transfer rules
    of mg_dev.h, run on the CPU), or None without hipcc."""
    import os
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        return None
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path_factory.mktemp("transfer") / "transfer_host")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-std=c++17", f"-I{root}/cupymag_amd/csrc",
                    f"-I{root}/include", os.path.join(root, "tests", "transfer_host.hip"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def test_device_transfer_rules_equal_the_host_prolongation(tmp_path_factory):
    """The per-node restriction and prolongation of mg_dev.h (compiled for
    the host) equal P^T r and P e with the host P (which the Galerkin and
    brute-force tests above pin): nested and non-nested size pairs, a box,
    with and without the anchor node."""
    import subprocess
    import pytest
    exe = _transfer_host(tmp_path_factory)
    if exe is None:
        pytest.skip("hipcc not available")
    rng = np.random.default_rng(17)
    for nf3, nc3 in (((16, 16, 16), (8, 8, 8)), ((9, 9, 9), (5, 5, 5)), ((13, 13, 13), (7, 7, 7)),
                     ((17, 17, 17), (9, 9, 9)), ((9, 7, 5), (5, 4, 3)), ((10, 10, 10), (5, 5, 5))):
        for anchor in (False, True):
            P = prolongation(nf3, anchor=anchor, dims_c=nc3)
            nf, nc = P.shape
            r = rng.standard_normal(nf)
            e = rng.standard_normal(nc)
            out = subprocess.run([exe, *map(str, nf3), *map(str, nc3), str(int(anchor))],
                                 input=r.tobytes() + e.tobytes(), capture_output=True, check=True).stdout
            got = np.frombuffer(out, dtype=np.float64)
            Rr, Pe = got[:nc], got[nc:]
            np.testing.assert_allclose(Rr, P.T @ r, rtol=0, atol=1e-13 * np.abs(P.T @ r).max())
            np.testing.assert_allclose(Pe, P @ e, rtol=0, atol=1e-13 * np.abs(P @ e).max())


def test_half_stored_kuhn_stencil_equals_the_full_form(tmp_path_factory):
    """FP16 level forms and FP32 operator forms of a symmetric Kuhn stencil: the half layout
    (make_half_row of mg_dev.h: forward slots only, backward blocks read
    transposed from the neighbour rows) gives the full layout's product bit
    for bit, its compensated diagonal blocks agree to rounding, and both are
    the FP64 product to FP16 precision (random slot-symmetric operators,
    1x1 and 3x3 blocks, grids with padding rows and the smallest dims)."""
    import os
    import shutil
    import subprocess
    import pytest
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("hipcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path_factory.mktemp("kuhn_half") / "kuhn_half_host")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-std=c++17", f"-I{root}/cupymag_amd/csrc",
                    f"-I{root}/include", os.path.join(root, "tests", "kuhn_half_host.hip"), "-o", exe],
                   check=True, capture_output=True)
    for dims, bs in (((3, 3, 3), 3), ((5, 6, 7), 3), ((8, 8, 8), 1), ((4, 5, 64), 3), ((9, 3, 11), 1)):
        out = subprocess.run([exe, *map(str, dims), str(bs), "7"], capture_output=True, check=True, text=True).stdout
        l16, l32 = out.strip().split("\n")
        dy, ym, dd, dm, e64, y64 = map(float, l16.split())
        assert dy == 0.0 and ym > 0.0, (dims, bs, dy)
        assert dd <= 1e-14 * dm, (dims, bs, dd, dm)
        assert 0.0 < e64 <= 1e-3 * y64, (dims, bs, e64, y64)
        # FP32 forms (Sell::half32, kuhn_half32 of sell_dev.h): the same, and
        # the symmetry check tells a perturbed operator apart
        dy, ym, dd, dm, sym, sym2 = map(float, l32.split())
        assert dy == 0.0 and ym > 0.0, (dims, bs, dy)
        assert dd <= 1e-14 * dm, (dims, bs, dd, dm)
        assert sym == 1.0 and sym2 == 0.0
