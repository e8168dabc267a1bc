"""CPU tests of the product's host side: native assembly against the oracle,
periodic map, Kuhn meshes, parameters, and the C-ABI exports."""
import re
from pathlib import Path

import numpy as np
import pytest

from tests.cases import config, oracle_setup

ROOT = Path(__file__).resolve().parent.parent


def _lib_or_skip():
    from cupymag_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libcupymag_hip.so not built")
    return _lib.lib()


def test_library_exports_every_declared_symbol():
    L = _lib_or_skip()
    header = (ROOT / "include" / "cupymag_hip.h").read_text()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(cpm_\w+)\s*\(", header, re.M))
    assert len(names) > 20
    for name in sorted(names):
        assert hasattr(L, name), name
    assert L.cpm_abi_version() == int(re.search(r"CPM_ABI_VERSION (\d+)", header).group(1))


@pytest.mark.parametrize("grid,rot", [("Hex", False), ("Tet", True)])
def test_native_assembly_matches_oracle(grid, rot):
    _lib_or_skip()
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.physics.host_ops import HostOps, voigt_C
    arr = kuhn_mesh(5, jitter=0.15, seed=1) if grid == "Tet" else None
    cfg = config(grid, n=6, me=True, rot=rot, defect_params=1.6)
    p, ops, Avg, (coords, el, gid) = oracle_setup(cfg, arr)
    H = HostOps(coords, el, gid, ops["n"], ops["DefDOF"], p["alpha1"], p["alpha2"],
                voigt_C(p["c11"], p["c12"], p["c44"], p["rot111"]))
    for name in ("A_demag", "Fx", "Fy", "Fz", "A1", "A2", "F1", "A_el", "F_el"):
        A, B = ops[name], H.scipy(name)
        A.sort_indices()
        assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices), name
        assert abs(A - B).max() <= 2e-15 * abs(A).max(), name
    H.free()


def test_vectorised_periodic_map_matches_reference_loop():
    from cupymag_amd.mesh.gridHex import HexGrid
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.mesh.periodic import periodic_node_map
    from oracle import mesh as om
    coords, tets = kuhn_mesh(4, jitter=0.2, seed=9)
    ref = om.periodic_map(coords, om.tet_bounds(coords))
    assert np.array_equal(periodic_node_map(coords, coords.min(0), coords.max(0)), ref)
    g = HexGrid(3, 4, 5, 1, 1, 1)
    c, el = g.std_fem_mesh()
    rc, rel_, b = om.hex_grid(3, 4, 5, 1, 1, 1)
    assert np.array_equal(c, rc) and np.array_equal(el, rel_)
    assert np.array_equal(g.build_periodic_node_map(c), om.periodic_map(rc, b))


def test_kuhn_mesh_orientation_and_dof_numbering():
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.mesh.periodic import periodic_node_map
    n = 6
    coords, tets = kuhn_mesh(n, jitter=0.15, seed=4)
    x = coords[tets]
    det = np.linalg.det(x[:, 1:] - x[:, :1])
    assert (det > 0).all()
    assert np.allclose(det.sum() / 6.0, n ** 3)
    gid = periodic_node_map(coords, coords.min(0), coords.max(0))
    i, j, k = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    expect = ((i % n) * n * n + (j % n) * n + (k % n)).ravel()
    assert np.array_equal(gid, expect)


def test_parameters_match_oracle():
    from cupymag_amd.core.parameters import from_dict
    from oracle import params
    cfg = config("Tet", me=True, rot=True, sigma11=3e7, dt=5e-12)
    P = from_dict(cfg)
    p = params.reduced(cfg)
    for a, b in (("dt", "dt"), ("alpha", "alpha1"), ("K1", "K1"), ("elasC1", None), ("s11", "s11")):
        if b:
            assert getattr(P, a) == p[b]
    assert (P.elasC1, P.elasC2, P.elasC3) == p["elasC_raw"]
    from cupymag_amd.utils.sigma_matrices import get_Ebar_sigma
    assert np.array_equal(get_Ebar_sigma(P), p["Ebar_sigma"])


def test_config_loader_and_cli_errors(tmp_path):
    import yaml

    from cupymag_amd.cli import main
    from cupymag_amd.core.parameters import load_parameters
    assert main([str(tmp_path / "missing.yaml")]) == 1
    cfg = config("Hex")
    cfg["output"]["directory"] = "out"
    path = tmp_path / "c.yaml"
    path.write_text(yaml.safe_dump(cfg))
    P = load_parameters(path)
    assert P.output_dir == (tmp_path / "out").resolve()
    cfg["grid"]["type"] = "Prism"
    path.write_text(yaml.safe_dump(cfg))
    with pytest.raises(NotImplementedError):
        load_parameters(path)


def test_coo_api_matches_native_csr():
    _lib_or_skip()
    from scipy.sparse import coo_matrix

    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.physics.assemble_coo import AssembleDemag, AssembleElasticity
    from oracle import mesh as om
    cfg = config("Tet", me=True, defect_params=1.2)
    arr = kuhn_mesh(4, jitter=0.1, seed=2)
    p, ops, Avg, (coords, el, gid) = oracle_setup(cfg, arr)
    n = ops["n"]
    r, c, v = AssembleDemag(coords, el, gid).build_coo_matrix_A_numba()
    A = coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    assert abs(A - ops["A_demag"]).max() < 1e-14 * abs(A).max()
    E = AssembleElasticity(coords, el, gid, p["c11"], p["c12"], p["c44"], p["lambda100"], p["lambda111"])
    r, c, v = E.build_coo_matrix_A_numba()
    A = coo_matrix((v, (r, c)), shape=(3 * n, 3 * n)).tocsr()
    assert abs(A - ops["A_el"]).max() < 1e-13 * abs(A).max()
    r, c, v = E.build_coo_matrices_F_numba()
    F = coo_matrix((v, (r, c)), shape=(3 * n, 6 * n)).tocsr()
    assert abs(F - ops["F_el"]).max() < 1e-13 * abs(F).max()


def test_vtu_writer_roundtrip(tmp_path):
    import base64
    import struct
    import xml.etree.ElementTree as ET

    from cupymag_amd.utils.volume_average import write_vtu
    pts = np.random.default_rng(0).standard_normal((9, 3))
    cells = np.array([[0, 1, 2, 3], [4, 5, 6, 7]])
    write_vtu(tmp_path / "a.vtu", pts, cells, {"m1": np.arange(9.0)}, {"defect_flag": np.array([0, 1], np.int32)})
    root = ET.parse(tmp_path / "a.vtu").getroot()
    arrays = {a.get("Name"): a for a in root.iter("DataArray")}

    def dec(a, dt):
        raw = base64.b64decode(a.text)
        nbytes = struct.unpack("<I", raw[:4])[0]
        return np.frombuffer(raw[4:4 + nbytes], dtype=dt)
    assert np.array_equal(dec(arrays["m1"], "<f8"), np.arange(9.0))
    assert np.array_equal(dec(arrays["connectivity"], "<i8").reshape(2, 4), cells)
    assert np.array_equal(dec(arrays["defect_flag"], "<i4"), [0, 1])


def _blend_brute_force(coords, el, gid, f, alpha, eps=1e-14):
    """The ParaView point data of volume_average.py:219-318, node by node
    with explicit loops: element value = sum_g |detJ_g| w_g f(x_g) /
    sum_g |detJ_g| w_g (Jacobians from coords[gid], the oracle's volume
    average), or the corner mean when that volume is <= eps; node value =
    mean over the elements holding the node (the DOF value when none);
    then the mean over the periodic images of the DOF and the alpha mix."""
    from oracle import fem
    npe = el.shape[1] - 1
    V = fem.VolumeAverage(coords, el, gid)
    vw = V.detJ * V.W
    f = f.reshape(f.shape[0], -1)
    ev = np.zeros((el.shape[0], f.shape[1]))
    for e in range(el.shape[0]):
        dofs = gid[el[e, :npe]]
        den = sum(vw[e, g] for g in range(vw.shape[1]))
        if den > eps:
            ev[e] = sum(vw[e, g] * sum(V.Ng[g, b] * f[dofs[b]] for b in range(npe))
                        for g in range(vw.shape[1])) / den
        else:
            ev[e] = sum(f[d] for d in dofs) / npe
    node = np.zeros((coords.shape[0], f.shape[1]))
    for i in range(coords.shape[0]):
        hold = [e for e in range(el.shape[0]) if i in el[e, :npe]]
        node[i] = sum(ev[e] for e in hold) / len(hold) if hold else f[gid[i]]
    out = np.zeros_like(node)
    for i in range(coords.shape[0]):
        imgs = np.flatnonzero(gid == gid[i])
        out[i] = alpha * f[gid[i]] + (1.0 - alpha) * node[imgs].mean(axis=0)
    return out


@pytest.mark.parametrize("case", ["hex4_defect", "hex3_orphan_node", "hex2_degenerate", "kuhn3_jitter"])
def test_vtu_blend_matches_per_node_evaluation(case, tmp_path):
    """The blend operator built once per mesh gives the reference's point
    values: checked against the node-by-node evaluation on small meshes,
    including the reference's Jacobians from coords[gid] (wrong volumes on
    periodic images), a node in no element, a grid whose every element
    volume is zero (corner-mean branch) and a jittered tet mesh."""
    from oracle import mesh as om
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.utils.volume_average import VolumeAverage
    if case.startswith("hex"):
        n = int(case[3])
        coords, el, b = om.hex_grid(n, n, n, 2, 2, 2)
        if case == "hex3_orphan_node":
            el = el[1:]                                  # node 0 (a box corner) is in element 0 only
        gid = om.periodic_map(coords, b)
    else:
        coords, tets = kuhn_mesh(3, jitter=0.15, seed=5)
        flags = om.tet_defect_flags(coords, tets, om.sphere(0.8))
        el = np.concatenate([tets, flags[:, None]], axis=1).astype(np.int32)
        gid = om.periodic_map(coords, om.tet_bounds(coords))
    ndof = int(gid.max()) + 1
    rng = np.random.default_rng(11)
    fields = {"m": rng.standard_normal((ndof, 3)), "U": rng.standard_normal(ndof)}
    V = VolumeAverage(coords, el, gid)
    for alpha in (0.05, 0.5, 1.0):
        pd = V.blended_point_data(fields, alpha=alpha)
        for name, f in fields.items():
            want = _blend_brute_force(coords, el, gid, f, alpha)
            got = (np.stack([pd[f"{name}_{c + 1}"] for c in range(3)], axis=1) if f.ndim == 2
                   else pd[name][:, None])
            assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max()), (case, alpha, name)
    if case == "hex2_degenerate":
        assert np.all(V.vw.sum(axis=1) <= 1e-14)
    if case == "hex3_orphan_node":
        assert not np.any(el[:, :8] == 0)
    V.write_to_paraview(fields, str(tmp_path / "f.vtu"), alpha=0.5)
    assert (tmp_path / "f.vtu").stat().st_size > 0


def test_cfg2_golden_file_matches_its_case():
    """The committed oracle file of the 64^3 magnetoelastic workload belongs
    to tests.cases.cfg2_case (same config, DOF count, sample DOFs), and its
    oracle magnetization is physical: |m| = 1 or 0 (defect) at the samples,
    block sums of |m|-bounded entries, and CG iterations of a plausible
    magnitude for this mesh (elasticity far above demag above Gauss-Seidel)."""
    import json
    import os

    from tests.cases import CFG2_BLOCK, CFG2_SAMPLES, cfg2_case
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg2_fp64_oracle.npz"), allow_pickle=False)
    cfg, (coords, tets), _ = cfg2_case()
    assert json.loads(str(fx["cfg"])) == json.loads(json.dumps(cfg, sort_keys=True))
    n = int(fx["n_dof"])
    assert n == 64 ** 3 and len(tets) == 6 * n
    idx = np.sort(np.random.default_rng(99).choice(n, CFG2_SAMPLES, replace=False))
    for k in (1, int(fx["steps"])):
        np.testing.assert_array_equal(fx[f"s{k}_idx"], idx)
        nrm = np.linalg.norm(fx[f"s{k}_m_samples"], axis=1)
        assert np.all((np.abs(nrm - 1) < 1e-12) | (nrm == 0)) and (nrm == 0).any()
        assert np.abs(fx[f"s{k}_m_blocks"]).max() <= CFG2_BLOCK
        assert fx[f"s{k}_u_blocks"].shape == (n // CFG2_BLOCK, 3)
    it = np.asarray(fx["iters"])
    assert it.shape == (int(fx["steps"]), 10)
    assert np.all(it[:, 0] > it[:, 1]) and np.all(it[:, 1] > it[:, 2:].max(axis=1))


def _sweep_ctl_host(tmp_path):
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    root = Path(__file__).resolve().parent.parent
    exe = str(tmp_path / "sweep_ctl_host")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-std=c++17", f"-I{root}/cupymag_amd/csrc",
                    f"-I{root}/include", str(root / "tests" / "sweep_ctl_host.hip"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def _reference_loop(seq, H0, dH, acc, initf, minstep, stop_dir, stop_value, fields=None, spp=0):
    """Micromagnetics.py:109,299-335 (and, with a field list and a per-point
    cap, the sweep server's rule) step by step: per step use_init,
    point_done, state, point, nstep, H of the step, H of the next step."""
    H = np.array(fields[0] if fields is not None else H0, dtype=np.float64)
    nstep = point = 0
    state = 0
    out = []
    for sumsq, avg in seq:
        if state:
            break
        nstep += 1
        use_init = sumsq < initf * acc
        Hs = H.copy()
        conv = sumsq < acc and nstep > minstep
        done = conv or (spp > 0 and nstep >= spp)
        if done:
            point += 1
            nstep = 0
            if fields is not None:
                if point >= len(fields):
                    state = 2
                else:
                    H = np.array(fields[point], dtype=np.float64)
            elif conv:
                H = H - np.asarray(dH)
        if state == 0 and not (np.dot(avg, stop_dir) > stop_value):
            state = 1
        out.append([float(use_init), float(done), float(state), float(point), float(nstep), *Hs, *H])
    return np.array(out)


@pytest.mark.parametrize("kind", ["field_step", "field_list"])
def test_device_sweep_control_rule_matches_the_reference_loop(kind, tmp_path):
    """The hysteresis control the device applies after every step
    (sweep_ctl.h, built for the host) gives, step by step, the reference
    loop's use_init, field points, field steps and stop (Micromagnetics.py
    :109,299-335) on a scripted sumsq / avg_m sequence that converges,
    stalls, reverses and crosses the stop value; and the server's rule with
    a field list and a per-point step cap."""
    import subprocess
    exe = _sweep_ctl_host(tmp_path)
    rng = np.random.default_rng(4)
    nst = 700
    acc, initf, minstep = 1e-3, 100.0, 10
    sumsq = np.where(rng.random(nst) < 0.7, 10.0 ** rng.uniform(-6, -2, nst), 10.0 ** rng.uniform(-1, 2, nst))
    sumsq[:40] = 1e-9                      # converges as soon as nstep > minstep
    avg = np.zeros((nst, 3))
    avg[:, 0] = np.linspace(1.0, -1.0, nst)  # reverses: the stop value is crossed near the end
    avg[:, 1] = 0.1
    stop_dir, stop_value = np.array([1.0, 0.0, 0.0]), -0.9
    H0, dH = np.array([0.05, 0.0, 0.01]), np.array([0.002, 0.0, 0.0])
    fields, spp = None, 0
    if kind == "field_list":
        fields = np.linspace(-0.08, 0.08, 9)[:, None] * np.array([[0.6, 0.0, 0.8]])
        spp, stop_value = 37, -2.0
    c = np.zeros(64)
    c[32] = 1.0
    c[33:36] = fields[0] if fields is not None else H0
    c[36:39] = dH
    c[39], c[40] = acc, initf
    c[41:44], c[44] = stop_dir, stop_value
    c[45], c[46] = minstep, spp
    c[47] = 0 if fields is None else len(fields)
    c[53:56] = c[33:36]
    blob = c.tobytes() + np.array([c[47]]).tobytes()
    if fields is not None:
        blob += np.ascontiguousarray(fields).tobytes()
    blob += np.array([float(nst)]).tobytes() + np.concatenate([sumsq[:, None], avg], axis=1).tobytes()
    got = np.frombuffer(subprocess.run([exe], input=blob, capture_output=True, check=True).stdout).reshape(nst, 11)
    want = _reference_loop(zip(sumsq, avg), H0, dH, acc, initf, minstep, stop_dir, stop_value, fields, spp)
    m = len(want)
    assert m > 100 and want[:, 1].sum() >= 3              # several field points
    assert want[-1, 2] == (1 if kind == "field_step" else 2)  # ends on the stop value / the list's end
    np.testing.assert_array_equal(got[:m], want)
    assert np.all(got[m:, :5] == got[m - 1, :5])           # once stopped, the state stays


@pytest.fixture(scope="module")
def _asan_assembler(tmp_path_factory):
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    root = Path(__file__).resolve().parent.parent
    exe = str(tmp_path_factory.mktemp("asan") / "assembly_sanitize")
    subprocess.run([gxx, "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-std=c++17", "-pthread", f"-I{root}/include",
                    f"-I{root}/cupymag_amd/csrc", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    str(root / "cupymag_amd" / "csrc" / "host_assembly.cpp"), str(root / "tests" / "assembly_sanitize.cpp"),
                    "-o", exe], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("grid,elastic,anchors,threads", [("Hex", False, True, 1), ("Hex", True, True, 4),
                                                          ("Tet", True, True, 4), ("Tet", True, False, 3)])
def test_host_assembler_under_address_and_ub_sanitizers(_asan_assembler, grid, elastic, anchors, threads):
    """The native host assembler built with AddressSanitizer and
    UndefinedBehaviorSanitizer (no recovery: any finding aborts the run)
    assembles a hex grid and a jittered Kuhn mesh with a defect, with and
    without the elasticity operators and anchors, on several threads; every
    array equals the library build's bit for bit."""
    import os
    import subprocess

    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.mesh.setup_FEM_mesh import FEMMesh
    from cupymag_amd.physics.host_ops import HostOps, voigt_C
    cfg = config(grid, n=5, me=elastic, ndef=2, defect_params=1.3)
    P = from_dict(cfg)
    M = FEMMesh(P, kuhn_mesh(5, jitter=0.15, seed=8) if grid == "Tet" else None)
    coords, el, gid, n = M.node_coords_np, M.elements_np, M.global_id_np, M.n_dof
    Cv = voigt_C(P.c11, P.c12, P.c44, False) if elastic else None
    npe = el.shape[1] - 1
    mask = np.zeros(n)
    mask[np.asarray(M.DefDOF)] = 1
    coef = np.zeros(38)
    coef[0], coef[1] = P.alpha, P.alpha * P.alpha_damp
    if elastic:
        coef[2:] = Cv.ravel()
    flags = int(elastic) | (0 if anchors else 2)
    blob = np.concatenate([[coords.shape[0]], coords.ravel(), [el.shape[0], npe, el.shape[1]], el.ravel(), [n],
                           gid, mask, coef, [flags, threads]]).astype(np.float64).tobytes()
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([_asan_assembler], input=blob, capture_output=True, env=env)
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    out = r.stdout
    H = HostOps(coords, el, gid, n, M.DefDOF, coef[0], coef[1], Cv, nthreads=threads, anchors=anchors)
    names = ["PTR", "IDX", "ADEMAG", "FX", "FY", "FZ", "GS_PTR", "GS_IDX", "A1", "A2", "F1_PTR", "F1_IDX", "F1",
             "EL_PTR", "EL_IDX", "AEL", "FE_PTR", "FE_IDX", "FEL"]
    pos = 0
    for name in names:
        ln = int(np.frombuffer(out[pos:pos + 8], dtype=np.int64)[0])
        pos += 8
        dt = np.int32 if name.endswith("IDX") else (np.int64 if name.endswith("PTR") else np.float64)
        got = np.frombuffer(out[pos:pos + ln * np.dtype(dt).itemsize], dtype=dt)
        pos += ln * np.dtype(dt).itemsize
        if name.startswith(("EL", "FE", "AEL")) and not elastic:
            assert ln == 0
            continue
        np.testing.assert_array_equal(got, np.asarray(H.array(name)), err_msg=name)
    assert pos < len(out)  # the element matrices follow


def test_xcd_order_of_the_stencil_launches_is_a_permutation(tmp_path):
    """The XCD order of the workgroups of the stencil launches (device.h:
    xcd_perm) is a bijection of the workgroup positions for every grid size
    and group, and gives every XCD consecutive workgroups."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    root = Path(__file__).resolve().parent.parent
    exe = str(tmp_path / "row_order_host")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-std=c++17", f"-I{root}/cupymag_amd/csrc",
                    f"-I{root}/include", str(root / "tests" / "row_order_host.hip"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe, "300"], check=True, capture_output=True, text=True).stdout.split("\n")
    perms = [line.split() for line in out if line]
    assert len(perms) == 300 * 8
    assert all(p[3] == "1" for p in perms)
    assert all(p[4] == "1" for p in perms if int(p[2]) > 1 or int(p[2]) < 0)
