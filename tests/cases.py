"""Shared test inputs: YAML-equivalent configs and small meshes."""
import copy

import numpy as np
import scipy.sparse as sp

BASE = {
    "physics": {"magnetic": {"ms": 1.2e6, "exchange_stiffness": 1.5e-11, "anisotropy_k1": -1e4},
                "elastic": {"c11": 2.4e11, "c12": 1.4e11, "c44": 1.2e11},
                "magnetostriction": {"lambda100": 2.6e-4, "lambda111": -3e-5},
                "demagnetization": {"n_x": 1 / 3, "n_y": 1 / 3, "n_z": 1 / 3},
                "external_field": {"hext1": 5e4, "dhext1": 1e4, "hext2": 0.0, "hext3": 0.0},
                "external_stress": {"sigma11": 0.0},
                "rotation_111": False, "magnetoelastic_coupling": False,
                "characteristic_length": 2e-9},
    "grid": {"type": "Hex", "cubic_mesh": {"nx": 8, "ny": 8, "nz": 8, "ndx": 2, "ndy": 2, "ndz": 2}},
    "simulation": {"precision": "DP",
                   "llg": {"timestep": 1e-12, "alpha_damping": 0.5, "llg_accuracy_factor": 1e-30},
                   "solver": {"tolerance": 1e-10, "max_iterations": 5000, "initial_factor": 100.0},
                   "stop_condition": {"direction": [1.0, 0.0, 0.0], "value": -0.95}},
    "output": {"write_magnetization": False},
}


def config(grid="Hex", n=8, me=False, rot=False, dt=1e-12, alpha=0.5, sigma11=0.0, ndef=2, tol=1e-10,
           hext=(5e4, 0.0, 0.0), K1=-1e4, defect="Sphere", defect_params=2.0):
    c = copy.deepcopy(BASE)
    c["physics"]["magnetoelastic_coupling"] = me
    c["physics"]["rotation_111"] = rot
    c["physics"]["external_stress"]["sigma11"] = sigma11
    c["physics"]["magnetic"]["anisotropy_k1"] = K1
    c["physics"]["external_field"].update(hext1=hext[0], hext2=hext[1], hext3=hext[2])
    c["simulation"]["llg"].update(timestep=dt, alpha_damping=alpha)
    c["simulation"]["solver"]["tolerance"] = tol
    if grid == "Hex":
        c["grid"] = {"type": "Hex", "cubic_mesh": {"nx": n, "ny": n, "nz": n, "ndx": ndef, "ndy": ndef,
                                                   "ndz": ndef}}
    else:
        c["grid"] = {"type": "Tet", "tet_mesh": {"mesh_file": None, "defect_shape": defect,
                                                 "defect_parameters": defect_params}}
    return c


def tilted_m(n, seed=0, max_deg=5.0):
    """+x with per-DOF random tilts <= max_deg (seeded)."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(max_deg) * rng.random(n)
    ph = 2 * np.pi * rng.random(n)
    return np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], axis=1)


def oracle_mesh(cfg, tet_arrays=None):
    """(coords, elements, gid) as the oracle builds them."""
    from oracle import mesh as om
    g = cfg["grid"]
    if g["type"] == "Hex":
        cm = g["cubic_mesh"]
        coords, el, b = om.hex_grid(cm["nx"], cm["ny"], cm["nz"], cm["ndx"], cm["ndy"], cm["ndz"])
        return coords, el, om.periodic_map(coords, b)
    coords, tets = tet_arrays
    tm = g["tet_mesh"]
    p = tm["defect_parameters"]
    shape = om.SHAPES[tm["defect_shape"]](p if tm["defect_shape"] != "RotatedEllipsoid" else
                                          (p, np.array(tm["rotation_matrix"])))
    flags = om.tet_defect_flags(coords, tets, shape)
    el = np.concatenate([tets, flags[:, None]], axis=1).astype(np.int32)
    return coords, el, om.periodic_map(coords, om.tet_bounds(coords))


def oracle_setup(cfg, tet_arrays=None):
    from oracle import fem, params
    p = params.reduced(cfg)
    coords, el, gid = oracle_mesh(cfg, tet_arrays)
    ops = fem.build_operators(coords, el, gid, p)
    Avg = fem.VolumeAverage(coords, el, gid)
    return p, ops, Avg, (coords, el, gid)


def unanchor(A, bs):
    """The periodic stiffness K of an anchored operator A (node 0's rows and
    columns replaced by the identity): K annihilates the bs translations, so
    K's block in node 0's column of row i != 0 is minus the sum of the
    row's other blocks, row 0 is the transpose of that column (K symmetric)
    and its diagonal block minus the sum of its other blocks.  Entries only
    where A's pattern holds them."""
    A = A.tocsr().astype(np.float64)
    N = A.shape[0]
    comp = np.arange(N) % bs
    Ac = A.tocsc()
    rows0 = set()
    for e in range(bs):
        rows0 |= set(Ac.indices[Ac.indptr[e]:Ac.indptr[e + 1]].tolist())
    rows0 = np.array(sorted(r for r in rows0 if r >= bs), dtype=np.int64)
    B = A.tolil()
    B[:bs, :] = 0.0
    B[:, :bs] = 0.0
    B = B.tocsr()
    col = np.zeros((N, bs))
    for e in range(bs):
        s_e = np.asarray(B[:, np.flatnonzero(comp == e)].sum(axis=1)).ravel()
        col[rows0, e] = -s_e[rows0]
    Cm = sp.lil_matrix((N, N))
    for e in range(bs):
        Cm[rows0, e] = col[rows0, e].reshape(-1, 1)
    Cm = Cm.tocsr()
    D0 = np.zeros((bs, bs))
    for d in range(bs):
        for e in range(bs):
            D0[d, e] = -col[rows0[comp[rows0] == e], d].sum()
    Dm = sp.lil_matrix((N, N))
    Dm[:bs, :bs] = D0
    return (B + Cm + Cm.T + Dm.tocsr()).tocsr()


def cfg2_case(n=64):
    """The 64^3 magnetoelastic workload: Kuhn tets (6 per cube) with seeded
    node jitter <= 0.15 h, a spherical non-magnetic defect of radius 8 h at
    the centre, magnetoelastic coupling on (c11 2.4e11, c12 1.4e11, c44
    1.2e11 Pa, lambda100 2.6e-4, lambda111 -3e-5), Hext (5e4, 0, 0) A/m,
    dt 5 ps, alpha 0.5, CG tol 1e-10; m = +x tilted <= 5 deg per DOF (seed 7).
    Returns (cfg, (coords, tets), m0 as a function of the DOF count)."""
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    from cupymag_amd.serving.workloads import instance_config
    cfg = instance_config((5e4, 0.0, 0.0), me=True, defect_shape="Sphere", defect_params=8.0 * n / 64.0,
                          dt=5e-12, alpha=0.5)
    return cfg, kuhn_mesh(n, jitter=0.15, seed=3), (lambda ndof: tilted_m(ndof, seed=7))


CFG2_SAMPLES = 4096   # DOFs sampled per field in the golden file
CFG2_BLOCK = 256      # DOFs per block sum


def cfg2_summary(n_dof, fields):
    """What the golden file keeps of full fields (dict name -> (n_dof, C)
    array): the values at CFG2_SAMPLES seeded DOFs and the sums over blocks
    of CFG2_BLOCK consecutive DOFs (a local error anywhere moves its block's
    sum)."""
    idx = np.sort(np.random.default_rng(99).choice(n_dof, CFG2_SAMPLES, replace=False))
    out = {"idx": idx.astype(np.int32)}
    nb = n_dof // CFG2_BLOCK
    for k, v in fields.items():
        v = np.asarray(v, dtype=np.float64).reshape(n_dof, -1)
        out[f"{k}_samples"] = v[idx]
        out[f"{k}_blocks"] = v[:nb * CFG2_BLOCK].reshape(nb, CFG2_BLOCK, -1).sum(axis=1)
    return out
