"""Seeded synthetic workloads for batches and the sweep server.

Both use the reference's YAML schema (parameters.py:60-174) for every
instance, Kuhn tetrahedral meshes of n^3 cells (6 tets per cube, h = 1 in
mesh units, defect sizes in the same units) and the material of the
headline workload: Ms = 1.2e6 A/m, A = 1.5e-11 J/m, K1 = -1e4 J/m^3,
demag factors (1/3, 1/3, 1/3), c11 = 2.4e11, c12 = 1.4e11, c44 = 1.2e11 Pa,
lambda100 = 2.6e-4, lambda111 = -3e-5, dt = 10 ps, alpha = 0.1, CG tol
1e-10.

* ``batch_workload``: independent simulations with n log-uniform in
  [16, 96], a spherical defect of radius uniform in [2h, 0.3 n h], an
  external field of magnitude uniform in [0, 1e5] A/m along a random unit
  direction, magnetoelastic coupling on in every other instance.
* ``sweep_requests``: hysteresis-sweep requests arriving as a Poisson
  stream, n log-uniform in [20, 128], a random rotated-ellipsoid defect
  (semi-axes uniform in [2h, 0.3 n h], random rotation), field points
  linspace(-1e5, 1e5) A/m along a random unit axis, a fixed number of
  steps per point.
"""
import copy

import numpy as np

MATERIAL = {
    "magnetic": {"ms": 1.2e6, "exchange_stiffness": 1.5e-11, "anisotropy_k1": -1e4},
    "elastic": {"c11": 2.4e11, "c12": 1.4e11, "c44": 1.2e11},
    "magnetostriction": {"lambda100": 2.6e-4, "lambda111": -3e-5},
    "demagnetization": {"n_x": 1 / 3, "n_y": 1 / 3, "n_z": 1 / 3},
}


def instance_config(hext, me, defect_shape, defect_params, rot111=False, dt=1e-11, alpha=0.1, tol=1e-10,
                    maxiter=20000, llg_accuracy_factor=1e-30):
    """A YAML-equivalent config dict of one Kuhn-mesh instance (the mesh
    itself is passed as arrays)."""
    return {
        "physics": {**copy.deepcopy(MATERIAL),
                    "external_field": {"hext1": float(hext[0]), "hext2": float(hext[1]), "hext3": float(hext[2]),
                                       "dhext1": 0.0},
                    "external_stress": {},
                    "rotation_111": bool(rot111), "magnetoelastic_coupling": bool(me),
                    "characteristic_length": 2e-9},
        "grid": {"type": "Tet", "tet_mesh": {"mesh_file": None, "defect_shape": defect_shape,
                                             "defect_parameters": defect_params}},
        "simulation": {"precision": "DP",
                       "llg": {"timestep": dt, "alpha_damping": alpha,
                               "llg_accuracy_factor": llg_accuracy_factor},
                       "solver": {"tolerance": tol, "max_iterations": maxiter, "initial_factor": 100.0},
                       "stop_condition": {"direction": [1.0, 0.0, 0.0], "value": -2.0}},
        "output": {"write_magnetization": False},
    }


def headline_config(n=128, tol=1e-12, seed=2024):
    """The single large instance: a Kuhn mesh of n^3 cells (12.6M tets at
    n = 128), a rotated-ellipsoid defect with semi-axes (20, 12, 8) h * n/128
    and a seeded rotation (QR of a Gaussian 3x3, det +1), magnetoelastic
    coupling on, the crystal x axis along [111], Hext (5e4, 0, 0) A/m,
    dt 10 ps, alpha 0.1.  tol: CG tolerance -- at 1e-10 the trajectory of
    this strongly driven state (m reverses within 50 steps) is itself
    determined only to ~1e-6 after 50 steps by the FP64 solves (their
    difference from the same solves at 1e-13), so 1e-12 is the default."""
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    s = n / 128.0
    cfg = instance_config((5e4, 0.0, 0.0), me=True, defect_shape="RotatedEllipsoid",
                          defect_params=[20.0 * s, 12.0 * s, 8.0 * s], rot111=True, dt=1e-11, alpha=0.1, tol=tol)
    cfg["grid"]["tet_mesh"]["rotation_matrix"] = q.tolist()
    cfg["simulation"]["stop_condition"]["value"] = -0.95
    return cfg


def _unit(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def _rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _log_uniform_int(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def batch_workload(count=64, seed=4, nmin=16, nmax=96, **kw):
    """``count`` instance descriptions: dict(n, cfg, m_seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = _log_uniform_int(rng, nmin, nmax)
        r = rng.uniform(2.0, max(2.0, 0.3 * n))
        hext = rng.uniform(0.0, 1e5) * _unit(rng)
        out.append(dict(n=n, cfg=instance_config(hext, me=(k % 2 == 0), defect_shape="Sphere", defect_params=r, **kw),
                        m_seed=int(rng.integers(1 << 30))))
    return out


def sweep_requests(count, rate=20.0, seed=5, nmin=20, nmax=128, points=41, hmax=1e5, steps_per_point=200, me=True,
                   **kw):
    """``count`` requests: dict(rid, arrival [s], n, cfg, fields (points, 3)
    A/m, steps_per_point, m_seed).  Arrivals are a Poisson process of
    ``rate`` per second; the sweep starts saturated along -axis."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        n = _log_uniform_int(rng, nmin, nmax)
        abc = [float(rng.uniform(2.0, max(2.0, 0.3 * n))) for _ in range(3)]
        rot = _rotation(rng)
        axis = _unit(rng)
        fields = np.linspace(-hmax, hmax, points)[:, None] * axis[None, :]
        cfg = instance_config(fields[0], me=me, defect_shape="RotatedEllipsoid", defect_params=abc, **kw)
        cfg["grid"]["tet_mesh"]["rotation_matrix"] = rot.tolist()
        out.append(dict(rid=k, arrival=t, n=n, cfg=cfg, fields=fields, steps_per_point=int(steps_per_point),
                        m_seed=int(rng.integers(1 << 30))))
    return out


def initial_m(ndof, seed, max_deg=5.0, direction=(1.0, 0.0, 0.0)):
    """Unit vectors within ``max_deg`` of ``direction`` (seeded)."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(max_deg) * rng.random(ndof)
    ph = 2 * np.pi * rng.random(ndof)
    m = np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], axis=1)
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    if np.allclose(d, [1.0, 0.0, 0.0]):
        return m
    # rotate +x onto d (Rodrigues)
    ex = np.array([1.0, 0.0, 0.0])
    v = np.cross(ex, d)
    c = float(ex @ d)
    if np.linalg.norm(v) < 1e-12:
        return -m if c < 0 else m
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    R = np.eye(3) + vx + vx @ vx / (1.0 + c)
    return m @ R.T
