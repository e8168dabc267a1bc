"""Mesh inputs of the GSPM step (oracle, test infrastructure).

* ``hex_grid``      -- ``cupymag/mesh/gridHex.py:68-142``: unit cubes on
  [0,Nx]x[0,Ny]x[0,Nz], node id i*(Ny+1)*(Nz+1)+j*(Nz+1)+k, element order
  i-major, k fastest, defect flag = cell centre in the half-open box
  [c-Nd/2, c+Nd/2).
* ``read_nas``      -- ``cupymag/mesh/gridTet.py:88-152``: comma-separated
  GRID/CTETRA cards, '$' comments and blank lines skipped, stop at ENDDATA,
  node ids remapped through argsort/searchsorted.
* ``tet_defect_flags`` -- ``gridTet.py:72-86,155-178`` with the shape
  functions of ``cupymag/utils/defect_shapes.py:73-134``.
* ``periodic_map``  -- ``gridHex.py:179-213`` / ``gridTet.py:180-225``: the
  reference's own dictionary loop over nodes in order, first-appearance
  numbering of ``round(wrapped coordinate, 12)`` keys.

Pinning: ``periodic_map`` is checked against a brute-force pairwise search
on small grids (tests/test_oracle.py); ``hex_grid`` node/element numbering
against the closed form above; ``read_nas`` by a write/read round trip of
meshes with shuffled node ids.
"""
import numpy as np


def hex_grid(Nx, Ny, Nz, Ndx, Ndy, Ndz, defect_center=None):
    x = np.linspace(0, Nx, Nx + 1)
    y = np.linspace(0, Ny, Ny + 1)
    z = np.linspace(0, Nz, Nz + 1)
    i3, j3, k3 = np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1),
                             np.arange(Nz + 1), indexing="ij")
    coords = np.stack((x[i3.ravel()], y[j3.ravel()], z[k3.ravel()]), axis=1)
    if defect_center is None:
        cx, cy, cz = Nx / 2.0, Ny / 2.0, Nz / 2.0
    else:
        cx, cy, cz = defect_center
    lo = (cx - Ndx / 2.0, cy - Ndy / 2.0, cz - Ndz / 2.0)
    hi = (cx + Ndx / 2.0, cy + Ndy / 2.0, cz + Ndz / 2.0)
    elements = np.zeros((Nx * Ny * Nz, 9), dtype=np.int32)
    e = 0
    for i in range(Nx):
        for j in range(Ny):
            for k in range(Nz):
                def nid(a, b, c):
                    return a * (Ny + 1) * (Nz + 1) + b * (Nz + 1) + c
                elements[e, :8] = (nid(i, j, k), nid(i + 1, j, k),
                                   nid(i + 1, j + 1, k), nid(i, j + 1, k),
                                   nid(i, j, k + 1), nid(i + 1, j, k + 1),
                                   nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1))
                xc, yc, zc = i + 0.5, j + 0.5, k + 0.5
                elements[e, 8] = int(lo[0] <= xc < hi[0] and lo[1] <= yc < hi[1]
                                     and lo[2] <= zc < hi[2])
                e += 1
    return coords, elements, (x[0], x[-1], y[0], y[-1], z[0], z[-1])


def read_nas(path):
    grids, tets = [], []
    with open(path, "r") as fh:
        for raw in fh:
            if raw.startswith("$") or raw.isspace():
                continue
            if raw.lstrip().startswith("ENDDATA"):
                break
            line = raw.rstrip()
            if line.startswith("GRID"):
                grids.append([t.strip() for t in line.split(",")])
            elif line.startswith("CTETRA"):
                tets.append([t.strip() for t in line.split(",")])
    if not grids or not tets:
        raise ValueError("File contains no GRID or CTETRA cards")
    node_ids = np.array([int(g[1]) for g in grids], dtype=np.int32)
    xyz = np.array([[float(g[3]), float(g[4]), float(g[5])] for g in grids],
                   dtype=np.float64)
    en = np.array([[int(t[3]), int(t[4]), int(t[5]), int(t[6])] for t in tets],
                  dtype=np.int32)
    order = np.argsort(node_ids)
    idx = order[np.searchsorted(node_ids[order], en, side="left")]
    return xyz, idx.astype(np.int32)


def sphere(R):
    R2 = float(R) * float(R)
    return lambda rel: np.sum(rel ** 2, axis=1) < R2


def ellipsoid(abc):
    a, b, c = abc
    inv2 = np.array([1.0 / a ** 2, 1.0 / b ** 2, 1.0 / c ** 2])
    return lambda rel: np.sum(rel ** 2 * inv2[None, :], axis=1) < 1.0


def rotated_ellipsoid(abc, rot):
    a, b, c = abc
    R = np.asarray(rot, dtype=float)
    Q = R @ np.diag([1.0 / a ** 2, 1.0 / b ** 2, 1.0 / c ** 2]) @ R.T
    return lambda rel: np.sum((rel @ Q) * rel, axis=1) < 1.0


SHAPES = {"Sphere": lambda p: sphere(p),
          "Ellipsoid": lambda p: ellipsoid(p),
          "RotatedEllipsoid": lambda p: rotated_ellipsoid(p[0], p[1])}


def tet_defect_flags(coords, tets, shape_fn, defect_center=None):
    if defect_center is None:
        defect_center = (coords.min(axis=0) + coords.max(axis=0)) / 2
    centers = coords[tets].mean(axis=1)
    rel = centers - np.asarray(defect_center, dtype=np.float64)
    return shape_fn(rel).astype(np.int32)


def periodic_map(coords, bounds, tol=1e-12):
    """bounds = (x0, x1, y0, y1, z0, z1); a coordinate within tol of the upper
    bound is wrapped to the lower one."""
    x0, x1, y0, y1, z0, z1 = (float(b) for b in bounds)

    def wrap(v, a, b):
        return float(a) if abs(v - b) < tol else float(v)

    canon = {}
    gid = np.zeros(coords.shape[0], dtype=np.int32)
    nxt = 0
    for n in range(coords.shape[0]):
        key = (round(wrap(coords[n, 0], x0, x1), 12),
               round(wrap(coords[n, 1], y0, y1), 12),
               round(wrap(coords[n, 2], z0, z1), 12))
        if key not in canon:
            canon[key] = nxt
            nxt += 1
        gid[n] = canon[key]
    return gid


def tet_bounds(coords):
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    return (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])


def defect_dofs(elements, gid):
    """extract_defect_dofs, final_assembly.py:27-57."""
    mask = elements[:, -1] == 1
    if not mask.any():
        return np.zeros(0, dtype=np.int32)
    nodes = np.unique(elements[mask, :-1].ravel())
    return np.unique(gid[nodes]).astype(np.int32)
