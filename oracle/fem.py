"""Element integrals, global operators and the volume average (oracle).

Restates, for an element array at a time:

* shape tables and Gauss rules -- ``cupymag/mesh/ShapeHex.py:254-393`` and
  ``cupymag/mesh/ShapeTet.py:205-255``.  The hex tables are the shape
  functions at +-0.57735026919 rounded to 12 decimals and the tet value table
  the Gauss abscissae rounded to 7 decimals; both are regenerated here from
  those definitions (this reproduces the reference's tables digit for digit).
* element matrices -- ``cupymag/physics/assemble_demag.py:95-184``,
  ``assemble_Gauss_Seidel.py:88-160``, ``assemble_elasticity.py:194-348``:
  signed detJ, J^-1 applied to reference gradients, zero element when
  |detJ| < 1e-14 at any Gauss point.
* COO layout, anchors and SciPy COO->CSR -- ``assemble_demag.py:186-341``,
  ``assemble_elasticity.py:350-527``, ``final_assembly.py:59-205``,
  ``core/assemble_simulation_operators.py:42-107``.
* volume average -- ``cupymag/utils/volume_average.py:69-216`` including the
  Jacobians taken from ``coords[global_id[elements]]``.

Pinning (tests/test_oracle.py): P1 tet stiffness against V*grad(N_a).grad(N_b)
and the mass matrix against the quadrature closed form with the 7-digit
table; unit-cube hex stiffness / mass against the textbook trilinear
matrices; isotropic elasticity (c11-c12 = 2 c44) against the Lame closed form
B^T C B; row sums K.1 = 0 and F_k.1 = 0 on periodic meshes; sum(M) = volume;
volume average of a constant field = the constant, and against a brute
dense re-evaluation on a 2^3 mesh.
"""
import numpy as np
from scipy.sparse import coo_matrix

HEX_G = 0.57735026919
TET_A, TET_B = 0.58541020, 0.13819660
_SX = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
_SY = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
_SZ = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)


def hex_points():
    pts = [(rx, ry, rz) for rx in (-HEX_G, HEX_G) for ry in (-HEX_G, HEX_G)
           for rz in (-HEX_G, HEX_G)]
    return np.array(pts, dtype=np.float64), np.ones(8)


def hex_N(r, s, t):
    return (1 + _SX * r) * (1 + _SY * s) * (1 + _SZ * t) / 8.0


def hex_dN(r, s, t):
    return np.stack([_SX * (1 + _SY * s) * (1 + _SZ * t) / 8.0,
                     _SY * (1 + _SX * r) * (1 + _SZ * t) / 8.0,
                     _SZ * (1 + _SX * r) * (1 + _SY * s) / 8.0], axis=1)


def tet_points():
    a, b = TET_A, TET_B
    return (np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]], dtype=np.float64),
            np.ones(4) / 24.0)


def tet_N(r, s, t):
    return np.array([1.0 - r - s - t, r, s, t])


TET_DN = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                   [0.0, 0.0, 1.0]])


def assembly_tables(npe):
    """Per-Gauss-point (N, dN, w) used by the assemblers (get_N / get_dN)."""
    if npe == 8:
        pts, w = hex_points()
        N = [np.round(hex_N(*p), 12) for p in pts]
        dN = [np.round(hex_dN(*p), 12) for p in pts]
    else:
        pts, w = tet_points()
        a7, b7 = round(TET_A, 7), round(TET_B, 7)
        tab = np.full((4, 4), b7)
        np.fill_diagonal(tab, a7)
        N = [tab[g] for g in range(4)]
        dN = [TET_DN for _ in range(4)]
    return N, dN, w


def _geometry(xe, dN_g, npe):
    """J, detJ, J^-1 per element (element_jacobian in Shape*.py)."""
    if npe == 8:
        J = np.einsum("eij,ik->ejk", xe, dN_g)
    else:
        J = xe[:, 1:, :] - xe[:, 0:1, :]
    det = np.linalg.det(J)
    bad = np.abs(det) < 1e-14
    Jsafe = J.copy()
    Jsafe[bad] = np.eye(3)
    inv = np.linalg.inv(Jsafe)
    return det, inv, bad


def _grads(inv, dN_g):
    # gx_a = Jinv[0,0] dN[a,0] + Jinv[0,1] dN[a,1] + Jinv[0,2] dN[a,2], etc.
    g = np.empty((inv.shape[0], dN_g.shape[0], 3))
    for j in range(3):
        g[:, :, j] = (inv[:, j, 0:1] * dN_g[None, :, 0] + inv[:, j, 1:2] * dN_g[None, :, 1]
                      + inv[:, j, 2:3] * dN_g[None, :, 2])
    return g


def element_matrices(coords, elements, want=("K", "M", "F")):
    """Scalar element matrices K_e, M_e, (Fx,Fy,Fz)_e for every element."""
    npe = elements.shape[1] - 1
    xe = coords[elements[:, :npe]]
    Ne = xe.shape[0]
    N, dN, w = assembly_tables(npe)
    K = np.zeros((Ne, npe, npe))
    M = np.zeros((Ne, npe, npe))
    F = np.zeros((3, Ne, npe, npe))
    anybad = np.zeros(Ne, dtype=bool)
    for g in range(len(w)):
        det, inv, bad = _geometry(xe, dN[g], npe)
        anybad |= bad
        dw = det * w[g]
        gr = _grads(inv, dN[g])
        if "K" in want:
            dot = np.einsum("eak,ebk->eab", gr, gr)
            K += dot * dw[:, None, None]
        if "M" in want:
            M += (N[g][:, None] * N[g][None, :])[None] * dw[:, None, None]
        if "F" in want:
            for c in range(3):
                F[c] += N[g][None, :, None] * gr[:, None, :, c] * dw[:, None, None]
    K[anybad] = 0.0
    M[anybad] = 0.0
    F[:, anybad] = 0.0
    return K, M, F


def voigt_C(c11, c12, c44, rot111=False, Mrot=None):
    C = np.zeros((6, 6))
    C[:3, :3] = c12
    C[0, 0] = C[1, 1] = C[2, 2] = c11
    C[3, 3] = C[4, 4] = C[5, 5] = c44
    if rot111:
        C = Mrot @ C @ Mrot.T        # assemble_elasticity.py:61-98,190-192
    return C


def _B(gr):
    Ne, npe, _ = gr.shape
    B = np.zeros((Ne, 6, 3 * npe))
    for a in range(npe):
        dx, dy, dz = gr[:, a, 0], gr[:, a, 1], gr[:, a, 2]
        c = 3 * a
        B[:, 0, c + 0] = dx
        B[:, 1, c + 1] = dy
        B[:, 2, c + 2] = dz
        B[:, 3, c + 0] = dy
        B[:, 3, c + 1] = dx
        B[:, 4, c + 1] = dz
        B[:, 4, c + 2] = dy
        B[:, 5, c + 0] = dz
        B[:, 5, c + 2] = dx
    return B


def elastic_element_matrices(coords, elements, C):
    """K_e = sum_g B^T C B detJ w  (3npe x 3npe);  F_e = sum_g B^T C S_g detJ w
    (3npe x 6npe), assemble_elasticity.py:194-348."""
    npe = elements.shape[1] - 1
    xe = coords[elements[:, :npe]]
    Ne = xe.shape[0]
    N, dN, w = assembly_tables(npe)
    K = np.zeros((Ne, 3 * npe, 3 * npe))
    Fe = np.zeros((Ne, 3 * npe, 6 * npe))
    anybad = np.zeros(Ne, dtype=bool)
    for g in range(len(w)):
        det, inv, bad = _geometry(xe, dN[g], npe)
        anybad |= bad
        dw = det * w[g]
        B = _B(_grads(inv, dN[g]))
        BtC = np.einsum("eki,kj->eij", B, C)
        K += np.einsum("eik,ekj->eij", BtC, B) * dw[:, None, None]
        for b in range(npe):
            Fe[:, :, 6 * b:6 * b + 6] += BtC * (N[g][b] * dw)[:, None, None]
    K[anybad] = 0.0
    Fe[anybad] = 0.0
    return K, Fe


def scalar_coo(elements, gid, Ke):
    npe = elements.shape[1] - 1
    g = gid[elements[:, :npe]]
    rows = np.repeat(g[:, :, None], npe, axis=2).ravel()
    cols = np.repeat(g[:, None, :], npe, axis=1).ravel()
    return rows.astype(np.int64), cols.astype(np.int64), Ke.ravel().copy()


def anchor_A(rows, cols, vals, anchors):
    """impose_anchor_node_dof0_coo: zero row/col entries, append 1 on diag."""
    vals = vals.copy()
    for a in anchors:
        vals[(rows == a) | (cols == a)] = 0.0
        rows = np.append(rows, a)
        cols = np.append(cols, a)
        vals = np.append(vals, 1.0)
    return rows, cols, vals


def to_csr(rows, cols, vals, shape):
    """final_assembly.py:176 / :200 -- SciPy COO -> CSR (duplicates summed,
    explicit zeros kept)."""
    return coo_matrix((vals, (rows, cols)), shape=shape).tocsr()


def enforce_A(A, ddofs):
    """final_assembly.py:59-112: drop defect rows/cols, identity on them."""
    if ddofs is None or len(ddofs) == 0:
        return A
    c = A.tocoo()
    isdef = np.zeros(A.shape[0], dtype=bool)
    isdef[ddofs] = True
    keep = ~(isdef[c.row] | isdef[c.col])
    r = np.concatenate([c.row[keep], ddofs])
    cc = np.concatenate([c.col[keep], ddofs])
    v = np.concatenate([c.data[keep], np.ones(len(ddofs))])
    return coo_matrix((v, (r, cc)), shape=A.shape).tocsr()


def enforce_F(F, ddofs):
    """final_assembly.py:115-158: drop defect rows."""
    if ddofs is None or len(ddofs) == 0:
        return F
    c = F.tocoo()
    isdef = np.zeros(F.shape[0], dtype=bool)
    isdef[ddofs] = True
    keep = ~isdef[c.row]
    return coo_matrix((c.data[keep], (c.row[keep], c.col[keep])), shape=F.shape).tocsr()


def build_operators(coords, elements, gid, p):
    """All CSR operators of SimulationOperators (assemble_simulation_operators.py:42-107)."""
    n = int(gid.max()) + 1
    dd = None
    from .mesh import defect_dofs
    dd = defect_dofs(elements, gid)
    K, M, F = element_matrices(coords, elements)
    r, c, vK = scalar_coo(elements, gid, K)
    _, _, vM = scalar_coo(elements, gid, M)
    ops = {"n": n, "DefDOF": dd}
    ra, ca, va = anchor_A(r, c, vK, [0])
    ops["A_demag"] = to_csr(ra, ca, va, (n, n))
    for k, name in enumerate(("Fx", "Fy", "Fz")):
        _, _, vF = scalar_coo(elements, gid, F[k])
        ops[name] = to_csr(r, c, vF, (n, n))
    ops["A1"] = enforce_A(to_csr(r, c, vK * p["alpha1"] + vM, (n, n)), dd)
    ops["A2"] = enforce_A(to_csr(r, c, vK * p["alpha2"] + vM, (n, n)), dd)
    ops["F1"] = enforce_F(to_csr(r, c, vM, (n, n)), dd)
    if p["ME"]:
        from .params import rot111_M
        C = voigt_C(p["c11"], p["c12"], p["c44"], p["rot111"], rot111_M())
        Ke, Fe = elastic_element_matrices(coords, elements, C)
        npe = elements.shape[1] - 1
        g = gid[elements[:, :npe]].astype(np.int64)
        # rows gid_a + da*n, cols gid_b + db*n  (assemble_elasticity.py:389-398)
        rowd = (g[:, :, None] + n * np.arange(3)[None, None, :]).reshape(-1, 3 * npe)
        rows = np.repeat(rowd[:, :, None], 3 * npe, axis=2).ravel()
        cols = np.repeat(rowd[:, None, :], 3 * npe, axis=1).ravel()
        ra, ca, va = anchor_A(rows, cols, Ke.ravel().copy(), [0, n, 2 * n])
        ops["A_el"] = to_csr(ra, ca, va, (3 * n, 3 * n))
        # cols 6*gid_b + vb (assemble_elasticity.py:444-454)
        cold = (6 * g[:, :, None] + np.arange(6)[None, None, :]).reshape(-1, 6 * npe)
        rows = np.repeat(rowd[:, :, None], 6 * npe, axis=2).ravel()
        cols = np.repeat(cold[:, None, :], 3 * npe, axis=1).ravel()
        vals = Fe.ravel().copy()
        vals[np.isin(rows, [0, n, 2 * n])] = 0.0
        ops["F_el"] = to_csr(rows, cols, vals, (3 * n, 6 * n))
    return ops


class VolumeAverage:
    """volume_average.py:69-216, Jacobians from coords[global_id[elements]]."""

    def __init__(self, coords, elements, gid):
        npe = elements.shape[1] - 1
        self.npe = npe
        self.corner = gid[elements[:, :npe]]
        if npe == 8:
            pts, w = hex_points()
            Nf, dNf = hex_N, hex_dN
        else:
            pts, w = tet_points()
            Nf, dNf = tet_N, (lambda r, s, t: TET_DN)
        cc = coords[self.corner]                       # the reference's indexing
        det = np.zeros((cc.shape[0], len(w)))
        for g, (r, s, t) in enumerate(pts):
            J = np.einsum("enk,nj->ekj", cc, dNf(r, s, t)[:npe])
            det[:, g] = np.abs(J[:, 0, 0] * (J[:, 1, 1] * J[:, 2, 2] - J[:, 1, 2] * J[:, 2, 1])
                               - J[:, 0, 1] * (J[:, 1, 0] * J[:, 2, 2] - J[:, 1, 2] * J[:, 2, 0])
                               + J[:, 0, 2] * (J[:, 1, 0] * J[:, 2, 1] - J[:, 1, 1] * J[:, 2, 0]))
        self.detJ = det
        self.Ng = np.array([Nf(*p)[:npe] for p in pts])
        self.W = w
        self.defect = elements[:, -1]

    def average(self, f, defect_flag=None):
        scalar = f.ndim == 1
        if scalar:
            f = f.reshape(-1, 1)
        cu = f[self.corner]
        mgp = np.einsum("ebk,gb->egk", cu, self.Ng)
        vw = self.detJ * self.W
        pie = np.sum(mgp * vw[:, :, None], axis=1)
        pv = np.sum(vw, axis=1)
        if defect_flag is not None:
            nd = 1.0 - defect_flag.astype(np.float64)
            pie = pie * nd[:, None]
            tv = np.sum(pv * nd)
        else:
            tv = np.sum(pv)
        ti = np.sum(pie, axis=0)
        if abs(tv) < 1e-6:
            tv = max(abs(tv), 1e-6)
        avg = ti / tv
        return avg[0] if scalar else avg
