"""CG solver, field operators and the GSPM LLG step (oracle).

* ``cg``  -- ``cupymag/solvers/linear_solvers.py:20-74`` calling
  ``cupyx.scipy.sparse.linalg.cg`` (third-party, ``cupy>=10.0.0``, not in
  the reference tree).  CuPy's published CG is restated: x0 = 0 unless given;
  ``if ||b|| == 0: return b``; atol = tol*||b||; r = b - A x; then
  rho = r.r, p = r (first) or r + (rho/rho_prev) p, q = A p,
  alpha = rho/(p.q), x += alpha p, r -= alpha q, iters += 1, stop when
  ||r|| <= atol; info = iters if maxiter was reached without convergence.
  Unlike SciPy's cg it always does at least one iteration and checks after
  the update.
* ``build_E0``           -- ``cupymag/utils/final_assembly.py:208-266``.
* ``hd_from_U``, ``strain_from_u`` -- ``cupymag/utils/compute_derivatives.py:36-125``.
* ``step``               -- ``cupymag/core/Micromagnetics.py:109-307``.
* ``run``                -- the hysteresis control of ``Micromagnetics.py:299-337``.

Pinning (tests/test_oracle.py): ``cg`` against ``scipy.sparse.linalg.spsolve``
on small SPD systems and its iteration count against the Krylov dimension
on matrices with few distinct eigenvalues (CG terminates in k iterations
for k distinct eigenvalues); ``step`` against the exact uniform-field
(macrospin) reduction -- for uniform m, no defect and ME off, the FEM solves
collapse to (M+aK)^-1 M c = c and the step becomes a closed-form 3-vector
recurrence -- and the invariants |m| = 1 off the defect, m = 0 on it;
``strain_from_u`` on affine displacement fields (constant strain, exact
for P1/Q1 elements); ``build_E0`` against the tensor definition
3/2 lambda (m_i m_j - delta_ij/3).  The full-field trajectory and CuPy's
exact iteration counts on a defect mesh depend on cuSPARSE/cuBLAS
summation order and cannot be imported here: that part is parity unpinned.
"""
import numpy as np

from .params import rot111_R, rot111_M


class CGError(RuntimeError):
    pass


def cg(A, b, x0=None, tol=1e-7, maxiter=5000, M=None):
    """Return (x, info, iters).  M (cupyx cg's preconditioner, applied as
    z = M r): rho = r.z, p = z or z + (rho/rho1) p; M None is the identity."""
    n = A.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    bn = np.sqrt(np.dot(b, b))
    if bn == 0:
        return b.copy(), 0, 0
    atol = tol * float(bn)
    r = b - A @ x
    it = 0
    rho = 0.0
    p = None
    resid = np.inf
    while it < maxiter:
        z = r if M is None else M @ r
        rho1 = rho
        rho = np.dot(r, z)
        if it == 0:
            p = z.copy()
        else:
            p = z + (rho / rho1) * p
        q = A @ p
        alpha = rho / np.dot(p, q)
        x = x + alpha * p
        r = r - alpha * q
        it += 1
        resid = np.sqrt(np.dot(r, r))
        if resid <= atol:
            break
    info = 0
    if it == maxiter and not (resid <= atol):
        info = it
    return x, info, it


def solve_cg(A, b, x0=None, tol=1e-7, maxiter=5000, use_init=False, system=None, log=None):
    x, info, it = cg(A, b, x0 if use_init else None, tol, maxiter)
    if log is not None:
        log.append((system, it))
    if info != 0:
        rn = np.linalg.norm(b - A @ x)
        who = f"CG for {system}" if system is not None else "CG"
        raise CGError(f"Error! {who} did not converge. info={info}, residual={rn}.")
    return x


def build_E0(lam100, lam111, m):
    N = m.shape[0]
    E0 = np.zeros(6 * N)
    mx, my, mz = m[:, 0], m[:, 1], m[:, 2]
    t = 1.0 / 3.0
    E0[0::6] = 1.5 * lam100 * (mx * mx - t)
    E0[1::6] = 1.5 * lam100 * (my * my - t)
    E0[2::6] = 1.5 * lam100 * (mz * mz - t)
    E0[3::6] = 2.0 * (1.5 * lam111 * (mx * my))
    E0[4::6] = 2.0 * (1.5 * lam111 * (my * mz))
    E0[5::6] = 2.0 * (1.5 * lam111 * (mx * mz))
    return E0


def hd_from_U(ops, U):
    return -(ops["Fx"] @ U), -(ops["Fy"] @ U), -(ops["Fz"] @ U)


def strain_from_u(ops, n, u, R=None):
    Dx, Dy, Dz = ops["Fx"], ops["Fy"], ops["Fz"]
    ux, uy, uz = u[:n], u[n:2 * n], u[2 * n:]
    if R is None:
        E = [Dx @ ux, Dy @ uy, Dz @ uz, Dy @ ux + Dx @ uy, Dz @ uy + Dy @ uz,
             Dz @ ux + Dx @ uz]
    else:
        dxx, dyx, dzx = Dx @ ux, Dy @ ux, Dz @ ux
        dxy, dyy, dzy = Dx @ uy, Dy @ uy, Dz @ uy
        dxz, dyz, dzz = Dx @ uz, Dy @ uz, Dz @ uz
        E11 = R[0, 0] * dxx + R[1, 0] * dyx + R[2, 0] * dzx
        E22 = R[0, 1] * dxy + R[1, 1] * dyy + R[2, 1] * dzy
        E33 = R[0, 2] * dxz + R[1, 2] * dyz + R[2, 2] * dzz
        E12 = (R[0, 1] * dxx + R[1, 1] * dyx + R[2, 1] * dzx) + \
              (R[0, 0] * dxy + R[1, 0] * dyy + R[2, 0] * dzy)
        E23 = (R[0, 2] * dxy + R[1, 2] * dyy + R[2, 2] * dzy) + \
              (R[0, 1] * dxz + R[1, 1] * dyz + R[2, 1] * dzz)
        E13 = (R[0, 2] * dxx + R[1, 2] * dyx + R[2, 2] * dzx) + \
              (R[0, 0] * dxz + R[1, 0] * dyz + R[2, 0] * dzz)
        E = [E11, E22, E33, E12, E23, E13]
    return np.stack(E, axis=-1)


def _aniso(p, m, E):
    """First derivative of the cubic anisotropy + magnetoelastic energy,
    Micromagnetics.py:181-189 (without the field terms)."""
    K1, c1, c2, c3, l100 = p["K1"], p["elasC1"], p["elasC2"], p["elasC3"], p["lambda100"]
    m1, m2, m3 = m[:, 0], m[:, 1], m[:, 2]
    s1, s2, s3 = m1 * m1, m2 * m2, m3 * m3
    v1 = (-2.0 * K1 * m1 * (s2 + s3) - 2.0 * c1 * m1 * (s2 + s3)
          - 2.0 * c2 * (E[:, 0] - l100) * m1 - c3 * (E[:, 3] * m2 + E[:, 5] * m3))
    v2 = (-2.0 * K1 * m2 * (s3 + s1) - 2.0 * c1 * m2 * (s3 + s1)
          - 2.0 * c2 * (E[:, 1] - l100) * m2 - c3 * (E[:, 4] * m3 + E[:, 3] * m1))
    v3 = (-2.0 * K1 * m3 * (s1 + s2) - 2.0 * c1 * m3 * (s1 + s2)
          - 2.0 * c2 * (E[:, 2] - l100) * m3 - c3 * (E[:, 5] * m1 + E[:, 4] * m2))
    return v1, v2, v3


def effective_field(p, m, E, Hbar, Ht, H):
    """f_i = anisotropic/elastic part + 1/2 Hbar_i + 1/2 Htilde_i + Hext_i,
    Micromagnetics.py:175-211 (rot111 evaluates the first part in the [100]
    frame and rotates back with R^T)."""
    if not p["rot111"]:
        # same left-to-right order as the reference expression
        K1, c1, c2, c3, l100 = p["K1"], p["elasC1"], p["elasC2"], p["elasC3"], p["lambda100"]
        m1, m2, m3 = m[:, 0], m[:, 1], m[:, 2]
        s1, s2, s3 = m1 * m1, m2 * m2, m3 * m3
        f1 = (-2.0 * K1 * m1 * (s2 + s3) + 0.5 * Hbar[0] + 0.5 * Ht[0] + H[0]
              - 2.0 * c1 * m1 * (s2 + s3) - 2.0 * c2 * (E[:, 0] - l100) * m1
              - c3 * (E[:, 3] * m2 + E[:, 5] * m3))
        f2 = (-2.0 * K1 * m2 * (s3 + s1) + 0.5 * Hbar[1] + 0.5 * Ht[1] + H[1]
              - 2.0 * c1 * m2 * (s3 + s1) - 2.0 * c2 * (E[:, 1] - l100) * m2
              - c3 * (E[:, 4] * m3 + E[:, 3] * m1))
        f3 = (-2.0 * K1 * m3 * (s1 + s2) + 0.5 * Hbar[2] + 0.5 * Ht[2] + H[2]
              - 2.0 * c1 * m3 * (s1 + s2) - 2.0 * c2 * (E[:, 2] - l100) * m3
              - c3 * (E[:, 5] * m1 + E[:, 4] * m2))
        return [f1, f2, f3]
    else:
        R = rot111_R()
        v = np.stack(_aniso(p, m @ R, E), axis=-1) @ R.T
        return [v[:, i] + 0.5 * Hbar[i] + 0.5 * Ht[i] + H[i] for i in range(3)]


def step(state, ops, Avg, p, log=None):
    """One GSPM LLG step, Micromagnetics.py:110-307. Mutates ``state``:
    m, avg_m, use_init and the warm-start vectors."""
    m = state["m"]
    n = ops["n"]
    avg_m = state["avg_m"]
    tol, maxit, ui = p["tol"], p["maxiter"], state["use_init"]
    R = rot111_R() if p["rot111"] else None
    Mr = rot111_M() if p["rot111"] else None
    if p["rot111"]:
        m100 = m @ R
    if p["ME"]:
        if p["rot111"]:
            E0 = (build_E0(p["lambda100"], p["lambda111"], m100).reshape(n, 6) @ Mr.T).reshape(6 * n)
        else:
            E0 = build_E0(p["lambda100"], p["lambda111"], m)
        b_el = ops["F_el"] @ E0
        state["u"] = solve_cg(ops["A_el"], b_el, state.get("u"), tol, maxit, ui, "elasticity", log)
        E = strain_from_u(ops, n, state["u"], R)
        E = E - Avg.average(E)
        l100, l111 = p["lambda100"], p["lambda111"]
        Es = p["Ebar_sigma"]
        am = avg_m if not p["rot111"] else Avg.average(m100)
        Eb = np.array([1.5 * l100 * (am[0] ** 2 - 1 / 3) + Es[0],
                       1.5 * l100 * (am[1] ** 2 - 1 / 3) + Es[1],
                       1.5 * l100 * (am[2] ** 2 - 1 / 3) + Es[2],
                       3.0 * l111 * am[0] * am[1] + Es[3],
                       3.0 * l111 * am[1] * am[2] + Es[4],
                       3.0 * l111 * am[0] * am[2] + Es[5]])
        if p["rot111"]:
            Eb = Mr @ Eb
        E = E + Eb[None, :]
        if p["rot111"]:
            E = E @ np.linalg.inv(Mr).T
    else:
        E = np.zeros((n, 6))
    mt = m - avg_m
    b = -(ops["Fx"] @ mt[:, 0] + ops["Fy"] @ mt[:, 1] + ops["Fz"] @ mt[:, 2])
    b[0] = 0.0
    state["U"] = solve_cg(ops["A_demag"], b, state.get("U"), tol, maxit, ui, "demag", log)
    Ht = hd_from_U(ops, state["U"])
    Nd = p["N"]
    Hbar = (-Nd[0] * avg_m[0], -Nd[1] * avg_m[1], -Nd[2] * avg_m[2])
    H = p["H_now"]
    dt = p["dt"]
    A1, A2, F1 = ops["A1"], ops["A2"], ops["F1"]
    f1, f2, f3 = effective_field(p, m, E, Hbar, Ht, H)

    def gs(name, rhs, A):
        state[name] = solve_cg(A, F1 @ rhs, state.get(name), tol, maxit, ui,
                               "Gauss-Seidel " + name, log)
        return state[name]

    g1n = gs("g1n", f1 * dt + m[:, 0], A1)
    g2n = gs("g2n", f2 * dt + m[:, 1], A1)
    g3n = gs("g3n", f3 * dt + m[:, 2], A1)
    m1s = m[:, 0] + g2n * m[:, 2] - g3n * m[:, 1]
    g1s = gs("g1star", f1 * dt + m1s, A1)
    m2s = m[:, 1] + g3n * m1s - g1s * m[:, 2]
    g2s = gs("g2star", f2 * dt + m2s, A1)
    m3s = m[:, 2] + g1s * m2s - g2s * m1s
    dt2 = dt * p["alpha_damp"]
    ms_ = np.stack([m1s, m2s, m3s], axis=-1)
    h = effective_field(p, ms_, E, Hbar, Ht, H)
    mss = [gs(nm, h[i] * dt2 + ms_[:, i], A2)
           for i, nm in enumerate(("m1starstar", "m2starstar", "m3starstar"))]
    mag = np.sqrt(mss[0] ** 2 + mss[1] ** 2 + mss[2] ** 2)
    mnew = np.empty_like(m)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(3):
            mnew[:, i] = np.where(np.abs(mag) < 1e-14, 0.0, mss[i] / mag)
    mnew[ops["DefDOF"], :] = 0.0
    sumsq = float(np.sum((mnew - m) ** 2))
    state["use_init"] = sumsq < p["use_init_factor"] * (n * p["llg_accuracy_factor"])
    state["m"] = mnew
    state["avg_m"] = Avg.average(mnew)
    state["E"] = E
    state["Ht"] = Ht
    state["sumsq"] = sumsq
    return state


def init_state(ops, Avg, m0):
    m = np.array(m0, dtype=np.float64)
    m[ops["DefDOF"], :] = 0.0
    return {"m": m, "avg_m": Avg.average(m), "use_init": False}


def uniform_m(n, direction=(1.0, 0.0, 0.0)):
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    return np.tile(d, (n, 1))


def run(state, ops, Avg, p, max_steps, on_converge=None, log=None):
    """Hysteresis control (Micromagnetics.py:109,299-335) for at most
    ``max_steps`` steps. Returns the list of (Hext, avg_m) field points."""
    n = ops["n"]
    acc = n * p["llg_accuracy_factor"]
    llg_min = 10 if p["ME"] else 100
    H = list(p["H"])
    p = dict(p)
    p["H_now"] = H
    nstep = 0
    pts = []
    for _ in range(max_steps):
        if not (state["avg_m"] @ p["stop_direction"] > p["stop_value"]):
            break
        nstep += 1
        step(state, ops, Avg, p, log)
        if state["sumsq"] < acc and nstep > llg_min:
            pts.append((tuple(H), state["avg_m"].copy()))
            H = [H[i] - p["dH"][i] for i in range(3)]
            p["H_now"] = H
            nstep = 0
    state["H"] = H
    return pts
