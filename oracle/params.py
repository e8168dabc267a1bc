"""Reduced-unit parameters of the GSPM step (oracle, test infrastructure).

Restates ``cupymag/core/parameters.py:60-224`` (YAML keys, defaults,
validation and SI -> reduced-unit normalisation), ``cupymag/core/constants.py``
and the two small matrix helpers ``cupymag/utils/rot_111_matrices.py:21-102``
and ``cupymag/utils/sigma_matrices.py:25-45``.

Pinning: the reduced constants are checked in ``tests/test_oracle.py``
against hand-computed values for the BASELINE material (psi0, a*, dt, alpha)
and the rotation matrices against orthogonality / Voigt-rotation identities
(R R^T = I, M(R) applied to a rotated strain equals the strain rotated as a
tensor).
"""
import math

import numpy as np

MU0 = 1.25663706127e-6      # constants.py:17
GAMMA = 1.76e11             # constants.py:20
SQRT2, SQRT3, SQRT6 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(6.0)


def rot111_R():
    """3x3 rotation [100] -> [111] (rot_111_matrices.py:53-57)."""
    i2, i3, i6 = 1.0 / SQRT2, 1.0 / SQRT3, 1.0 / SQRT6
    return np.array([[i3, i3, i3],
                     [-i2, i2, 0.0],
                     [-i6, -i6, 2.0 * i6]], dtype=np.float64)


def rot111_M():
    """6x6 Voigt strain rotation for the same frame (rot_111_matrices.py:94-101)."""
    s2, s3, i6 = SQRT2, SQRT3, 1.0 / SQRT6
    return np.array([
        [1 / 3, 1 / 3, 1 / 3, 2 / 3, 2 / 3, 2 / 3],
        [1 / 2, 1 / 2, 0.0, 0.0, 0.0, -1.0],
        [1 / 6, 1 / 6, 2 / 3, -2 / 3, -2 / 3, 1 / 3],
        [s3 / 6, -s3 / 6, 0.0, s3 / 3, -s3 / 3, 0.0],
        [-s2 / 6, -s2 / 6, s2 / 3, s2 / 6, s2 / 6, -s2 / 3],
        [-i6, i6, 0.0, i6, -i6, 0.0]], dtype=np.float64)


def _div(a, b, what):
    # parameters.py:39-42
    if abs(b) < 1e-30:
        raise ValueError(f"Division by zero in {what}: {b}")
    return a / b


def reduced(cfg):
    """Return a dict of reduced-unit parameters from a parsed YAML dict.

    Follows parameters.py:60-224 key by key (same defaults, same validation
    order, same arithmetic).
    """
    ph = cfg["physics"]
    ms = ph["magnetic"]["ms"]
    AA_raw = ph["magnetic"]["exchange_stiffness"]
    K1_raw = ph["magnetic"]["anisotropy_k1"]
    c11r, c12r, c44r = (ph["elastic"][k] for k in ("c11", "c12", "c44"))
    lam100 = ph["magnetostriction"]["lambda100"]
    lam111 = ph["magnetostriction"]["lambda111"]
    dm = ph["demagnetization"]
    fld = ph["external_field"]
    st = ph["external_stress"]
    sig = [st.get(k, 0.0) for k in ("sigma11", "sigma22", "sigma33",
                                      "sigma12", "sigma23", "sigma13")]
    sim = cfg["simulation"]
    llg = sim["llg"]
    sol = sim["solver"]
    timestep = llg["timestep"]
    alpha_damp = llg["alpha_damping"]
    ld = ph["characteristic_length"]
    # parameters.py:26-36
    if ms <= 0:
        raise ValueError("Saturation magnetization must be positive")
    if ld <= 0:
        raise ValueError("Characteristic length must be positive")
    if timestep <= 0:
        raise ValueError("Timestep must be positive")
    if AA_raw <= 0:
        raise ValueError("Exchange constant must be positive")
    if alpha_damp < 0:
        raise ValueError("Damping constant alpha must be non-negative")
    psi0 = MU0 * ms ** 2
    K1 = _div(K1_raw, psi0, "K1")
    AA = _div(AA_raw, psi0, "AA")
    c11 = _div(c11r, psi0, "c11")
    c12 = _div(c12r, psi0, "c12")
    c44 = _div(c44r, psi0, "c44")
    sig = [_div(s, psi0, "sigma") for s in sig]
    den = c11 ** 2 + c11 * c12 - 2 * c12 ** 2
    s11 = (c11 + c12) / den
    s12 = -c12 / den
    s44 = 1 / c44
    elasC1 = 2 * c44 * (1.5 * lam111 ** 2) - (c11 - c12) * (1.5 * lam100) ** 2
    elasC2 = -1.5 * lam100 * (c11 - c12)
    elasC3 = -6.0 * lam111 * c44
    dt = _div(timestep * (MU0 * GAMMA) * ms, 1 + alpha_damp ** 2, "dt")
    astar = _div(2 * AA, ld ** 2, "astar")
    alpha = astar * dt
    H = [_div(fld["hext1"], ms, "H"), _div(fld.get("hext2", 0.0), ms, "H"),
         _div(fld.get("hext3", 0.0), ms, "H")]
    dH = [_div(fld["dhext1"], ms, "H"), _div(fld.get("dhext2", 0.0), ms, "H"),
          _div(fld.get("dhext3", 0.0), ms, "H")]
    # stop condition, parameters.py:147-167
    sc = sim["stop_condition"]
    comp = sc.get("component", None)
    if comp is None:
        d = sc.get("direction", None)
        if d is None:
            sdir = np.array([1.0, 0.0, 0.0])
        else:
            sdir = np.asarray(d, dtype=np.float64)
            nrm = np.linalg.norm(sdir)
            if nrm > 0:
                sdir = sdir / nrm
            else:
                raise ValueError("Direction vector cannot be zero.")
    else:
        sdir = np.zeros(3)
        sdir[comp] = 1.0
    sval = sc.get("value", -0.95)
    if not (0 > sval > -1):
        sval = -0.95
    # E_bar from external stress, sigma_matrices.py:25-45
    S = np.zeros((6, 6))
    S[0, 0] = S[1, 1] = S[2, 2] = s11
    S[3, 3] = S[4, 4] = S[5, 5] = s44
    S[0, 1] = S[0, 2] = S[1, 0] = S[1, 2] = S[2, 0] = S[2, 1] = s12
    Ebar_sigma = S @ np.array(sig)
    ME = bool(ph["magnetoelastic_coupling"])
    p = dict(ms=ms, psi0=psi0, K1=K1, AA=AA, c11=c11, c12=c12, c44=c44,
             s11=s11, s12=s12, s44=s44, sigma=sig, Ebar_sigma=Ebar_sigma,
             lambda100=lam100, lambda111=lam111,
             elasC1=elasC1 if ME else 0.0, elasC2=elasC2 if ME else 0.0,
             elasC3=elasC3 if ME else 0.0,
             elasC_raw=(elasC1, elasC2, elasC3),
             N=(dm["n_x"], dm["n_y"], dm["n_z"]),
             H=H, dH=dH, rot111=bool(ph["rotation_111"]), ME=ME, ld=ld,
             dt=dt, astar=astar, alpha=alpha, alpha_damp=alpha_damp,
             alpha1=alpha, alpha2=alpha * alpha_damp,
             llg_accuracy_factor=llg["llg_accuracy_factor"],
             tol=sol["tolerance"], maxiter=sol["max_iterations"],
             use_init_factor=sol["initial_factor"],
             stop_direction=sdir, stop_value=sval)
    return p
