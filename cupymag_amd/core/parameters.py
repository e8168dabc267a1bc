"""Simulation parameters: YAML schema, defaults, validation and SI ->
reduced units (cupymag/core/parameters.py:26-224).

The reference evaluates all of this into module globals at import time from
``sys.argv[1]`` / ``$CUPYMAG_CONFIG_PATH``; here ``load_parameters`` returns
the same names as attributes of one object, so several simulations can
coexist in a process.  ``from_dict`` accepts an already parsed YAML mapping.
"""
import os
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np

from .config_loader import ConfigLoader
from .constants import GAMMA, MU0


def _validate_parameters(ms, ld, timestep, AA_raw, alpha_damp):
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


def _safe_divide(numerator, denominator, operation_name):
    if abs(denominator) < 1e-30:
        raise ValueError(f"Division by zero in {operation_name}: {denominator}")
    return numerator / denominator


def default_config_path():
    if len(sys.argv) > 1 and not sys.argv[1].startswith("-"):
        return Path(sys.argv[1]).resolve()
    return Path(os.environ.get("CUPYMAG_CONFIG_PATH", "config.yaml")).resolve()


def load_parameters(config_path=None):
    path = Path(config_path).resolve() if config_path is not None else default_config_path()
    loader = ConfigLoader(path)
    P = from_dict(loader.config, base_dir=path.parent)
    P.config_path = path
    P.output_dir = loader.get_output_dir()
    return P


def from_dict(config, base_dir=None):
    P = SimpleNamespace()
    P.config = config
    physics = config["physics"]
    P.ms = physics["magnetic"]["ms"]
    P.AA_raw = physics["magnetic"]["exchange_stiffness"]
    P.K1_raw = physics["magnetic"]["anisotropy_k1"]
    P.c11_raw = physics["elastic"]["c11"]
    P.c12_raw = physics["elastic"]["c12"]
    P.c44_raw = physics["elastic"]["c44"]
    P.lambda100 = physics["magnetostriction"]["lambda100"]
    P.lambda111 = physics["magnetostriction"]["lambda111"]
    demag = physics["demagnetization"]
    P.N_x, P.N_y, P.N_z = demag["n_x"], demag["n_y"], demag["n_z"]
    field = physics["external_field"]
    Hext1_raw, dHext1_raw = field["hext1"], field["dhext1"]
    Hext2_raw, dHext2_raw = field.get("hext2", 0.0), field.get("dhext2", 0.0)
    Hext3_raw, dHext3_raw = field.get("hext3", 0.0), field.get("dhext3", 0.0)
    st = physics["external_stress"]
    sig = [st.get(k, 0.0) for k in ("sigma11", "sigma22", "sigma33", "sigma12", "sigma23", "sigma13")]
    P.rot111 = physics["rotation_111"]
    P.ME = physics["magnetoelastic_coupling"]
    P.ld = physics["characteristic_length"]

    grid = config["grid"]
    P.grid_type = grid["type"]
    dc = grid.get("defect_center")
    P.defect_center = None if dc is None else np.asarray(dc, dtype=np.float64)
    if P.grid_type == "Hex":
        cm = grid["cubic_mesh"]
        P.Nx, P.Ny, P.Nz = cm.get("nx"), cm.get("ny"), cm.get("nz")
        P.Ndx, P.Ndy, P.Ndz = cm.get("ndx"), cm.get("ndy"), cm.get("ndz")
    elif P.grid_type == "Tet":
        tm = grid["tet_mesh"]
        P.mesh_file = tm["mesh_file"]
        if base_dir is not None and P.mesh_file is not None and not Path(P.mesh_file).is_absolute() \
                and not Path(P.mesh_file).exists():
            cand = Path(base_dir) / P.mesh_file
            if cand.exists():
                P.mesh_file = str(cand)
        P.defect_shape = tm["defect_shape"]
        P.shape_params = np.array(tm["defect_parameters"])
        rm = tm.get("rotation_matrix", None)
        if rm is not None:
            P.shape_params = (P.shape_params, np.array(rm))
    else:
        raise NotImplementedError(f"Grid type '{P.grid_type}' is not supported.")

    sim = config["simulation"]
    P.precision = sim["precision"]
    from ..utils.precision_select import get_float_type
    get_float_type(P.precision)  # the reference's ValueError for anything but SP / DP
    llg = sim["llg"]
    P.timestep = llg["timestep"]
    P.alpha_damp = llg["alpha_damping"]
    P.LLG_accuracy_factor = llg["llg_accuracy_factor"]
    solver = sim["solver"]
    P.tol = solver["tolerance"]
    P.maxiter = solver["max_iterations"]
    P.use_init_factor = solver["initial_factor"]
    P.restart = sim.get("restart", False)
    P.restart_m = sim.get("restart_magnetization", None)
    P.initial_magnetization = sim.get("initial_magnetization", [1.0, 0.0, 0.0]) if P.restart is False else None

    sc = sim["stop_condition"]
    comp = sc.get("component", None)
    if comp is None:
        d = sc.get("direction", None)
        if d is None:
            sd = np.array([1.0, 0.0, 0.0])
        else:
            sd = np.asarray(d, dtype=np.float64)
            nrm = np.linalg.norm(sd)
            if nrm > 0:
                sd = sd / nrm
            else:
                raise ValueError("Direction vector cannot be zero.")
    else:
        sd = np.zeros(3)
        sd[comp] = 1.0
    P.stop_direction = sd
    sv = sc.get("value", -0.95)
    P.stop_value = sv if (0 > sv > -1) else -0.95

    output = config.get("output", {}) or {}
    P.write_m = output.get("write_magnetization", True)
    P.save_frequency = int(output.get("save_frequency", 1000))
    P.vtu_alpha = output.get("vtu_blend_alpha", 0.05)
    P.output_dir = Path(base_dir or ".").resolve() if output.get("directory") is None else \
        (Path(output["directory"]) if Path(output["directory"]).is_absolute()
         else (Path(base_dir or ".") / output["directory"]).resolve())

    _validate_parameters(P.ms, P.ld, P.timestep, P.AA_raw, P.alpha_damp)
    psi0 = MU0 * P.ms ** 2
    P.psi0 = psi0
    P.K1 = _safe_divide(P.K1_raw, psi0, "K1 normalization")
    P.AA = _safe_divide(P.AA_raw, psi0, "AA normalization")
    P.c11 = _safe_divide(P.c11_raw, psi0, "c11 normalization")
    P.c12 = _safe_divide(P.c12_raw, psi0, "c12 normalization")
    P.c44 = _safe_divide(P.c44_raw, psi0, "c44 normalization")
    P.sigma = [_safe_divide(s, psi0, f"sigma normalization") for s in sig]
    den = P.c11 ** 2 + P.c11 * P.c12 - 2 * P.c12 ** 2
    P.s11 = (P.c11 + P.c12) / den
    P.s12 = -P.c12 / den
    P.s44 = 1 / P.c44
    P.elasC1 = 2 * P.c44 * (1.5 * P.lambda111 ** 2) - (P.c11 - P.c12) * (1.5 * P.lambda100) ** 2
    P.elasC2 = -1.5 * P.lambda100 * (P.c11 - P.c12)
    P.elasC3 = -6.0 * P.lambda111 * P.c44
    P.dt = _safe_divide(P.timestep * (MU0 * GAMMA) * P.ms, (1 + P.alpha_damp ** 2), "Time step dt")
    P.astar = _safe_divide(2 * P.AA, P.ld ** 2, "Exchange parameter")
    P.alpha = P.astar * P.dt
    P.Hext1 = _safe_divide(Hext1_raw, P.ms, "Hext1 normalization")
    P.dHext1 = _safe_divide(dHext1_raw, P.ms, "dHext1 normalization")
    P.Hext2 = _safe_divide(Hext2_raw, P.ms, "Hext2 normalization")
    P.dHext2 = _safe_divide(dHext2_raw, P.ms, "dHext2 normalization")
    P.Hext3 = _safe_divide(Hext3_raw, P.ms, "Hext3 normalization")
    P.dHext3 = _safe_divide(dHext3_raw, P.ms, "dHext3 normalization")
    return P
