"""Rotation of the crystal x axis to [111] (cupymag/utils/rot_111_matrices.py:21-102):
R (3x3, y = R x) and the Voigt strain map M (6x6)."""
import numpy as np

from ..core.constants import inv_sqrt2, inv_sqrt3, inv_sqrt6, sqrt2, sqrt3


def get_R_matrix(backend="numpy", dtype=None):
    if backend not in ("numpy", "cupy", "torch"):
        raise ValueError("Backend must be 'numpy' or 'cupy'")
    return np.array([[inv_sqrt3, inv_sqrt3, inv_sqrt3],
                     [-inv_sqrt2, inv_sqrt2, 0.0],
                     [-inv_sqrt6, -inv_sqrt6, 2 * inv_sqrt6]], dtype=dtype or np.float64)


def get_M_matrix(backend="numpy", dtype=None):
    if backend not in ("numpy", "cupy", "torch"):
        raise ValueError("Backend must be 'numpy' or 'cupy'")
    return np.array([
        [1 / 3, 1 / 3, 1 / 3, 2 / 3, 2 / 3, 2 / 3],
        [1 / 2, 1 / 2, 0.0, 0.0, 0.0, -1.0],
        [1 / 6, 1 / 6, 2 / 3, -2 / 3, -2 / 3, 1 / 3],
        [sqrt3 / 6, -sqrt3 / 6, 0.0, sqrt3 / 3, -sqrt3 / 3, 0.0],
        [-sqrt2 / 6, -sqrt2 / 6, sqrt2 / 3, sqrt2 / 6, sqrt2 / 6, -sqrt2 / 3],
        [-inv_sqrt6, inv_sqrt6, 0.0, inv_sqrt6, -inv_sqrt6, 0.0]], dtype=dtype or np.float64)
