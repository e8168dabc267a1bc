"""Precision selection (cupymag/utils/precision_select.py:16-53)."""
import numpy as np


def get_float_type(precision, backend="numpy"):
    if precision not in ("SP", "DP"):
        raise ValueError(f"Invalid precision '{precision}'. Choose 'SP' or 'DP'.")
    if backend == "numpy":
        return np.float32 if precision == "SP" else np.float64
    if backend == "torch":
        import torch
        return torch.float32 if precision == "SP" else torch.float64
    raise ValueError(f"Unsupported backend '{backend}'. Choose 'numpy' or 'torch'.")
