"""Magnetization checkpoint and initial state
(cupymag/utils/magnetization_io.py:21-124).

``write_array`` / ``read_array`` keep the reference's HDF5 layout (dataset
``data``, gzip level 6, root attributes ``shape`` and ``dtype``): through
h5py when it is importable, otherwise through the module's own HDF5 writer
and reader (``h5lite``), so ``last_m.h5`` is an HDF5 file either way.
"""
from pathlib import Path

import numpy as np
import torch

from . import h5lite

try:  # optional dependency of the output format
    import h5py
except ImportError:  # pragma: no cover - depends on the image
    h5py = None


def write_array(array, filename):
    filename = Path(filename)
    if filename.suffix != ".h5":
        filename = filename.with_suffix(".h5")
    data = array.detach().cpu().numpy() if isinstance(array, torch.Tensor) else np.asarray(array)
    if h5py is not None:
        with h5py.File(filename, "w") as f:
            f.create_dataset("data", data=data, compression="gzip", compression_opts=6)
            f.attrs["shape"] = data.shape
            f.attrs["dtype"] = str(data.dtype)
    else:
        h5lite.write(filename, data, attrs={"shape": data.shape, "dtype": str(data.dtype)}, level=6)


def read_array(filename, nDOF):
    if filename is None:
        print("Initial magnetization file not found, start with uniform magnetization.")
        return None
    filename = Path(filename)
    if filename.suffix != ".h5":
        filename = filename.with_suffix(".h5")
    if h5py is not None:
        with h5py.File(filename, "r") as f:
            data = f["data"][:]
    else:
        data = h5lite.read(filename, "data")
    if data.shape != (nDOF, 3):
        print(f"Initial magnetization shape mismatch: expected ({nDOF}, 3), got {data.shape}.")
        print("Start with uniform magnetization.")
        return None
    return np.asarray(data, dtype=np.float64)


def initialize_m(restart_m, restart, initial_m, nDOF, DefDOF, float_dtype=np.float64):
    m = read_array(restart_m, nDOF) if restart else None
    if m is None:
        if initial_m is None:
            im = np.array([1.0, 0.0, 0.0])
        else:
            im = np.array(initial_m, dtype=np.float64)
            nrm = np.linalg.norm(im)
            if nrm > 0:
                im = im / nrm
            else:
                raise ValueError("Initial magnetization vector cannot be zero.")
        m = np.zeros((nDOF, 3), dtype=np.float64)
        m[:, 0], m[:, 1], m[:, 2] = im[0], im[1], im[2]
    m = np.array(m, dtype=np.float64)
    m[np.asarray(DefDOF, dtype=np.int64), :] = 0.0
    return m
