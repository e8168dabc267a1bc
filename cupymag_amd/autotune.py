"""SpMV schedule autotuning by size class (cpm_mat_autotune).

``tune(A)`` times every candidate launch schedule of a device operator
(CSR lanes per row x workgroup cap, SELL workgroup cap, the format of the
standalone product) and keeps the fastest.  The result is stored per size
class — (format family, block size, index-free or not, rounded nonzeros per
row, rows in half-octaves) — so that operators of the same class (the A1 /
A2 / A_demag triple of a mesh, the meshes of a batch that fall into one
class) reuse it without timing again.
"""
import ctypes as C
import math
import threading

from . import _lib
from ._lib import check, lib

_cache = {}
# one tuning at a time: operators prepared on several host threads (batch
# and serving preparation) share the context stream, so concurrent timings
# would include each other's kernels; the lock also tunes a class once
_lock = threading.Lock()


def size_class(A):
    rows = max(1, A.shape[0])
    nnz_row = A.nnz / rows
    sell = getattr(A, "sell_bs", 0)
    stencil = A.sell_info()["stencil"] if sell else False
    return (sell, bool(stencil), int(round(nnz_row)), int(round(2 * math.log2(rows))), A.nsets > 1)


def get_schedule(A):
    s = _lib.cpm_schedule()
    check(lib().cpm_mat_get_schedule(A.h, C.byref(s)), "cpm_mat_get_schedule")
    return dict(lanes=s.lanes, grid_cap=s.grid_cap, sell_grid_cap=s.sell_grid_cap, spmv_sell=s.spmv_sell)


def set_schedule(A, sched):
    s = _lib.cpm_schedule(int(sched["lanes"]), int(sched["grid_cap"]),
                          int(sched["sell_grid_cap"]) or 8192, int(sched["spmv_sell"]))
    check(lib().cpm_mat_set_schedule(A.h, C.byref(s)), "cpm_mat_set_schedule")


def tune(A, use_cache=True):
    """Tune operator A (a DeviceCSR, with its SELL form if any).  Returns
    (schedule, timings); timings is None when the size class was already
    tuned and its schedule was applied."""
    key = size_class(A)
    with _lock:
        if use_cache and key in _cache:
            set_schedule(A, _cache[key])
            return _cache[key], None
        cap = 64
        arr = (_lib.cpm_tune_entry * cap)()
        n = C.c_int()
        A.ctx.sync()
        check(lib().cpm_mat_autotune(A.h, arr, cap, C.byref(n)), "cpm_mat_autotune")
        timings = [dict(format=_lib.FMT[e.format], lanes=e.lanes, grid_cap=e.grid_cap, us=float(e.us))
                   for e in arr[:min(n.value, cap)]]
        sched = get_schedule(A)
        _cache[key] = sched
        return sched, timings


def clear_cache():
    _cache.clear()
