"""ctypes binding of libcupymag_hip.so (the C ABI declared in
include/cupymag_hip.h).

The library is built in-tree by ``cupymag_amd.build`` (``__graft_entry__``
calls it).  Loading fails loudly when the library is missing: there is no
CPU fallback for any device operation.
"""
import atexit
import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libcupymag_hip.so"

CPM_OK = 0
CPM_EINVAL, CPM_ENOMEM, CPM_EHIP, CPM_EINTERNAL, CPM_ENODEV = -1, -2, -3, -4, -5

# host-array ids of cpm_host_ops_array
HA = dict(PTR=0, IDX=1, ADEMAG=2, FX=3, FY=4, FZ=5, GS_PTR=6, GS_IDX=7, A1=8, A2=9,
          F1_PTR=10, F1_IDX=11, F1=12, EL_PTR=13, EL_IDX=14, AEL=15, FE_PTR=16, FE_IDX=17, FEL=18)
FIELD = dict(M=0, HTILDE=1, E=2, U=3, UEL=4)
MODE_FP64, MODE_FAST64, MODE_MIXED = 0, 1, 2
MODES = {"fp64": MODE_FP64, "fast64": MODE_FAST64, "mixed": MODE_MIXED}
BATCH_GRAPH = 1
PRECOND_EL_DEMAG = 1  # elasticity and demag solves: multigrid-preconditioned CG
PRECOND_GS = 2        # Gauss-Seidel solves on A1: multigrid-preconditioned CG
FMT = {0: "csr", 1: "sell64", 2: "sell32"}


class CpmError(RuntimeError):
    """A C-ABI call returned a negative status."""


class cpm_sim_ops(C.Structure):
    _fields_ = [("A_demag", C.c_void_p), ("F3", C.c_void_p), ("A1", C.c_void_p), ("A2", C.c_void_p),
                ("F1", C.c_void_p), ("A_el", C.c_void_p), ("F_el", C.c_void_p), ("mg_el", C.c_void_p),
                ("mg_demag", C.c_void_p), ("fel", C.c_void_p), ("mg_a1", C.c_void_p), ("fder", C.c_void_p)]


class cpm_sim_params(C.Structure):
    _fields_ = [("n", C.c_int32), ("me", C.c_int32), ("rot111", C.c_int32), ("maxiter", C.c_int32),
                ("tol", C.c_double), ("dt", C.c_double), ("alpha_damp", C.c_double), ("K1", C.c_double),
                ("elasC", C.c_double * 3), ("lambda100", C.c_double), ("lambda111", C.c_double),
                ("Nd", C.c_double * 3), ("Ebar_sigma", C.c_double * 6), ("R", C.c_double * 9),
                ("Mv", C.c_double * 36), ("MvinvT", C.c_double * 36), ("mode", C.c_int32),
                ("warm_start", C.c_int32), ("rel_delta", C.c_double), ("precond", C.c_int32),
                ("overlap", C.c_int32), ("reserved", C.c_int32 * 2), ("rel_delta_gs", C.c_double), ("rel_delta_mg", C.c_double)]


class cpm_step_report(C.Structure):
    _fields_ = [("sumsq", C.c_double), ("avg_m", C.c_double * 3), ("iters", C.c_int32 * 10),
                ("failed", C.c_int32), ("resid", C.c_double), ("point_done", C.c_int32), ("state", C.c_int32),
                ("point", C.c_int32), ("nstep", C.c_int32), ("use_init", C.c_int32), ("pad", C.c_int32),
                ("H", C.c_double * 3), ("H_next", C.c_double * 3)]


class cpm_sweep_cfg(C.Structure):
    _fields_ = [("H0", C.c_double * 3), ("dH", C.c_double * 3), ("llg_accuracy", C.c_double),
                ("initial_factor", C.c_double), ("stop_dir", C.c_double * 3), ("stop_value", C.c_double),
                ("llg_min_step", C.c_int32), ("steps_per_point", C.c_int32), ("npoints", C.c_int32),
                ("pad", C.c_int32)]


class cpm_schedule(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("grid_cap", C.c_int32), ("sell_grid_cap", C.c_int32),
                ("spmv_sell", C.c_int32)]


class cpm_tune_entry(C.Structure):
    _fields_ = [("format", C.c_int32), ("lanes", C.c_int32), ("grid_cap", C.c_int32), ("us", C.c_float)]


_P = C.c_void_p
_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double
_SIGS = {
    "cpm_last_error": (C.c_char_p, []),
    "cpm_abi_version": (C.c_int, []),
    "cpm_ctx_create": (C.c_int, [C.c_int, _P, C.POINTER(_P)]),
    "cpm_ctx_destroy": (C.c_int, [_P]),
    "cpm_ctx_synchronize": (C.c_int, [_P]),
    "cpm_mat_create_csr": (C.c_int, [_P, _I32, _I32, _I64, _P, _P, _P, C.c_int, C.POINTER(_P)]),
    "cpm_mat_destroy": (C.c_int, [_P]),
    "cpm_mat_shape": (C.c_int, [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I64)]),
    "cpm_spmv": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "cpm_cg": (C.c_int, [_P, _P, _P, _P, C.c_int, _D, C.c_int, C.POINTER(_I32), C.POINTER(_D)]),
    "cpm_cg_precond": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _D, C.c_int, C.POINTER(_I32), C.POINTER(_D)]),
    "cpm_cg_noalloc": (C.c_int, [_P, _P, _P, _P, C.c_int, _D, C.c_int, C.POINTER(_I32), C.POINTER(_D)]),
    "cpm_prof_enable": (C.c_int, [_P, C.c_int]),
    "cpm_prof_only": (C.c_int, [_P, _P]),
    "cpm_mat_prof": (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64), C.c_int]),
    "cpm_mat_prof_outer": (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64), C.c_int]),
    "cpm_fel_strain": (C.c_int, [_P, _P, _P, _P]),
    "cpm_fel_hd": (C.c_int, [_P, _P, _P]),
    "cpm_fel_dmul": (C.c_int, [_P, C.c_int, _P, _P]),
    "cpm_mat_add_sell": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "cpm_mat_sell_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(_I64)]),
    "cpm_cg_sell": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, _D, C.c_int, C.c_int, _D, C.POINTER(_I32),
                              C.POINTER(_D)]),
    "cpm_spmv_sell": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "cpm_hd_from_u": (C.c_int, [_P, _P, _P, _P]),
    "cpm_strain_from_u": (C.c_int, [_P, _P, _P, _P, _P]),
    "cpm_build_e0": (C.c_int, [_P, _I32, _D, _D, _P, _P, _P, _P]),
    "cpm_weighted_sum": (C.c_int, [_P, _I32, C.c_int, _P, _P, _P]),
    "cpm_sim_create": (C.c_int, [_P, C.POINTER(cpm_sim_ops), C.POINTER(cpm_sim_params), _P, _D, _P, C.POINTER(_P)]),
    "cpm_sim_destroy": (C.c_int, [_P]),
    "cpm_sim_set_m": (C.c_int, [_P, _P, _P]),
    "cpm_sim_get": (C.c_int, [_P, C.c_int, _P]),
    "cpm_sim_step": (C.c_int, [_P, _P, C.c_int, C.POINTER(cpm_step_report)]),
    "cpm_sim_control": (C.c_int, [_P, C.POINTER(cpm_sweep_cfg), _P]),
    "cpm_mat_autotune": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_int)]),
    "cpm_mat_set_schedule": (C.c_int, [_P, C.POINTER(cpm_schedule)]),
    "cpm_mat_get_schedule": (C.c_int, [_P, C.POINTER(cpm_schedule)]),
    "cpm_mg_create": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, C.c_int, _D, C.POINTER(_P)]),
    "cpm_mg_create_smoothed": (C.c_int, [_P, C.c_int, _P, C.c_int, C.c_int, _D, C.c_int, _D, C.POINTER(_P)]),
    "cpm_mg_set_coarse_degree": (C.c_int, [_P, C.c_int]),
    "cpm_mg_destroy": (C.c_int, [_P]),
    "cpm_mg_info": (C.c_int, [_P, C.POINTER(C.c_int), _P, C.c_int]),
    "cpm_mg_forms": (C.c_int, [_P, _P, C.c_int]),
    "cpm_mg_apply": (C.c_int, [_P, _P, _P]),
    "cpm_cg_mg": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, _D, C.c_int, _D, _P, _P]),
    "cpm_mg_set_graphs": (C.c_int, [_P, C.c_int]),
    "cpm_cg_mg_batch": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, _P, _D, C.c_int, _P, _P, _P]),
    "cpm_fel_create": (C.c_int, [_P, _I32, _P, _P, C.POINTER(_P)]),
    "cpm_fel_destroy": (C.c_int, [_P]),
    "cpm_fel_apply": (C.c_int, [_P, _P, _P]),
    "cpm_batch_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "cpm_batch_destroy": (C.c_int, [_P]),
    "cpm_batch_add": (C.c_int, [_P, _P]),
    "cpm_batch_remove": (C.c_int, [_P, _P]),
    "cpm_batch_size": (C.c_int, [_P, C.POINTER(_I32)]),
    "cpm_batch_step": (C.c_int, [_P, _P, _P, C.POINTER(cpm_step_report)]),
    "cpm_host_assemble": (C.c_int, [_P, _I64, _P, _I64, C.c_int, C.c_int, _P, _I32, _P, _P, C.c_int, C.c_int,
                                    C.POINTER(_P)]),
    "cpm_host_ops_array": (C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(_I64)]),
    "cpm_host_ops_free": (C.c_int, [_P]),
    "cpm_host_element_matrices": (C.c_int, [_P, _I64, _P, _I64, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
}

_lib = None
_finalizing = False


@atexit.register
def _at_exit():
    # Objects still alive when the interpreter finalizes (for instance kept
    # by a traceback) are not destroyed through the library: the device
    # runtime may already be shutting down, and process exit releases
    # everything anyway.
    global _finalizing
    _finalizing = True


def alive():
    """True while handles may still be destroyed through the library."""
    return _lib is not None and not _finalizing


def release(name, h):
    """Destroy a native handle from a ``__del__`` (``name``: the library's
    destroy function).  A no-op once the interpreter finalizes, including
    when this module's globals have already been cleared."""
    try:
        if alive():
            getattr(_lib, name)(h)
    except (AttributeError, TypeError):
        pass


def lib():
    """Load (once) and return the shared library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("CUPYMAG_HIP_LIB", LIB_PATH))
    if not path.exists():
        raise ImportError(f"{path} not found: build it with `python -m cupymag_amd.build` "
                          "(the HIP engine has no CPU fallback)")
    L = C.CDLL(str(path))
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().cpm_last_error().decode(errors="replace")


def check(rc, what=""):
    """Raise on negative status; return positive (CG info) values."""
    if rc < 0:
        kind = {CPM_EINVAL: ValueError, CPM_ENOMEM: MemoryError}.get(rc, CpmError)
        raise kind(f"{what}: {last_error()} (status {rc})" if what else f"{last_error()} (status {rc})")
    return rc


def ptr(a):
    """Raw address of a torch tensor or numpy array (contiguous)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
