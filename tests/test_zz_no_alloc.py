"""cpm_cg_noalloc (the CG entry point that allocates no device memory)
against cpm_cg, through the C ABI: the same solution bit for bit, the same
iteration count and residual, on two calls in a row (the second with a
smaller system, so the shared scratch holds the first call's data), with
and without a warm start; no change of the device's free memory over the
call; a system above the stated maximum is refused."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.cases import config  # noqa: E402


def _cg(fn, ctx, A, b, x0=None):
    from cupymag_amd._lib import lib, ptr
    x = x0.clone() if x0 is not None else torch.empty_like(b)
    it, res = C.c_int32(), C.c_double()
    rc = getattr(lib(), fn)(ctx.h, A.h, ptr(b), ptr(x), int(x0 is not None), 1e-10, 5000, C.byref(it),
                            C.byref(res))
    return rc, x, it.value, res.value


def test_noalloc_cg_equals_cg_on_two_calls():
    from cupymag_amd.core.assemble_simulation_operators import SimulationOperators
    from cupymag_amd.core.parameters import from_dict
    from cupymag_amd.device import context
    from cupymag_amd.mesh.kuhn import kuhn_mesh
    ctx = context(0)
    rng = np.random.default_rng(17)
    systems = []
    for n, which in ((16, "A_el"), (8, "A1")):   # the second call is smaller
        P = from_dict(config("Tet", n=n, me=True, defect_params=2.0))
        S = SimulationOperators(P, mesh_arrays=kuhn_mesh(n, jitter=0.1, seed=n), ctx=ctx)
        A = S.csr_elasticity()[0] if which == "A_el" else S.csr_GS()[0]
        b = torch.as_tensor(rng.standard_normal(A.shape[0]), device="cuda")
        if which == "A_el":
            b[[0, A.shape[0] // 3, 2 * A.shape[0] // 3]] = 0.0
        systems.append((S, A, b))
    for S, A, b in systems:
        for warm in (False, True):
            x0 = torch.as_tensor(rng.standard_normal(A.shape[0]), device="cuda") if warm else None
            rc0, x_ref, it0, res0 = _cg("cpm_cg", ctx, A, b, x0)
            torch.cuda.synchronize()
            free0, _ = torch.cuda.mem_get_info()
            rc1, x, it1, res1 = _cg("cpm_cg_noalloc", ctx, A, b, x0)
            torch.cuda.synchronize()
            free1, _ = torch.cuda.mem_get_info()
            assert rc0 == rc1 == 0
            assert it0 == it1 and res0 == res1
            assert torch.equal(x, x_ref)
            assert free1 == free0
            assert res1 <= 1e-10 * float(b.norm())   # CuPy's test, on the recurrence residual


def test_noalloc_cg_refuses_systems_above_its_maximum():
    from cupymag_amd._lib import lib
    from cupymag_amd.device import DeviceCSR, context
    ctx = context(0)
    n = (1 << 20) + 1
    A = DeviceCSR(np.arange(n + 1), np.arange(n, dtype=np.int32), np.ones(n), (n, n), ctx)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    rc, x, it, _ = _cg("cpm_cg_noalloc", ctx, A, b)
    assert rc < 0
    assert "exceed the maximum of 1048576" in lib().cpm_last_error().decode()
    rc, x, it, _ = _cg("cpm_cg", ctx, A, b)
    assert rc == 0 and it == 1 and float((x - 1.0).abs().max()) == 0.0
