"""Host checks of bench.py's roofline inputs (no GPU): the algorithmic bytes
per launch of the roofline kernels at 128^3, against closed forms of the
periodic Kuhn stencil, and the HBM-traffic figures that bench.py reads from
the rocprofv3 summaries under profiles/."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

# 128^3 periodic Kuhn lattice: every node has 14 neighbours, so A_el holds
# 15 3x3 blocks per block row (the anchored rows keep the pattern)
NODES = 128 ** 3
ROWS = 3 * NODES
NNZ = 15 * 9 * NODES


def test_stencil_counts_match_the_profiled_operator():
    # the operator size quoted in the sweep summary
    text = open(os.path.join(ROOT, "profiles", "mg_sweep_A_el_128cubed_hbm_traffic.md")).read()
    assert "nnz(A_el) = 283,115,520" in text and "nodes = 2,097,152" in text
    assert NNZ == 283_115_520 and NODES == 2_097_152


@pytest.mark.parametrize("args,kw,per_node", [
    # level-0 FP16 Chebyshev step, half-stored: 7 FP16 blocks (18 B), FP32
    # diagonal block (36 B), 18 FP32 vector entries (72 B)
    (("mixed", True, 16), dict(half=True), 7 * 18 + 36 + 72),
    # FP64 Chebyshev step, every block stored: 14 FP64 blocks, the FP64
    # diagonal block, the stored D^-1 (9 values) and 18 vector entries
    (("fast64", True, 64), dict(), 14 * 72 + 72 + (9 + 18) * 8),
    # SELL products, half-stored: 7 blocks, FP64 diagonal block, p and q
    (("mixed", False, None), dict(half32=True), 7 * 36 + 72 + 2 * 3 * 4),
    (("fast64", False, None), dict(half64=True), 7 * 72 + 72 + 2 * 3 * 8),
    # indexed SELL form: 14 blocks with a 4-byte column index each
    (("fast64", False, None), dict(stencil=False), 14 * 76 + 72 + 2 * 3 * 8),
    # CSR FP64: values and column indices per nonzero, row pointers, p and q
    (("fp64", False, None), dict(), None),
])
def test_kernel_bytes_closed_forms(args, kw, per_node):
    _, _, b = bench.kernel_bytes(*args, NNZ, ROWS, **kw)
    if per_node is None:
        assert b == 12.0 * NNZ + 4.0 * (ROWS + 1) + 16.0 * ROWS
    else:
        assert b == per_node * NODES


def test_outer_product_bytes():
    # 7 FP32 blocks (36 B), FP64 compensated diagonal block, FP32 p and q
    assert bench.outer_bytes(NNZ, ROWS, half32=True) == (7 * 36 + 72 + 2 * 3 * 4) * NODES
    assert bench.outer_bytes(NNZ, ROWS, half32=False) == (14 * 36 + 72 + 2 * 3 * 4) * NODES


def _quoted_algorithmic(path):
    """The 'algorithmic bytes' column of a summary's one-row table."""
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and re.fullmatch(r"\d+", cells[4]):
            return int(cells[4]), int(cells[3])
    return None


@pytest.mark.parametrize("fn,roc,bytes_fn", [
    ("mg_sweep_A_el_128cubed_hbm_traffic.md", "k_mg_sweep<3, 1, float, 0, true, _Float16>",
     lambda: bench.kernel_bytes("mixed", True, 16, NNZ, ROWS, half=True)[2]),
    ("pcg_outer_product_A_el_128cubed_hbm_traffic.md", "k_pcg_spmv_rel<3, 1, true>",
     lambda: bench.outer_bytes(NNZ, ROWS, half32=True)),
])
def test_profiled_traffic_matches_its_summary(fn, roc, bytes_fn):
    path = os.path.join(ROOT, "profiles", fn)
    got = bench.profiled_traffic(roc, 128)
    assert got is not None and got["source"] == os.path.join("profiles", fn)
    alg, traffic = _quoted_algorithmic(path)
    # the summary's table quotes the bench's own algorithmic bytes and the
    # traffic line it reads
    assert alg == bytes_fn()
    assert traffic == got["bytes"]
    # measured traffic covers the algorithmic bytes (re-reads across XCDs
    # add to them; a figure below them would mean a stale or wrong summary)
    assert 1.0 <= traffic / alg <= 1.6
    assert bench.profiled_traffic(roc, 64) is None


@pytest.mark.parametrize("roc,half", [("k_sell_spmv<3, 1, float, true, false, false, true>", "half32"),
                                      ("k_sell_spmv<3, 1, double, true, false, false, true>", "half64")])
def test_profiled_traffic_of_the_sell_products(roc, half):
    base = "mixed" if half == "half32" else "fast64"
    name, rocname, alg = bench.kernel_bytes(base, False, None, NNZ, ROWS, **{half: True})
    assert rocname == roc
    got = bench.profiled_traffic(roc, 128)
    assert got is not None
    assert 1.0 <= got["bytes"] / alg <= 1.6
