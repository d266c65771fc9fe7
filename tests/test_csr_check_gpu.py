"""loops_csr_check and the check at the start of loops_spmv_plan_create.

Well-formed matrices (empty rows, repeated columns, descents across row
boundaries, a 200k-row power-law matrix) give no flags; a descent inside a
row, a column id outside [0, cols) and offsets that do not run from 0 to
nnz without decreasing are each flagged.  A plan of an unsorted matrix is
refused before anything is built (tested with the row-block layout, whose
build reads and writes only inside A and the plan on such input, so the
test stays in bounds even if the check let it through; the plan is never
executed)."""
import ctypes

import numpy as np
import pytest
import torch

import edge_loops_amd as loops
from conftest import random_csr
from edge_loops_amd import _capi, synth

pytestmark = pytest.mark.gpu


def _csr(offsets, indices, cols):
    off = np.asarray(offsets, np.int32)
    idx = np.asarray(indices, np.int32)
    return loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx,
                                 np.ones(idx.shape[0], np.float32))


CASES = [
    ("sample", [0, 0, 2, 3, 4], [0, 1, 2, 1], 4, []),
    ("repeats, empty rows, descents at row starts", [0, 3, 3, 3, 6], [1, 1, 2, 0, 0, 3], 4, []),
    ("one row", [0, 4], [0, 1, 1, 3], 4, []),
    ("descent inside a row", [0, 3, 5], [0, 2, 1, 0, 3], 4, ["unsorted"]),
    ("descent at the last two entries", [0, 2, 4], [0, 1, 3, 2], 4, ["unsorted"]),
    ("descent after an empty last row start", [0, 2, 2], [3, 1], 4, ["unsorted"]),
    ("column = cols", [0, 2, 3], [0, 4, 1], 4, ["column_range"]),
    ("negative column", [0, 1, 2], [-1, 0], 4, ["column_range"]),
    ("offsets decrease", [0, 3, 2, 4], [0, 1, 2, 3], 4, ["bad_offsets"]),
    ("offsets end short of nnz", [0, 1, 3], [0, 1, 2, 3], 4, ["bad_offsets"]),
    ("offsets start above 0", [1, 2, 4], [0, 1, 2, 3], 4, ["bad_offsets"]),
    ("unsorted and out of range", [0, 3], [2, 1, 9], 4, ["column_range", "unsorted"]),
]


@pytest.mark.parametrize("name,off,idx,cols,want", CASES, ids=[c[0] for c in CASES])
def test_csr_check_cases(name, off, idx, cols, want):
    assert loops.csr_check(_csr(off, idx, cols)) == want


def test_csr_check_large():
    A = synth.powerlaw_csr(200000, 150000, 1.8, 1, 20000, 2000000, seed=71)
    off, idx = A.offsets.copy(), A.indices.copy()
    assert loops.csr_check(loops.csr_t.from_host(A.rows, A.cols, off, idx, A.values)) == []
    # swap two entries of the longest row
    r = int(np.argmax(np.diff(off)))
    idx[off[r] + 5], idx[off[r] + 6] = idx[off[r] + 6], idx[off[r] + 5]
    assert loops.csr_check(loops.csr_t.from_host(A.rows, A.cols, off, idx, A.values)) == ["unsorted"]


def test_plan_refuses_unsorted():
    off, idx, val = random_csr(3000, 2500, 40, seed=72)
    r = int(np.argmax(np.diff(off)))
    idx = idx.copy()
    idx[off[r]], idx[off[r] + 1] = idx[off[r] + 1], idx[off[r]]
    A = loops.csr_t.from_host(3000, 2500, off, idx, np.ones_like(val))
    with pytest.raises(loops.LoopsError) as e:
        loops.spmv_plan(A, layout="blocks")
    assert e.value.status == _capi.ERROR_INVALID_VALUE
    # the per-call entry points do not need sorted rows
    x = torch.ones(2500, device="cuda")
    y = torch.empty(3000, device="cuda")
    loops.spmv.merge_path_flat(A, x, y)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), np.diff(off).astype(np.float32))


def test_csr_check_through_the_abi_rejects_null_flags():
    A = _csr([0, 1], [0], 1)
    v = A.view()
    assert _capi.lib().loops_csr_check(ctypes.byref(v), None, None) == _capi.ERROR_INVALID_VALUE
