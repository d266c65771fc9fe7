"""SpMV at the largest matrices the C ABI accepts: rows + nnzs within a few
thousand of LOOPS_MAX_ROWS_PLUS_NNZS (2^31 - 1 - 2^20), about 2.1 billion
nonzeros (17 GB of CSR), generated on the device.

Entry j of row r sits in column (r + j) mod 2^20 with value 1, and
x[c] = c mod 2, so y[r] is the number of odd integers in [r, r + L_r):
(r + L_r) // 2 - r // 2, known exactly without an oracle run.  Two shapes:
2^20 - 1 rows of 2,046 entries (every partial sum an integer below 2^24,
so the GPU result must be exact), and 4 rows of 536,608,766 entries, each
spread over about 262,000 merge tiles and summed by the fix-up (held to the
1e-5 bar relative to sum |a x|, which is the count itself)."""
import ctypes

import numpy as np
import pytest
import torch

import edge_loops_amd as loops
from edge_loops_amd import _capi
from edge_loops_amd.algorithms import make_params

pytestmark = pytest.mark.gpu
LIMIT = 2**31 - 1 - 2**20
COLS = 2**20


def _matrix(rows, length):
    """rows x COLS CSR with `length` entries per row, built on the device in
    chunks of 2^28 entries."""
    nnz = rows * length
    assert rows + nnz <= LIMIT
    offsets = (torch.arange(rows + 1, dtype=torch.int64, device="cuda") * length).to(torch.int32)
    indices = torch.empty(nnz, dtype=torch.int32, device="cuda")
    step = 1 << 28
    for k0 in range(0, nnz, step):
        k = torch.arange(k0, min(nnz, k0 + step), dtype=torch.int64, device="cuda")
        r = k // length
        indices[k0:k0 + k.numel()] = ((r + (k - r * length)) % COLS).to(torch.int32)
        del k, r
    values = torch.ones(nnz, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    return loops.csr_t(rows, COLS, offsets, indices, values)


def _expected(rows, length):
    r = np.arange(rows, dtype=np.int64)
    return (r + length) // 2 - r // 2


@pytest.fixture(scope="module")
def x():
    return (torch.arange(COLS, device="cuda") % 2).to(torch.float32)


def test_max_size_many_rows(x):
    rows, length = 2**20 - 1, 2046
    print(f"max size: {rows} rows x {length}, rows + nnzs = {rows + rows * length}", flush=True)
    A = _matrix(rows, length)
    want = _expected(rows, length)
    for sched in ("merge_path_flat", "thread_mapped", "original"):
        y = torch.full((rows,), float("nan"), device="cuda")
        getattr(loops.spmv, sched)(A, x, y)
        torch.cuda.synchronize()
        got = y.cpu().numpy()
        assert np.array_equal(got, want.astype(np.float32)), (sched, np.flatnonzero(got != want)[:5])
    # the allocation-free entry point: the default tiles' fix-up slots of
    # this size fit its static array
    y = torch.full((rows,), float("nan"), device="cuda")
    v, p = A.view(), make_params()
    _capi.check(_capi.lib().loops_spmv_noalloc(
        _capi.handle(), _capi.SCHEDULES["merge_path_flat"], ctypes.byref(v), x.data_ptr(),
        y.data_ptr(), ctypes.byref(p), torch.cuda.current_stream().cuda_stream),
        "loops_spmv_noalloc")
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), want.astype(np.float32))
    del A
    torch.cuda.empty_cache()


def test_max_size_four_rows(x):
    rows = 4
    length = (LIMIT - rows) // rows
    print(f"max size: {rows} rows x {length}, rows + nnzs = {rows + rows * length}", flush=True)
    A = _matrix(rows, length)
    want = _expected(rows, length)
    y = torch.full((rows,), float("nan"), device="cuda")
    loops.spmv.merge_path_flat(A, x, y)
    torch.cuda.synchronize()
    got = y.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(got - want) <= 1e-5 * want), (got, want)
    del A
    torch.cuda.empty_cache()
