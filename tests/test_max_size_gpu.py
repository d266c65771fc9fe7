"""SpMV at the largest matrices the C ABI accepts: rows + nnzs within a few
thousand of LOOPS_MAX_ROWS_PLUS_NNZS (2^31 - 1 - 2^20), about 2.1 billion
nonzeros (17 GB of CSR), generated on the device with sorted columns (the
plans need them; loops_csr_check confirms it), values 1 and x[c] = c mod 2,
so y is a count of odd columns known exactly without an oracle run.
Plans are refused at this size (LOOPS_MAX_PLAN_ROWS_PLUS_NNZS, checked
before any device work): building them here faulted the GPU twice with the
cause not found, so those runs are kept below as skipped tests.

Two shapes:
- 2^20 - 1 rows of 2,046 entries, row r in the band of columns
  [r', r' + 2046), r' = r mod (2^20 - 2046): every partial sum is an
  integer below 2^24, so every result must be exact.  The per-call
  schedules and the allocation-free entry point run.
- 4 rows of 536,608,766 entries, column (j * 2^20) // L for entry j (each
  column about 512 times): each row spans about 262,000 merge tiles and is
  summed by the fix-up; merge_path_flat runs (one lane per row would walk
  536M entries), held to the 1e-5 bar relative to sum |a x|, which is the
  count itself."""
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


def _matrix(rows, length, column):
    """rows x COLS CSR with `length` entries per row, entry j of row r in
    column(r, j) (int64 tensors), built on the device in chunks of 2^28
    entries and checked with loops_csr_check."""
    nnz = rows * length
    assert rows + nnz <= LIMIT
    offsets = (torch.arange(rows + 1, dtype=torch.int64, device="cuda") * length).to(torch.int32)
    indices = torch.empty(nnz, dtype=torch.int32, device="cuda")
    step = 1 << 28
    for k0 in range(0, nnz, step):
        k = torch.arange(k0, min(nnz, k0 + step), dtype=torch.int64, device="cuda")
        r = k // length
        indices[k0:k0 + k.numel()] = column(r, k - r * length).to(torch.int32)
        del k, r
    values = torch.ones(nnz, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    A = loops.csr_t(rows, COLS, offsets, indices, values)
    assert loops.csr_check(A) == []
    return A


PLAN_FAULT = ("building a plan of about 2^31 rows + nonzeros faulted the GPU twice, cause "
              "not found; plans are refused above LOOPS_MAX_PLAN_ROWS_PLUS_NNZS (2^26)")


def _plans_refused(A):
    for layout in ("blocks", "panels"):
        with pytest.raises(loops.LoopsError) as e:
            loops.spmv_plan(A, layout=layout)
        assert e.value.status == _capi.ERROR_NOT_SUPPORTED


def _run_plans(A, x, rows, check):
    for layout in ("blocks", "panels"):
        plan = loops.spmv_plan(A, layout=layout)
        y = torch.full((rows,), float("nan"), device="cuda")
        plan.execute(x, y)
        torch.cuda.synchronize()
        got = y.cpu().numpy()
        plan.destroy()
        check(layout, got)


@pytest.fixture(scope="module")
def x():
    return (torch.arange(COLS, device="cuda") % 2).to(torch.float32)


def test_max_size_many_rows(x):
    rows, length = 2**20 - 1, 2046
    print(f"max size: {rows} rows x {length}, rows + nnzs = {rows + rows * length}", flush=True)
    A = _matrix(rows, length, lambda r, j: r % (COLS - length) + j)
    r0 = np.arange(rows, dtype=np.int64) % (COLS - length)
    want = ((r0 + length) // 2 - r0 // 2).astype(np.float32)  # odd integers in [r0, r0 + length)

    def check(what, got):
        assert np.array_equal(got, want), (what, np.flatnonzero(got != want)[:5])

    for sched in ("merge_path_flat", "work_oriented", "group_mapped", "thread_mapped",
                  "original"):
        y = torch.full((rows,), float("nan"), device="cuda")
        getattr(loops.spmv, sched)(A, x, y)
        torch.cuda.synchronize()
        check(sched, y.cpu().numpy())
    # the allocation-free entry point: the default tiles' fix-up slots of
    # this size fit its static array
    y = torch.full((rows,), float("nan"), device="cuda")
    v, p = A.view(), make_params()
    _capi.check(_capi.lib().loops_spmv_noalloc(
        _capi.handle(), _capi.SCHEDULES["merge_path_flat"], ctypes.byref(v), x.data_ptr(),
        y.data_ptr(), ctypes.byref(p), torch.cuda.current_stream().cuda_stream),
        "loops_spmv_noalloc")
    torch.cuda.synchronize()
    check("noalloc", y.cpu().numpy())
    _plans_refused(A)
    del A
    torch.cuda.empty_cache()


def test_max_size_four_rows(x):
    rows = 4
    length = (LIMIT - rows) // rows
    print(f"max size: {rows} rows x {length}, rows + nnzs = {rows + rows * length}", flush=True)
    A = _matrix(rows, length, lambda r, j: (j * COLS) // length)
    # column c holds entries j in [ceil(c L / C), ceil((c + 1) L / C))
    c = np.arange(COLS + 1, dtype=np.int64)
    first = (c * length + COLS - 1) // COLS
    want = np.full(rows, np.diff(first)[1::2].sum(), dtype=np.float64)

    def check(what, got):
        got = got.astype(np.float64)
        assert np.all(np.abs(got - want) <= 1e-5 * want), (what, got, want)

    y = torch.full((rows,), float("nan"), device="cuda")
    loops.spmv.merge_path_flat(A, x, y)
    torch.cuda.synchronize()
    check("merge_path_flat", y.cpu().numpy())
    _plans_refused(A)
    del A
    torch.cuda.empty_cache()


def test_rows_past_2_30(x):
    """2^30 + 2^21 rows, 96 nonzeros in the last rows (row ids past 2^30, so
    the sum of two row ids exceeds int32 in a binary search over rows): the
    five per-call schedules and the allocation-free entry point give the
    exact counts, and loops_csr_check, whose search for the row of each
    column descent runs over all rows, finds the descents at row starts
    well formed and one inside a row unsorted."""
    rows = 2**30 + 2**21
    nzrows = rows - 300 + 3 * np.arange(48)  # every third of the last 300 rows
    lens = 1 + np.arange(48) % 3
    cols_of = [(k % 4) + 2 * np.arange(L) for k, L in enumerate(lens)]  # descents at row starts
    idx = np.concatenate(cols_of).astype(np.int32)
    nnz = idx.size
    cum = np.cumsum(lens)
    offsets = torch.zeros(rows + 1, dtype=torch.int32, device="cuda")
    for i, r in enumerate(nzrows):
        hi = nzrows[i + 1] + 1 if i + 1 < len(nzrows) else rows + 1
        offsets[int(r) + 1:int(hi)] = int(cum[i])
    assert int(offsets[-1]) == nnz
    print(f"rows past 2^30: {rows} rows, {nnz} nonzeros", flush=True)
    A = loops.csr_t(rows, COLS, offsets, torch.from_numpy(idx).cuda(),
                    torch.ones(nnz, dtype=torch.float32, device="cuda"))
    assert loops.csr_check(A) == []
    want = torch.zeros(rows, dtype=torch.float32, device="cuda")
    want[torch.from_numpy(nzrows).cuda()] = torch.tensor(
        [float((c % 2).sum()) for c in cols_of], device="cuda")
    y = torch.empty(rows, dtype=torch.float32, device="cuda")
    for sched in ("merge_path_flat", "work_oriented", "group_mapped", "thread_mapped",
                  "original"):
        y.fill_(float("nan"))
        getattr(loops.spmv, sched)(A, x, y)
        torch.cuda.synchronize()
        assert torch.equal(y, want), sched
    y.fill_(float("nan"))
    v, p = A.view(), make_params()
    _capi.check(_capi.lib().loops_spmv_noalloc(
        _capi.handle(), _capi.SCHEDULES["merge_path_flat"], ctypes.byref(v), x.data_ptr(),
        y.data_ptr(), ctypes.byref(p), torch.cuda.current_stream().cuda_stream),
        "loops_spmv_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(y, want), "noalloc"
    _plans_refused(A)
    # a descent inside the last row of three entries
    last3 = int(np.flatnonzero(lens == 3)[-1])
    bad = idx.copy()
    a = int(cum[last3]) - 3
    bad[a], bad[a + 1] = bad[a + 1], bad[a]
    B = loops.csr_t(rows, COLS, offsets, torch.from_numpy(bad).cuda(), A.values)
    assert loops.csr_check(B) == ["unsorted"]
    del A, B, offsets, want, y
    torch.cuda.empty_cache()


@pytest.mark.skip(reason=PLAN_FAULT)
def test_max_size_plans_many_rows(x):
    rows, length = 2**20 - 1, 2046
    A = _matrix(rows, length, lambda r, j: r % (COLS - length) + j)
    r0 = np.arange(rows, dtype=np.int64) % (COLS - length)
    want = ((r0 + length) // 2 - r0 // 2).astype(np.float32)

    def check(what, got):
        assert np.array_equal(got, want), (what, np.flatnonzero(got != want)[:5])

    _run_plans(A, x, rows, check)


@pytest.mark.skip(reason=PLAN_FAULT)
def test_max_size_plans_four_rows(x):
    rows = 4
    length = (LIMIT - rows) // rows
    A = _matrix(rows, length, lambda r, j: (j * COLS) // length)
    c = np.arange(COLS + 1, dtype=np.int64)
    first = (c * length + COLS - 1) // COLS
    want = np.full(rows, np.diff(first)[1::2].sum(), dtype=np.float64)

    def check(what, got):
        got = got.astype(np.float64)
        assert np.all(np.abs(got - want) <= 1e-5 * want), (what, got, want)

    _run_plans(A, x, rows, check)
