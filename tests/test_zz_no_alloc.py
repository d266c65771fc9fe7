"""loops_spmv_noalloc: the allocation-free SpMV entry point (fix-up slots in
a static device array of the library, one per GPU), called through the C
ABI.  Each result must equal loops_spmv's with the same schedule and tuning
bit for bit, on two calls in a row, the second with a smaller matrix (the
array keeps the first call's contents), and meet the 1e-5 bar against the
FP64 oracle."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import _capi, synth  # noqa: E402
from edge_loops_amd.algorithms import make_params  # noqa: E402

CASES = [("merge_path_flat", {}), ("merge_path_flat", {"threads_per_block": 128,
                                                       "items_per_thread": 5}),
         ("work_oriented", {}), ("work_oriented", {"threads_per_block": 256, "grid_size": 77}),
         ("group_mapped", {}), ("thread_mapped", {}), ("original", {})]


def _noalloc(schedule, A, x, y, **tuning):
    v = A.view()
    p = make_params(**tuning)
    st = _capi.lib().loops_spmv_noalloc(
        _capi.handle(), _capi.SCHEDULES[schedule], ctypes.byref(v), x.data_ptr(),
        y.data_ptr(), ctypes.byref(p), torch.cuda.current_stream().cuda_stream)
    _capi.check(st, "loops_spmv_noalloc")


def _matrices():
    big = synth.powerlaw_csr(300000, 250000, 1.8, 1, 60000, 3000000, seed=61,
                             force_first_max=True)
    off, idx, val = random_csr(3000, 2500, 40, seed=62, long_rows=((7, 2500),))
    return [(big.offsets, big.indices, big.values, big.cols), (off, idx, val, 2500)]


@pytest.mark.parametrize("schedule,tuning", CASES,
                         ids=[f"{s}-{'-'.join(map(str, t.values())) or 'default'}"
                              for s, t in CASES])
def test_noalloc_equals_loops_spmv(schedule, tuning):
    for call, (off, idx, val, cols) in enumerate(_matrices()):
        rows = off.shape[0] - 1
        A = loops.csr_t.from_host(rows, cols, off, idx, val)
        xh = np.random.default_rng(70 + call).uniform(-1, 1, cols).astype(np.float32)
        x = torch.from_numpy(xh).cuda()
        y_ref = torch.full((rows,), float("nan"), device="cuda")
        ws = torch.full((max(loops.spmv_workspace_size(schedule, A, **tuning), 4),), 0x5a,
                        dtype=torch.uint8, device="cuda")
        loops.run_spmv(schedule, A, x, y_ref, workspace=ws, **tuning)
        y = torch.full((rows,), float("nan"), device="cuda")
        _noalloc(schedule, A, x, y, **tuning)
        torch.cuda.synchronize()
        got, ref = y.cpu().numpy(), y_ref.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (schedule, call)
        y64, a = oracle.spmv_f64(off, idx, val, xh)
        bad, worst = oracle.count_rel_errors(got, y64, a, 1e-5)
        assert bad == 0, (schedule, call, worst)


def test_noalloc_rejects_beyond_its_scratch():
    # 64 x 3 merge tiles over 250M rows need 20.8 MB of fix-up slots, more
    # than the static array: rejected on the host before anything runs.
    v = _capi.CsrView(250_000_000, 4, 0, 1, None, None, 0)
    p = make_params(threads_per_block=64, items_per_thread=3)
    st = _capi.lib().loops_spmv_noalloc(None, _capi.SCHEDULES["merge_path_flat"],
                                        ctypes.byref(v), None, 1, ctypes.byref(p), None)
    assert st == _capi.ERROR_WORKSPACE
