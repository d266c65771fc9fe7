"""GPU parity of SpMM C = A B through the C ABI.

thread_mapped sums every C(row, c) in CSR order, so it must equal the fp32
oracle (oracle_spmm_f32, the column-wise restatement of
algorithms/spmm/thread_mapped.cuh:41-52) bit for bit; group_mapped and
merge_path split long rows between lane ranges and are held to
|C - C64| <= 1e-5 * sum_j |a_ij b_jc| per entry against the FP64 oracle.  SpMM parity is pinned by the oracle's closed form
(column-wise SpMV); the reference has no SpMM validator.
"""
import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import synth  # noqa: E402

REL = 1e-5


def _mats():
    m = {}
    m["sample4"] = (np.array([0, 0, 2, 3, 4], np.int32), np.array([0, 1, 2, 1], np.int32),
                    np.array([5, 8, 3, 6], np.float32), 4)
    off, idx, val = random_csr(3000, 2500, 40, seed=1, long_rows=((7, 2500), (2999, 1700)))
    m["skewed"] = (off, idx, val, 2500)
    lens = np.array([3, 9001, 1, 7003, 2, 0, 4099, 5], np.int64)
    off = np.zeros(lens.shape[0] + 1, np.int32)
    np.cumsum(lens, out=off[1:])
    rng = np.random.default_rng(4)
    idx = np.concatenate([np.sort(rng.choice(12000, L, replace=False)) for L in lens]).astype(np.int32)
    m["long_misaligned"] = (off, idx, rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), 12000)
    B = synth.powerlaw_csr(20000, 30000, 2.0, 1, 8000, 200000, seed=9, force_first_max=True)
    m["powerlaw_20k"] = (B.offsets, B.indices, B.values, B.cols)
    m["no_nnz"] = (np.zeros(65, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 5)
    return m


MATS = _mats()


def _run(schedule, name, k, ld_pad=0, value_type="f32", **tuning):
    off, idx, val, cols = MATS[name]
    rows = off.shape[0] - 1
    A = loops.csr_t.from_host(rows, cols, off, idx, val, value_type=value_type)
    Bh = synth.dense(cols, k, seed=5 + k)
    Bfull = torch.zeros((cols, k + ld_pad), device="cuda")
    Bfull[:, :k] = torch.from_numpy(Bh).cuda()
    Cfull = torch.full((rows, k + ld_pad), float("nan"), device="cuda")
    B, C = Bfull[:, :k], Cfull[:, :k]
    loops.run_spmm(schedule, A, B, C, **tuning)
    torch.cuda.synchronize()
    if ld_pad:
        assert torch.isnan(Cfull[:, k:]).all()  # padding untouched
    return C.cpu().numpy(), Bh, A


@pytest.mark.parametrize("schedule", ["thread_mapped", "group_mapped", "merge_path"])
@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("k", [1, 3, 10, 16, 64, 100])
def test_spmm_parity(schedule, name, k):
    off, idx, val, cols = MATS[name]
    C, Bh, _ = _run(schedule, name, k)
    C32 = oracle.spmm_f32(off, idx, val, Bh)
    C64, A64 = oracle.spmm_f64(off, idx, val, Bh)
    assert np.all(np.abs(C - C64) <= REL * A64 + 1e-30), (schedule, name, k)
    if schedule == "thread_mapped":
        assert np.array_equal(C, C32), (schedule, name, k)


@pytest.mark.parametrize("schedule", ["thread_mapped", "group_mapped", "merge_path"])
@pytest.mark.parametrize("k,pad", [(8, 1), (32, 3), (64, 4), (130, 2)])
def test_spmm_leading_dimension(schedule, k, pad):
    off, idx, val, cols = MATS["skewed"]
    C, Bh, _ = _run(schedule, "skewed", k, ld_pad=pad)
    C64, A64 = oracle.spmm_f64(off, idx, val, Bh)
    assert np.all(np.abs(C - C64) <= REL * A64 + 1e-30)


@pytest.mark.parametrize("k", [1, 2, 3, 7, 16, 33, 64, 65, 129, 300])
@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_spmm_merge_path_tile_sizes(ipt, k):
    """Merge-path SpMM tiles of 4, 8 or 16 steps per range, lane groups of
    2 .. 64 lanes, 1 or 4 columns per lane, column blocks when k > 256."""
    off, idx, val, cols = MATS["powerlaw_20k"]
    C, Bh, _ = _run("merge_path", "powerlaw_20k", k, items_per_thread=ipt)
    C64, A64 = oracle.spmm_f64(off, idx, val, Bh)
    assert np.all(np.abs(C - C64) <= REL * A64 + 1e-30)


@pytest.mark.parametrize("k", [7, 33, 64, 300])
def test_spmm_merge_path_32_steps(k):
    """32 steps per range, for launches whose ranges hold 8 lanes or more
    (7 columns: 8 lanes x 1 column; 33 columns, B rows of 33 floats: 64
    lanes x 1 column; 64: 16 lanes x 4 columns; 300: four 64-column
    launches and one of 44, 16 lanes x 4 columns each)."""
    off, idx, val, cols = MATS["powerlaw_20k"]
    C, Bh, _ = _run("merge_path", "powerlaw_20k", k, items_per_thread=32)
    C64, A64 = oracle.spmm_f64(off, idx, val, Bh)
    assert np.all(np.abs(C - C64) <= REL * A64 + 1e-30)


def test_spmm_merge_path_rejects_other_tiles():
    with pytest.raises(loops.LoopsError):
        _run("merge_path", "powerlaw_20k", 8, items_per_thread=256)
    # 32 steps need ranges of 8 lanes: 2 columns (2 lanes x 1) and 16
    # columns (4 lanes x 4) are refused before any launch
    for k in (2, 16):
        with pytest.raises(loops.LoopsError) as e:
            _run("merge_path", "powerlaw_20k", k, items_per_thread=32)
        assert e.value.status == loops._capi.ERROR_NOT_SUPPORTED


def test_spmm_deterministic():
    for s in ("merge_path", "group_mapped"):
        runs = [_run(s, "powerlaw_20k", 32)[0] for _ in range(3)]
        assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[1], runs[2]), s


@pytest.mark.parametrize("tpb", [256, 512, 1024])
@pytest.mark.parametrize("k", [1, 5, 16, 64, 100])
@pytest.mark.parametrize("name", ["skewed", "long_misaligned", "no_nnz"])
def test_spmm_group_mapped_block_sizes(tpb, k, name):
    """block_mapped groups of 256, 512 or 1024 rows, their atoms shared by
    every lane range of the workgroup (rows longer than a share are summed
    in range order)."""
    off, idx, val, cols = MATS[name]
    C, Bh, _ = _run("group_mapped", name, k, threads_per_block=tpb)
    C64, A64 = oracle.spmm_f64(off, idx, val, Bh)
    assert np.all(np.abs(C - C64) <= REL * A64 + 1e-30), (tpb, k, name)


@pytest.mark.parametrize("vt", ["bf16", "f16"])
def test_spmm_low_precision_storage(vt):
    off, idx, val, cols = MATS["powerlaw_20k"]
    for s in ("thread_mapped", "group_mapped", "merge_path"):
        C, Bh, A = _run(s, "powerlaw_20k", 16, value_type=vt)
        stored = A.values.float().cpu().numpy()
        C64, A64 = oracle.spmm_f64(off, idx, stored, Bh)
        assert np.all(np.abs(C - C64) <= REL * A64 + 1e-30), s


@pytest.mark.parametrize("k", [8, 9, 13, 17, 30])
@pytest.mark.parametrize("schedule", ["merge_path", "group_mapped"])
def test_spmm_aligned_b_unaligned_c(schedule, k):
    """B rows 16-byte aligned (ldb = 32) and C rows of k floats: lanes take 4
    columns of B, and C (and the fix-up slots, k floats per tile) are
    written column by column where a row is not 16-byte aligned; the
    merge-path result equals the same product as a batch item bit for bit."""
    from edge_loops_amd.batch import Batch, BatchItem
    off, idx, val, cols = MATS["powerlaw_20k"]
    rows = off.shape[0] - 1
    A = loops.csr_t.from_host(rows, cols, off, idx, val)
    Bh = synth.dense(cols, 32, seed=40 + k)
    B = torch.from_numpy(Bh).cuda()[:, :k]
    C = torch.full((rows, k), float("nan"), device="cuda")
    loops.run_spmm(schedule, A, B, C)
    torch.cuda.synchronize()
    C64, A64 = oracle.spmm_f64(off, idx, val, np.ascontiguousarray(Bh[:, :k]))
    got = C.cpu().numpy()
    assert np.all(np.abs(got - C64) <= REL * A64 + 1e-30)
    if schedule == "merge_path":
        Cb = torch.full((rows, k), float("nan"), device="cuda")
        Batch([BatchItem(A, B, Cb, "merge_path_flat", 0)]).launch()
        torch.cuda.synchronize()
        assert np.array_equal(got.view(np.uint32), Cb.cpu().numpy().view(np.uint32))


@pytest.fixture(scope="module")
def c3():
    """The 1M x 1M Zipf(2.0) matrix with 16M nonzeros (longest row 200,000)
    and a dense 1M x 64 B, with the FP64 product of every row."""
    A_h = synth.c3_spmm_1m()
    B_h = synth.dense(A_h.cols, 64, seed=31)
    C64, A64 = oracle.spmm_f64(A_h.offsets, A_h.indices, A_h.values, B_h)
    return A_h, B_h, C64, A64


@pytest.mark.parametrize("schedule", ["merge_path", "group_mapped"])
def test_c3_full_size(c3, schedule):
    """Config-3 SpMM at full size: every entry of C within 1e-5 of
    sum_j |a_ij b_jc| of the FP64 product, the 200,000-nonzero row
    included."""
    A_h, B_h, C64, A64 = c3
    assert A_h.max_row() == 200000
    A = loops.csr_t.from_host(A_h.rows, A_h.cols, A_h.offsets, A_h.indices, A_h.values)
    B = torch.from_numpy(B_h).cuda()
    C = torch.full((A_h.rows, 64), float("nan"), device="cuda")
    loops.run_spmm(schedule, A, B, C)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    bad = np.abs(got - C64) > REL * A64 + 1e-30
    assert not bad.any(), (schedule, int(bad.sum()))
