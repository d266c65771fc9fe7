"""GPU parity of the row-block SpMV plan (layout "blocks"): A copied into
blocks of 15872 rows x column segments (4, or a multiple of 8 for more than
2^22 columns), each (block, segment) stream bucketed by 64-column chunks
and cut into rounds of at most 4096 entries sorted by (row, column); y
accumulated in LDS one round at a time; the segments' partial rows (and a
split row's pieces) added in a fixed order by the combine kernel (y starts
as NaN, so a row nobody writes fails).

Bar: per row |y - y64| <= 1e-5 * sum_j |a_ij x_j| against the FP64 oracle and
the reference validator's count of |y - y32| > 1e-2 equal to 0
(examples/spmv/helpers.hxx:97-146); repeated executions of one plan
bit-identical."""
import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import synth  # noqa: E402


def _mats():
    m = {}
    off, idx, val = random_csr(3000, 2500, 40, seed=1, long_rows=((7, 2500), (2999, 1700)))
    m["skewed"] = (off, idx, val, 2500)
    lens = np.array([3, 9001, 1, 7003, 2, 0, 4099, 5], np.int64)
    off = np.zeros(lens.shape[0] + 1, np.int32)
    np.cumsum(lens, out=off[1:])
    rng = np.random.default_rng(4)
    idx = np.concatenate([np.sort(rng.choice(12000, L, replace=False)) for L in lens]).astype(np.int32)
    m["long_misaligned"] = (off, idx, rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), 12000)
    off, idx, val = random_csr(700, 5000, 3, seed=2, empty_frac=0.7)
    m["mostly_empty"] = (off, idx, val, 5000)
    m["no_nnz"] = (np.zeros(129, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 7)
    off, idx, val = random_csr(40, 5, 5, seed=3, empty_frac=0.1)
    m["five_cols"] = (off, idx, val, 5)
    B = synth.powerlaw_csr(60000, 80000, 1.8, 1, 20000, 600000, seed=9, force_first_max=True)
    m["powerlaw_60k"] = (B.offsets, B.indices, B.values, B.cols)
    # several row blocks, the last one partial; 8 column segments
    B = synth.powerlaw_csr(50000, 5_000_000, 2.0, 1, 30000, 500000, seed=11,
                           force_first_max=True)
    m["wide_5m_cols"] = (B.offsets, B.indices, B.values, B.cols)
    # 64-column chunks with far more than 4096 entries per block: split rounds
    off, idx, val = random_csr(20000, 40, 40, seed=13, empty_frac=0.05)
    m["dense_narrow"] = (off, idx, val, 40)
    # rows of 0 or 1 entries only: each row a stream row of one segment
    # (compressed stream rows)
    rng = np.random.default_rng(17)
    lens = rng.integers(0, 2, 30000)
    off = np.zeros(lens.shape[0] + 1, np.int32)
    np.cumsum(lens, out=off[1:])
    idx = rng.integers(0, 700000, int(lens.sum()))
    m["one_entry_rows"] = (off, idx.astype(np.int32),
                           rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), 700000)
    # one row holding most of the nonzeros: its pieces fill whole blocks and
    # span block boundaries
    off, idx, val = random_csr(30000, 600000, 4, seed=31, long_rows=((777, 400000),))
    m["row_over_many_blocks"] = (off, idx, val, 600000)
    # every row in segments 1-3 (dense stream rows), every 300th also in
    # segment 0: that segment's rounds skip 300 rows between slots, the
    # others step by one
    r = np.arange(124000)
    has0 = r % 300 == 0
    lens = 3 + has0.astype(np.int64)
    off = np.zeros(r.shape[0] + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    idx = np.empty(int(off[-1]), np.int32)
    at = off[:-1]
    idx[at[has0]] = (r % 100)[has0]
    at = at + has0
    for u, mul in enumerate((7, 11, 13)):
        idx[at + u] = 150000 * (u + 1) + (r * mul) % 1000
    m["gapped_segment_rows"] = (off.astype(np.int32), idx,
                                rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), 600000)
    # rows of one, none and several entries interleaved, 4 and 8 column
    # segments
    for cols, name in ((700000, "mixed_lengths_4seg"), (5_000_000, "mixed_lengths_8seg")):
        lens = np.tile(np.array([1, 2, 0, 1, 1, 5, 1, 17], np.int64), 4000)
        off = np.zeros(lens.shape[0] + 1, np.int32)
        np.cumsum(lens, out=off[1:])
        idx = np.concatenate([np.sort(rng.choice(cols, L, replace=False)) for L in lens])
        m[name] = (off, idx.astype(np.int32),
                   rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), cols)
    return m


MATS = _mats()


def _check(y, off, idx, val, x):
    y32 = oracle.spmv_f32(off, idx, val, x)
    y64, a = oracle.spmv_f64(off, idx, val, x)
    assert oracle.count_abs_errors(y, y32) == 0
    bad, worst = oracle.count_rel_errors(y, y64, a, 1e-5)
    assert bad == 0, worst


def segments(cols, for_batches=False):
    """Column segments of a blocks plan: 4 while 4 segments of 2^20 columns
    hold the columns, else 8; plans cut for batches take one up to 2^20."""
    if for_batches and cols <= 2 ** 20:
        return 1
    return 4 if cols <= 4 * 2 ** 20 else 8


@pytest.mark.parametrize("for_batches", [False, True], ids=["alone", "for_batches"])
@pytest.mark.parametrize("name", list(MATS))
def test_blocks_parity(name, for_batches):
    off, idx, val, cols = MATS[name]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A, layout="blocks", for_batches=for_batches)
    assert plan.layout == "blocks"
    assert plan.panels == segments(cols, for_batches)
    if for_batches:  # the fewest blocks the LDS holds, not enough to fill the GPU
        alone = loops.spmv_plan(A, layout="blocks")
        assert plan.blocks <= alone.blocks
        alone.destroy()
    x_np = np.random.default_rng(3).uniform(-1, 1, cols).astype(np.float32)
    x = torch.from_numpy(x_np).cuda()
    ys = []
    for _ in range(2):
        y = torch.full((A.rows,), float("nan"), device="cuda")
        plan.execute(x, y)
        torch.cuda.synchronize()
        ys.append(y.cpu().numpy())
    _check(ys[0], off, idx, val, x_np)
    assert np.array_equal(ys[0], ys[1])
    plan.destroy()


def test_blocks_outlive_source():
    off, idx, val, cols = MATS["powerlaw_60k"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A, layout="blocks")
    plan.csr = None
    A.indices.fill_(0)
    A.values.fill_(0)
    del A
    x_np = np.random.default_rng(5).uniform(-1, 1, cols).astype(np.float32)
    y = torch.empty(off.shape[0] - 1, device="cuda")
    plan.execute(torch.from_numpy(x_np).cuda(), y)
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), off, idx, val, x_np)


@pytest.mark.parametrize("vt", ["bf16", "f16"])
def test_blocks_low_precision(vt):
    off, idx, val, cols = MATS["powerlaw_60k"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type=vt)
    stored = A.values.float().cpu().numpy()
    x_np = np.random.default_rng(6).uniform(-1, 1, cols).astype(np.float32)
    plan = loops.spmv_plan(A, layout="blocks")
    y = torch.empty(A.rows, device="cuda")
    plan.execute(torch.from_numpy(x_np).cuda(), y)
    torch.cuda.synchronize()
    y64, a = oracle.spmv_f64(off, idx, stored, x_np)
    bad, worst = oracle.count_rel_errors(y.cpu().numpy(), y64, a, 1e-5)
    assert bad == 0, worst


def test_blocks_graph_replays():
    off, idx, val, cols = MATS["powerlaw_60k"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A, layout="blocks")
    x_np = np.random.default_rng(8).uniform(-1, 1, cols).astype(np.float32)
    x = torch.from_numpy(x_np).cuda()
    y = torch.empty(A.rows, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.execute(x, y, stream=s.cuda_stream)
    for _ in range(3):
        y.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _check(y.cpu().numpy(), off, idx, val, x_np)


def test_blocks_row_modes():
    # rows of 0 or 1 entries: a row lives in at most one of the 4 segments,
    # so stream rows are compressed; the narrow dense matrix's rows have
    # entries in every segment (dense)
    for name, comp in (("one_entry_rows", True), ("dense_narrow", False),
                       ("mixed_lengths_4seg", True), ("gapped_segment_rows", False)):
        off, idx, val, cols = MATS[name]
        A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
        plan = loops.spmv_plan(A, layout="blocks")
        assert plan.compressed == comp, name
        assert plan.blocks >= 1
        plan.destroy()
    off, idx, val, cols = MATS["long_misaligned"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A, layout="blocks")
    assert plan.split_rows == 3  # the rows of 9001, 7003 and 4099 entries (over 4096)
    plan.destroy()


def test_blocks_c2_full_size_auto():
    A_h = synth.c2_powerlaw_2m()
    x_np = synth.uniform(A_h.cols, seed=21)
    A = loops.csr_t.from_host(A_h.rows, A_h.cols, A_h.offsets, A_h.indices, A_h.values)
    plan = loops.spmv_plan(A)
    assert plan.layout == "blocks" and plan.panels == 4
    # one wave of compressed blocks on a 256-CU MI355X (two workgroups per
    # CU, four segments: at most 128 blocks)
    assert plan.compressed and plan.blocks <= 2 * torch.cuda.get_device_properties(0).multi_processor_count // 4
    y = torch.empty(A_h.rows, device="cuda")
    plan.execute(torch.from_numpy(x_np).cuda(), y)
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), A_h.offsets, A_h.indices, A_h.values, x_np)


@pytest.mark.parametrize("shift", [0, 1, 3])
def test_blocks_ragged_rows_and_unaligned_y(shift):
    # 1001 rows (the combine's last group of 4 rows is partial) with rows cut
    # into pieces in the middle and in that last group; y starting `shift`
    # floats past a 16-byte boundary (shift != 0: the combine's per-row path)
    off, idx, val = random_csr(1001, 30000, 12, seed=23, long_rows=((500, 9000), (1000, 5000)))
    A = loops.csr_t.from_host(off.shape[0] - 1, 30000, off, idx, val)
    plan = loops.spmv_plan(A, layout="blocks")
    x_np = np.random.default_rng(29).uniform(-1, 1, 30000).astype(np.float32)
    buf = torch.full((A.rows + 8,), float("nan"), device="cuda")
    y = buf[shift:shift + A.rows]
    plan.execute(torch.from_numpy(x_np).cuda(), y)
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), off, idx, val, x_np)
    assert torch.isnan(buf[:shift]).all() and torch.isnan(buf[shift + A.rows:]).all()
    plan.destroy()


@pytest.mark.parametrize("rows,cols,nnz,zipf", [
    (2_000, 2_000_000, 600_000, 1.8),     # few rows, very wide: rows cut into pieces
    (10_000, 1_000_000, 800_000, 1.8),
    (200_000, 5_000, 700_000, 1.8),       # tall and narrow: one segment of columns in use
    (300_000, 40, 600_000, 0.0),          # 40 columns, uniform rows of 2
    (150_000, 150_000, 900_000, 2.2),
    (3_000, 20_000_000, 600_000, 1.8),    # 24 column segments of 2^20 columns
])
def test_blocks_shapes_of_the_auto_rule(rows, cols, nnz, zipf):
    # shapes the AUTO layout sends to row blocks (>= 2^12 columns, >= 2^19
    # nonzeros) or near its edges, built explicitly as blocks
    lmax = min(rows // 4, cols, 50_000) if zipf else max(1, nnz // rows)
    B = synth.powerlaw_csr(rows, cols, zipf, 1, lmax, nnz, seed=rows % 97)
    A = loops.csr_t.from_host(B.rows, B.cols, B.offsets, B.indices, B.values)
    plan = loops.spmv_plan(A, layout="blocks")
    x_np = np.random.default_rng(31).uniform(-1, 1, cols).astype(np.float32)
    y = torch.full((B.rows,), float("nan"), device="cuda")
    plan.execute(torch.from_numpy(x_np).cuda(), y)
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), B.offsets, B.indices, B.values, x_np)
    plan.destroy()


@pytest.mark.parametrize("name", ["dense_narrow", "row_over_many_blocks", "mixed_lengths_8seg",
                                  "powerlaw_60k"])
def test_blocks_builds_are_identical(name):
    """Two builds of a matrix give the same plan: buckets of more than a
    round (dense_narrow's 64-column chunks, a long row's pieces) are placed
    in (row, column) order before they are cut into rounds, so every row's
    sum is grouped the same way; y is compared bit for bit over four
    builds."""
    off, idx, val, cols = MATS[name]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    x_np = np.random.default_rng(8).uniform(-1, 1, cols).astype(np.float32)
    x = torch.from_numpy(x_np).cuda()
    ys = []
    for _ in range(4):
        plan = loops.spmv_plan(A, layout="blocks")
        y = torch.full((A.rows,), float("nan"), device="cuda")
        plan.execute(x, y)
        torch.cuda.synchronize()
        ys.append(y.cpu().numpy())
        plan.destroy()
    for y in ys[1:]:
        assert np.array_equal(y.view(np.uint32), ys[0].view(np.uint32))
    _check(ys[0], off, idx, val, x_np)
