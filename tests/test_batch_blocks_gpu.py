"""GPU parity of row-block plans inside batches (loops_batch_item_t.plan):
an item carrying a BLOCKS plan runs the plan's (block, segment) streams in
the batch's persistent kernel and its rows' segment partials in the batch's
combine kernel (plans cut for batches with one segment and no cut rows write
y from the streams and skip the combine).  Plans of both cuts are tested.

Bars: every plan item bit-identical to plan.execute on the same x (same
stream code, same summation order), in a batch of its own and mixed with
other plan items and CSR items; f32 items also meet the FP64 oracle's
per-row 1e-5 bar and the reference validator's 1e-2 count of 0; graph
replays equal eager launches; a plan twice in one batch, a panels plan, a
plan of another shape, a multi-column item with a plan and a destroyed plan
are rejected."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import random_csr
from test_blocks_gpu import MATS, _check

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import _capi, synth  # noqa: E402
from edge_loops_amd.batch import Batch, BatchItem  # noqa: E402


def _plan_items(names, vt="f32", seed=0, for_batches=True):
    rng = np.random.default_rng(seed)
    out = []
    for name in names:
        off, idx, val, cols = MATS[name]
        A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type=vt)
        plan = loops.spmv_plan(A, layout="blocks", for_batches=for_batches)
        xh = rng.uniform(-1, 1, cols).astype(np.float32)
        y = torch.full((A.rows,), float("nan"), device="cuda")
        out.append((BatchItem(A, torch.from_numpy(xh).cuda(), y, "blocks", plan=plan), name, xh))
    return out


def _csr_items(seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for i, (sched, ipt) in enumerate((("merge_path_flat", 8), ("thread_mapped", 0),
                                      ("merge_path_flat", 16))):
        off, idx, val = random_csr(2000 + 700 * i, 3000, 12, seed=60 + i)
        A = loops.csr_t.from_host(off.shape[0] - 1, 3000, off, idx, val)
        xh = rng.uniform(-1, 1, 3000).astype(np.float32)
        y = torch.full((A.rows,), float("nan"), device="cuda")
        out.append((BatchItem(A, torch.from_numpy(xh).cuda(), y, sched, ipt), (off, idx, val), xh))
    return out


def _executed(it):
    y = torch.full_like(it.y, float("nan"))
    it.plan.execute(it.x, y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("for_batches", [True, False], ids=["for_batches", "alone"])
def test_batch_plans_match_execute(for_batches):
    plans = _plan_items(list(MATS), for_batches=for_batches)
    csr = _csr_items()
    b = Batch([i for i, _, _ in plans] + [i for i, _, _ in csr])
    # streams; plans of several segments padded to multiples of 8
    want = sum(-(-p.blocks * p.panels // 8) * 8 if p.panels > 1 else p.blocks * p.panels
               for p in (i.plan for i, _, _ in plans) if p.blocks)
    assert b.info.blocks_work_items == want
    b.launch()
    torch.cuda.synchronize()
    mixed = [i.y.cpu().numpy().copy() for i, _, _ in plans]
    for (it, name, xh), ym in zip(plans, mixed):
        assert np.array_equal(ym, _executed(it)), name
        off, idx, val, _ = MATS[name]
        _check(ym, off, idx, val, xh)
        it.y.fill_(float("nan"))
        Batch([it]).launch()  # alone
        torch.cuda.synchronize()
        assert np.array_equal(it.y.cpu().numpy(), ym), name
    for it, (off, idx, val), xh in csr:
        r64, a64 = oracle.spmv_f64(off, idx, val, xh)
        assert np.all(np.abs(it.y.cpu().numpy() - r64) <= 1e-5 * a64 + 1e-30)


@pytest.mark.parametrize("for_batches", [True, False], ids=["for_batches", "alone"])
@pytest.mark.parametrize("vt", ["bf16", "f16"])
def test_batch_plans_low_precision(vt, for_batches):
    names = ["skewed", "powerlaw_60k", "mixed_lengths_8seg", "row_over_many_blocks"]
    plans = _plan_items(names, vt=vt, seed=5, for_batches=for_batches)
    Batch([i for i, _, _ in plans]).launch()
    torch.cuda.synchronize()
    for it, name, xh in plans:
        ym = it.y.cpu().numpy()
        assert np.array_equal(ym, _executed(it)), name
        off, idx, _, _ = MATS[name]
        stored = it.csr.values.float().cpu().numpy()
        r64, a64 = oracle.spmv_f64(off, idx, stored, xh)
        assert np.all(np.abs(ym - r64) <= 1e-5 * a64 + 1e-30), name


def test_batch_plans_graph_replay():
    plans = _plan_items(["skewed", "powerlaw_60k", "wide_5m_cols", "one_entry_rows",
                         "gapped_segment_rows", "no_nnz"], seed=7)
    its = [i for i, _, _ in plans] + [i for i, _, _ in _csr_items(seed=8)]
    b = Batch(its)
    g = b.capture()
    for rep in range(3):
        rng = np.random.default_rng(200 + rep)
        for it in its:
            it.x.copy_(torch.from_numpy(rng.uniform(-1, 1, it.x.shape[0]).astype(np.float32)))
            it.y.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        replayed = [it.y.cpu().numpy().copy() for it in its]
        b.launch()
        torch.cuda.synchronize()
        for it, yr in zip(its, replayed):
            assert np.array_equal(yr, it.y.cpu().numpy())
        for (it, name, _), yr in zip(plans, replayed):
            assert np.array_equal(yr, _executed(it)), name


def test_batch_of_batched_workload_matrices():
    """Matrices of the 512-matrix batched workload (square, rows
    log-uniform in [1e3, 2e5], Zipf(1.8) row lengths): the twelve largest
    as plan items, next to twelve others as merge-path items."""
    shapes = synth.c4_batch_shapes()
    order = np.argsort(shapes)
    big, small = order[-12:], order[:12]
    rng = np.random.default_rng(9)
    its = []
    for i in list(big) + list(small):
        m = synth.c4_batch_matrix(int(i), shapes[i])
        A = loops.csr_t.from_host(m.rows, m.cols, m.offsets, m.indices, m.values)
        x = torch.from_numpy(rng.uniform(-1, 1, m.cols).astype(np.float32)).cuda()
        y = torch.full((m.rows,), float("nan"), device="cuda")
        if i in big:
            plan = loops.spmv_plan(A, layout="blocks", for_batches=True)
            its.append((BatchItem(A, x, y, "blocks", plan=plan), m))
        else:
            its.append((BatchItem(A, x, y, "merge_path_flat", 8), m))
    Batch([it for it, _ in its]).launch()
    torch.cuda.synchronize()
    for it, m in its:
        ym = it.y.cpu().numpy()
        if it.plan is not None:
            assert np.array_equal(ym, _executed(it))
        r64, a64 = oracle.spmv_f64(m.offsets, m.indices, m.values, it.x.cpu().numpy())
        assert np.all(np.abs(ym - r64) <= 1e-5 * a64 + 1e-30)


def test_batch_plan_rejections():
    off, idx, val, cols = MATS["skewed"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    x = torch.zeros(cols, device="cuda")
    y = torch.zeros(A.rows, device="cuda")
    plan = loops.spmv_plan(A, layout="blocks")
    with pytest.raises(loops.LoopsError):  # the streams of one plan share its partials
        Batch([BatchItem(A, x, y, plan=plan), BatchItem(A, x, torch.zeros_like(y), plan=plan)])
    panels = loops.spmv_plan(A, layout="panels")
    with pytest.raises(ValueError):
        Batch([BatchItem(A, x, y, plan=panels)])
    # the C ABI refuses it too
    c = BatchItem(A, x, y, plan=plan).to_c()
    c.plan = panels._h.value
    arr = (_capi.BatchItem * 1)(c)
    info = _capi.BatchInfo()
    assert (_capi.lib().loops_batch_query(arr, 1, ctypes.byref(info))
            == _capi.ERROR_NOT_SUPPORTED)
    c.plan = plan._h.value
    c.A.cols = cols + 1
    arr = (_capi.BatchItem * 1)(c)
    assert (_capi.lib().loops_batch_query(arr, 1, ctypes.byref(info))
            == _capi.ERROR_INVALID_VALUE)
    off2, idx2, val2 = random_csr(50, cols, 4, seed=3)
    A2 = loops.csr_t.from_host(50, cols, off2, idx2, val2)
    with pytest.raises(ValueError):
        Batch([BatchItem(A2, x, torch.zeros(50, device="cuda"), plan=plan)])
    with pytest.raises(ValueError):
        Batch([BatchItem(A, torch.zeros(cols, 4, device="cuda"),
                         torch.zeros(A.rows, 4, device="cuda"), plan=plan)])
    with pytest.raises(ValueError):
        BatchItem(A, x, y, "blocks")
    item = BatchItem(A, x, y, plan=plan)
    plan.destroy()
    with pytest.raises(ValueError):
        Batch([item])
