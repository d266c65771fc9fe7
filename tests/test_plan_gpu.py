"""GPU parity of the merge-path SpMV plan (loops_spmv_plan_*): tile
coordinates computed once, and column panels (A split by column into P
panels, each panel's tiles on its own XCDs, partial results added in panel
order).  Bar: the per-row 1e-5 relative bound against the FP64 oracle and the
reference validator's 1e-2 count; repeated executions bit-identical."""
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
    B = synth.powerlaw_csr(60000, 80000, 1.8, 1, 20000, 600000, seed=9, force_first_max=True)
    m["powerlaw_60k"] = (B.offsets, B.indices, B.values, B.cols)
    return m


MATS = _mats()


def _check(y, off, idx, val, x):
    y32 = oracle.spmv_f32(off, idx, val, x)
    y64, a = oracle.spmv_f64(off, idx, val, x)
    assert oracle.count_abs_errors(y, y32) == 0
    bad, worst = oracle.count_rel_errors(y, y64, a, 1e-5)
    assert bad == 0, worst


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("panels", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("tpb,ipt", [(256, 8), (512, 4), (256, 16)])
def test_plan_parity(name, panels, tpb, ipt):
    off, idx, val, cols = MATS[name]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A, panels=panels, threads_per_block=tpb, items_per_thread=ipt)
    assert plan.panels == panels
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


def test_plan_outlives_source_with_panels():
    # With panels > 1 the plan holds its own copy of A.
    off, idx, val, cols = MATS["powerlaw_60k"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A, panels=4)
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
def test_plan_low_precision(vt):
    off, idx, val, cols = MATS["powerlaw_60k"]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type=vt)
    stored = A.values.float().cpu().numpy()
    x_np = np.random.default_rng(6).uniform(-1, 1, cols).astype(np.float32)
    for panels in (1, 4):
        plan = loops.spmv_plan(A, panels=panels)
        y = torch.empty(A.rows, device="cuda")
        plan.execute(torch.from_numpy(x_np).cuda(), y)
        torch.cuda.synchronize()
        y64, a = oracle.spmv_f64(off, idx, stored, x_np)
        bad, worst = oracle.count_rel_errors(y.cpu().numpy(), y64, a, 1e-5)
        assert bad == 0, (panels, worst)


def test_plan_c2_full_size_auto_panels():
    A_h = synth.c2_powerlaw_2m()
    x_np = synth.uniform(A_h.cols, seed=21)
    A = loops.csr_t.from_host(A_h.rows, A_h.cols, A_h.offsets, A_h.indices, A_h.values)
    plan = loops.spmv_plan(A, layout="panels")
    assert plan.panels == 4 and plan.layout == "panels"  # 8 MB of x, one panel per 2 MiB
    y = torch.empty(A_h.rows, device="cuda")
    plan.execute(torch.from_numpy(x_np).cuda(), y)
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), A_h.offsets, A_h.indices, A_h.values, x_np)


def test_plan_auto_layout():
    # AUTO: row blocks from 2^12 columns and 2^19 nonzeros, unless a panel
    # parameter is set; smaller matrices take column panels
    B = synth.powerlaw_csr(60000, 80000, 1.8, 1, 20000, 600000, seed=9, force_first_max=True)
    A = loops.csr_t.from_host(B.rows, B.cols, B.offsets, B.indices, B.values)
    for kw, want in (({}, "blocks"), ({"panels": 2}, "panels"),
                     ({"threads_per_block": 512}, "panels"), ({"layout": "panels"}, "panels")):
        plan = loops.spmv_plan(A, **kw)
        assert plan.layout == want, kw
        plan.destroy()
    off, idx, val, cols = MATS["skewed"]  # 3000 x 2500, under 2^12 columns
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    plan = loops.spmv_plan(A)
    assert plan.layout == "panels"
    plan.destroy()
