"""GPU parity of the SpMV schedules against the CPU oracle, through the C ABI.

Bar (per row): |y_gpu - y32| <= 1e-5 * sum_j |a_ij x_j| against the
reference's fp32 validator (the literal bar), the same against the FP64
oracle, and the reference validator's own check (|y_gpu - y32| > 1e-2
counted as an error, examples/spmv/helpers.hxx:136-139) must report 0.
The FP64 bar is the stricter judge of accuracy: y32 itself carries the
sequential fp32 rounding error (up to 2.7e-7 of sum |a x| on the 2M-row
matrix), so meeting the bar against FP64 bounds the GPU's own error.
thread_mapped and original sum each row in CSR order like the validator and
must match the fp32 oracle bit for bit.  Partitions must be bit-exact.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import synth  # noqa: E402

SCHEDULES = ["merge_path_flat", "work_oriented", "group_mapped", "thread_mapped", "original"]
REL = 1e-5


def _dev(off, idx, val, cols, value_type="f32"):
    return loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type=value_type)


def _run(schedule, A, x_np, **tuning):
    x = torch.from_numpy(x_np).cuda()
    y = torch.full((A.rows,), float("nan"), dtype=torch.float32, device="cuda")
    loops.run_spmv(schedule, A, x, y, **tuning)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _check(y, off, idx, val, x, exact=False, tag=""):
    y32 = oracle.spmv_f32(off, idx, val, x)
    y64, a = oracle.spmv_f64(off, idx, val, x)
    assert oracle.count_abs_errors(y, y32) == 0, tag
    bad, worst = oracle.count_rel_errors(y, y64, a, REL)
    assert bad == 0, f"{tag}: {bad} rows beyond {REL} (worst {worst:.3g})"
    bad, worst = oracle.count_rel_errors(y, y32.astype(np.float64), a, REL)
    assert bad == 0, f"{tag}: {bad} rows beyond {REL} of the fp32 validator (worst {worst:.3g})"
    if exact:
        assert np.array_equal(y, y32), tag


def _matrices():
    mats = {}
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "sample_4x4.npz"))
    mats["sample4"] = (g["offsets"], g["indices"], g["values"], 4)
    r, c, off, idx, val = oracle.mtx_read(os.path.join(os.path.dirname(__file__), "golden",
                                                       "chesapeake.mtx"))
    mats["chesapeake"] = (off, idx, val, c)
    A = synth.c1_uniform_4k()
    mats["c1"] = (A.offsets, A.indices, A.values, A.cols)
    off, idx, val = random_csr(3000, 2500, 40, seed=1, long_rows=((7, 2500), (2999, 1700)))
    mats["skewed"] = (off, idx, val, 2500)
    off, idx, val = random_csr(700, 5000, 3, seed=2, empty_frac=0.7)
    mats["mostly_empty"] = (off, idx, val, 5000)
    mats["one_row"] = (np.array([0, 3000], np.int32), np.arange(3000, dtype=np.int32),
                       np.random.default_rng(3).uniform(-1, 1, 3000).astype(np.float32), 3000)
    mats["no_nnz"] = (np.zeros(129, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 7)
    # rows far longer than a tile, starting off a 4-element boundary
    lens = np.array([3, 9001, 1, 7003, 2, 0, 4099, 5], np.int64)
    off = np.zeros(lens.shape[0] + 1, np.int32)
    np.cumsum(lens, out=off[1:])
    rng = np.random.default_rng(4)
    idx = np.concatenate([np.sort(rng.choice(12000, L, replace=False)) for L in lens]).astype(np.int32)
    mats["long_misaligned"] = (off, idx, rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), 12000)
    B = synth.powerlaw_csr(60000, 80000, 1.8, 1, 20000, 600000, seed=9, force_first_max=True)
    mats["powerlaw_60k"] = (B.offsets, B.indices, B.values, B.cols)
    return mats


MATS = _matrices()


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("name", list(MATS))
def test_spmv_parity(schedule, name):
    off, idx, val, cols = MATS[name]
    x = np.random.default_rng(11).uniform(-1, 1, cols).astype(np.float32)
    if name == "chesapeake":
        x = np.ones(cols, np.float32)
    A = _dev(off, idx, val, cols)
    y = _run(schedule, A, x)
    _check(y, off, idx, val, x, exact=schedule in ("thread_mapped", "original"),
           tag=f"{schedule}/{name}")
    if name == "chesapeake":
        deg = np.load(os.path.join(os.path.dirname(__file__), "golden", "chesapeake_degrees.npy"))
        assert np.array_equal(y, deg)


def test_sample_exact_all_schedules():
    off, idx, val, cols = MATS["sample4"]
    x = np.array([1, 2, 3, 4], np.float32)
    A = _dev(off, idx, val, cols)
    for s in SCHEDULES:
        assert np.array_equal(_run(s, A, x), np.array([0, 21, 9, 12], np.float32)), s


def test_empty_matrix():
    A = loops.csr_t.from_host(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32),
                              np.zeros(0, np.float32))
    for s in SCHEDULES:
        loops.run_spmv(s, A, torch.zeros(0, device="cuda"), torch.zeros(0, device="cuda"))
    torch.cuda.synchronize()


@pytest.mark.parametrize("tpb,ipt", [(64, 3), (64, 8), (128, 4), (128, 5), (128, 8), (128, 12),
                                     (128, 16), (256, 2), (256, 3), (256, 4), (256, 5), (256, 6),
                                     (256, 8), (256, 10), (256, 12), (256, 16), (512, 2),
                                     (512, 4), (512, 6), (512, 8)])
@pytest.mark.parametrize("name", ["skewed", "long_misaligned"])
def test_merge_path_tunings(tpb, ipt, name):
    off, idx, val, cols = MATS[name]
    x = np.random.default_rng(5).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    y = _run("merge_path_flat", A, x, threads_per_block=tpb, items_per_thread=ipt)
    _check(y, off, idx, val, x, tag=f"mp{tpb}x{ipt}")
    # precomputed tile coordinates give the same bits as the in-kernel search
    coords = loops.merge_path_partition(A, tpb * ipt)
    y2 = _run("merge_path_flat", A, x, threads_per_block=tpb, items_per_thread=ipt,
              tile_coordinates=coords)
    assert np.array_equal(y, y2)


@pytest.mark.parametrize("tpb,tpt", [(64, 64), (128, 128), (256, 256), (256, 128), (128, 64),
                                     (256, 64),
                                     (256, 32), (256, 16), (256, 8), (256, 4)])
def test_group_mapped_tunings(tpb, tpt):
    off, idx, val, cols = MATS["skewed"]
    x = np.random.default_rng(6).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    y = _run("group_mapped", A, x, threads_per_block=tpb, threads_per_tile=tpt)
    _check(y, off, idx, val, x, tag=f"gm{tpb}/{tpt}")


@pytest.mark.parametrize("tpb,grid", [(64, 1), (128, 3), (128, 512), (256, 77)])
def test_work_oriented_tunings(tpb, grid):
    off, idx, val, cols = MATS["powerlaw_60k"]
    x = np.random.default_rng(7).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    y = _run("work_oriented", A, x, threads_per_block=tpb, grid_size=grid)
    _check(y, off, idx, val, x, tag=f"wo{tpb}x{grid}")


@pytest.mark.parametrize("tpb,grid", [(128, 512), (64, 1), (256, 77)])
@pytest.mark.parametrize("name", ["skewed", "powerlaw_60k", "mostly_empty", "c1"])
def test_work_oriented_unsplit_rows_bit_exact(name, tpb, grid):
    """A row that lies inside one thread's share is summed by that thread
    alone, front to back, so it equals the reference's sequential fp32
    validator loop bit for bit; rows cut by a share boundary go through the
    fix-up and meet the usual bar."""
    off, idx, val, cols = MATS[name]
    x = np.random.default_rng(17).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    y = _run("work_oriented", A, x, threads_per_block=tpb, grid_size=grid)
    xy = oracle.wo_coords(off, tpb * grid)
    split = np.zeros(off.shape[0] - 1, bool)
    sx, sy = xy[:, 0].astype(np.int64), xy[:, 1].astype(np.int64)
    inside = (sx < split.shape[0]) & (sy > off[np.minimum(sx, split.shape[0] - 1)])
    split[sx[inside]] = True
    y32 = oracle.spmv_f32(off, idx, val, x)
    wpt = -(-(off.shape[0] - 1 + int(off[-1])) // (tpb * grid))  # merge items per share
    if wpt >= 64:
        assert (~split).mean() > 0.5, (name, wpt)
    assert np.array_equal(y[~split].view(np.uint32), y32[~split].view(np.uint32)), name
    _check(y, off, idx, val, x, tag=f"wo{tpb}x{grid}")


@pytest.mark.parametrize("name", ["c1", "skewed", "powerlaw_60k", "one_row", "mostly_empty"])
def test_partitions_bit_exact(name):
    off, idx, val, cols = MATS[name]
    A = _dev(off, idx, val, cols)
    for ti in (640, 2048, 7):
        got = loops.merge_path_partition(A, ti).cpu().numpy().astype(np.uint32)
        assert np.array_equal(got, oracle.mp_tile_coords(off, ti)), ti
    for nt in (1, 1000, 65536):
        got = loops.work_oriented_partition(A, nt).cpu().numpy().astype(np.uint32)
        assert np.array_equal(got, oracle.wo_coords(off, nt)), nt


def test_caller_workspace_reuse():
    # One caller workspace, never initialised and dirtied by other schedules,
    # serves every schedule and matrix in turn.
    ws = torch.full((1 << 22,), 7, dtype=torch.uint8, device="cuda")
    for name in ("powerlaw_60k", "skewed", "c1", "one_row", "powerlaw_60k"):
        off, idx, val, cols = MATS[name]
        x = np.random.default_rng(14).uniform(-1, 1, cols).astype(np.float32)
        A = _dev(off, idx, val, cols)
        xt = torch.from_numpy(x).cuda()
        for s in ("work_oriented", "merge_path_flat", "work_oriented", "merge_path_flat"):
            y = torch.empty(A.rows, device="cuda")
            for _ in range(2):
                loops.run_spmv(s, A, xt, y, workspace=ws)
            torch.cuda.synchronize()
            _check(y.cpu().numpy(), off, idx, val, x, tag=f"{name}/{s}")


def test_deterministic_repeat():
    off, idx, val, cols = MATS["powerlaw_60k"]
    x = np.random.default_rng(8).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    for s in SCHEDULES:
        ys = [_run(s, A, x) for _ in range(3)]
        assert np.array_equal(ys[0], ys[1]) and np.array_equal(ys[1], ys[2]), s


def test_overwrites_y():
    # y is overwritten (not accumulated) by every schedule.
    off, idx, val, cols = MATS["c1"]
    x = np.random.default_rng(9).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    xt = torch.from_numpy(x).cuda()
    for s in SCHEDULES:
        y = torch.full((A.rows,), 123.0, device="cuda")
        loops.run_spmv(s, A, xt, y)
        loops.run_spmv(s, A, xt, y)
        torch.cuda.synchronize()
        _check(y.cpu().numpy(), off, idx, val, x, tag=s)


def test_unaligned_views():
    # Sub-array views whose pointers are not 16-byte aligned take the scalar
    # load path of the merge-path kernel.
    off, idx, val, cols = MATS["skewed"]
    x = np.random.default_rng(10).uniform(-1, 1, cols).astype(np.float32)
    idx_d = torch.from_numpy(np.concatenate([[0], idx]).astype(np.int32)).cuda()[1:]
    val_d = torch.from_numpy(np.concatenate([[0], val]).astype(np.float32)).cuda()[1:]
    A = loops.csr_t(off.shape[0] - 1, cols, torch.from_numpy(off).cuda(), idx_d, val_d)
    assert A.indices.data_ptr() % 16 != 0
    for s in SCHEDULES:
        _check(_run(s, A, x), off, idx, val, x, tag=s)


def test_reference_style_entry_points():
    off, idx, val, cols = MATS["c1"]
    x = np.random.default_rng(12).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols)
    xt = torch.from_numpy(x).cuda()
    for name in ("original", "thread_mapped", "group_mapped", "work_oriented"):
        y = loops.vector_t(A.rows)
        getattr(loops.spmv, name)(A, xt, y)
        _check(y.cpu().numpy(), off, idx, val, x, tag=name)
    y = loops.vector_t(A.rows)
    t = loops.spmv.merge_path_flat(A, xt, y)
    assert t.milliseconds() > 0
    _check(y.cpu().numpy(), off, idx, val, x, tag="merge_path_flat")


@pytest.mark.parametrize("vt", ["bf16", "f16"])
def test_low_precision_storage(vt):
    # bf16 / fp16 value storage, fp32 accumulation.  Against the FP64 oracle
    # evaluated on the stored (rounded) values the 1e-5 row bar holds; the
    # storage rounding itself (2^-9 relative for bf16, 2^-12 for fp16 per
    # value) is reported against the fp32-valued matrix.
    off, idx, val, cols = MATS["powerlaw_60k"]
    x = np.random.default_rng(13).uniform(-1, 1, cols).astype(np.float32)
    A = _dev(off, idx, val, cols, value_type=vt)
    stored = A.values.float().cpu().numpy()
    u = 2.0 ** -8 if vt == "bf16" else 2.0 ** -11
    for s in SCHEDULES:
        y = _run(s, A, x)
        y64s, a_s = oracle.spmv_f64(off, idx, stored, x)
        bad, worst = oracle.count_rel_errors(y, y64s, a_s, REL)
        assert bad == 0, (s, worst)
        y64, a = oracle.spmv_f64(off, idx, val, x)
        bad, worst = oracle.count_rel_errors(y, y64, a, u)
        assert bad == 0, (s, worst)


def test_c2_full_size_merge_path():
    """2M-row power-law matrix: parity against the FP64 oracle at full size,
    plus linearity A(x1 + x2) = A x1 + A x2 within the same bar."""
    A_h = synth.c2_powerlaw_2m()
    x = synth.uniform(A_h.cols, seed=21)
    A = _dev(A_h.offsets, A_h.indices, A_h.values, A_h.cols)
    y = _run("merge_path_flat", A, x)
    _check(y, A_h.offsets, A_h.indices, A_h.values, x, tag="c2")
    y2 = _run("merge_path_flat", A, x, threads_per_block=128, items_per_thread=5)
    _check(y2, A_h.offsets, A_h.indices, A_h.values, x, tag="c2/128x5")
    x1 = synth.uniform(A_h.cols, seed=22)
    ya, yb = _run("merge_path_flat", A, x1), _run("merge_path_flat", A, (x + x1).astype(np.float32))
    _, a1 = oracle.spmv_f64(A_h.offsets, A_h.indices, A_h.values, x1)
    _, a0 = oracle.spmv_f64(A_h.offsets, A_h.indices, A_h.values, x)
    assert np.all(np.abs((y.astype(np.float64) + ya) - yb) <= 4 * REL * (a0 + a1) + 1e-30)


def test_c2_full_size_all_schedules():
    """The 2M-row, 40M-nonzero power-law matrix through every schedule at
    its default tuning: the bars above on every row; thread_mapped and
    original bit-identical to the fp32 validator."""
    A_h = synth.c2_powerlaw_2m()
    x = synth.uniform(A_h.cols, seed=23)
    A = _dev(A_h.offsets, A_h.indices, A_h.values, A_h.cols)
    for s in SCHEDULES:
        y = _run(s, A, x)
        _check(y, A_h.offsets, A_h.indices, A_h.values, x,
               exact=s in ("thread_mapped", "original"), tag=f"c2/{s}")
