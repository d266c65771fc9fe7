"""GPU parity of batches (loops_batch_*): many matrices, SpMV and SpMM mixed,
one persistent launch.  Each item must match the oracle (per-row / per-entry
1e-5 bar against FP64), be bit-identical to the same item in a batch of its
own ("alone and in the mix"), and, for merge-path items, bit-identical to
the single-matrix entry point with the same tile size.  Graph replays are
checked against eager launches."""
import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import synth  # noqa: E402
from edge_loops_amd.batch import Batch, BatchItem  # noqa: E402


def _mats():
    ms = []
    for seed in range(6):
        off, idx, val = random_csr(800 + 500 * seed, 1500 + 300 * seed, 10 + 7 * seed,
                                   seed=40 + seed, long_rows=((seed * 3, 1400),))
        ms.append((off, idx, val, 1500 + 300 * seed))
    B = synth.powerlaw_csr(20000, 25000, 1.8, 1, 4096, 150000, seed=12)
    ms.append((B.offsets, B.indices, B.values, B.cols))
    ms.append((np.zeros(101, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 9))
    return ms


MATS = _mats()
# (matrix, k, schedule, items_per_thread)
SPEC = [(0, 1, "merge_path_flat", 4), (1, 1, "merge_path_flat", 8), (2, 1, "merge_path_flat", 16),
        (3, 1, "thread_mapped", 0), (4, 2, "merge_path_flat", 0), (5, 5, "merge_path_flat", 4),
        (6, 1, "merge_path_flat", 8), (6, 16, "merge_path_flat", 16), (1, 33, "merge_path_flat", 8),
        (2, 64, "merge_path_flat", 0), (0, 100, "merge_path_flat", 16), (7, 1, "merge_path_flat", 8),
        (7, 3, "merge_path_flat", 0), (6, 1, "thread_mapped", 0), (3, 1, "group_mapped", 0),
        (6, 1, "group_mapped", 64), (1, 1, "group_mapped", 2), (7, 1, "group_mapped", 0),
        (4, 1, "group_mapped", 16), (6, 32, "merge_path_flat", 4), (5, 300, "merge_path_flat", 8),
        (2, 65, "merge_path_flat", 32), (0, 129, "merge_path_flat", 0), (3, 40, "merge_path_flat", 32)]


def _items(vt="f32", seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for m, k, sched, ipt in SPEC:
        off, idx, val, cols = MATS[m]
        A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type=vt)
        if k == 1:
            xh = rng.uniform(-1, 1, cols).astype(np.float32)
            y = torch.full((A.rows,), float("nan"), device="cuda")
        else:
            xh = rng.uniform(-1, 1, (cols, k)).astype(np.float32)
            y = torch.full((A.rows, k), float("nan"), device="cuda")
        out.append((BatchItem(A, torch.from_numpy(xh).cuda(), y, sched, ipt), m, xh))
    return out


def _ref(m, xh, stored=None):
    off, idx, val, cols = MATS[m]
    v = val if stored is None else stored
    if xh.ndim == 1:
        return oracle.spmv_f64(off, idx, v, xh)
    return oracle.spmm_f64(off, idx, v, xh)


@pytest.mark.parametrize("vt", ["f32", "bf16", "f16"])
def test_batch_parity_alone_and_in_mix(vt):
    its = _items(vt)
    b = Batch([i for i, _, _ in its])
    b.launch()
    torch.cuda.synchronize()
    mixed = [i.y.cpu().numpy().copy() for i, _, _ in its]
    for (it, m, xh), ym in zip(its, mixed):
        stored = it.csr.values.float().cpu().numpy()
        r64, a64 = _ref(m, xh, stored)
        assert np.all(np.abs(ym - r64) <= 1e-5 * a64 + 1e-30), (m, xh.shape)
        it.y.fill_(float("nan"))
        alone = Batch([it])
        alone.launch()
        torch.cuda.synchronize()
        assert np.array_equal(it.y.cpu().numpy(), ym), (m, xh.shape)


def test_batch_matches_single_matrix_entry_points():
    its = _items()
    Batch([i for i, _, _ in its]).launch()
    torch.cuda.synchronize()
    for it, m, xh in its:
        if it.schedule != "merge_path_flat":
            continue
        y = torch.full_like(it.y, float("nan"))
        if xh.ndim == 1:
            loops.run_spmv("merge_path_flat", it.csr, it.x, y, threads_per_block=256,
                           items_per_thread=it.items_per_thread or 8)
        else:
            loops.run_spmm("merge_path_flat", it.csr, it.x, y,
                           items_per_thread=it.items_per_thread or 0)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), it.y.cpu().numpy()), (m, xh.shape)


def test_batch_graph_replay():
    its = _items(seed=3)
    b = Batch([i for i, _, _ in its])
    g = b.capture()
    for rep in range(3):
        rng = np.random.default_rng(100 + rep)
        xs = []
        for it, m, xh in its:
            new = rng.uniform(-1, 1, xh.shape).astype(np.float32)
            it.x.copy_(torch.from_numpy(new))
            it.y.fill_(float("nan"))
            xs.append(new)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        replayed = [it.y.cpu().numpy().copy() for it, _, _ in its]
        b.launch()
        torch.cuda.synchronize()
        for (it, m, _), yr, xh in zip(its, replayed, xs):
            assert np.array_equal(yr, it.y.cpu().numpy())
            r64, a64 = _ref(m, xh)
            assert np.all(np.abs(yr - r64) <= 1e-5 * a64 + 1e-30)


def test_batch_rejects_unaligned_and_bad_groups():
    off, idx, val, cols = MATS[0]
    idx_d = torch.from_numpy(np.concatenate([[0], idx]).astype(np.int32)).cuda()[1:]
    val_d = torch.from_numpy(np.concatenate([[0], val]).astype(np.float32)).cuda()[1:]
    A = loops.csr_t(off.shape[0] - 1, cols, torch.from_numpy(off).cuda(), idx_d, val_d)
    x = torch.zeros(cols, device="cuda")
    y = torch.zeros(A.rows, device="cuda")
    with pytest.raises(loops.LoopsError):
        Batch([BatchItem(A, x, y)])
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    for bad in (3, 128):
        with pytest.raises(loops.LoopsError):
            Batch([BatchItem(A, x, y, "group_mapped", bad)])
    with pytest.raises(loops.LoopsError):
        Batch([BatchItem(A, x, y, "work_oriented")])
