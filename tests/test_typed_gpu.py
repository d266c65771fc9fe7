"""SpMV with double and int values (loops_spmv_typed): the reference's host
entries instantiated for type_t = double and int (its benchmark's value
types, benchmarks/spmv/work_oriented.cu:53), through the C ABI, the Python
mirror and the C++ drop-in CLI.

Bars: int results equal the oracle's wrap-around loop bit for bit for every
schedule (any summation order gives the same bits); double results of
thread_mapped and original equal the sequential double validator loop bit
for bit, and every schedule is within (n + 1) * 2^-52 * sum_j |a_ij x_j| of
the long-double reference per row (n: the row's nonzeros), the bound any
summation order of n products meets.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import synth  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edge_loops_amd", "bin")
SCHEDULES = ["merge_path_flat", "work_oriented", "group_mapped", "thread_mapped", "original"]


def _mats():
    m = {}
    g = np.load(os.path.join(ROOT, "tests", "golden", "sample_4x4.npz"))
    m["sample4"] = (g["offsets"], g["indices"], 4)
    A = synth.c1_uniform_4k()
    m["c1"] = (A.offsets, A.indices, A.cols)
    off, idx, _ = random_csr(3000, 2500, 40, seed=1, long_rows=((7, 2500), (2999, 1700)))
    m["skewed"] = (off, idx, 2500)
    m["one_row"] = (np.array([0, 3000], np.int32), np.arange(3000, dtype=np.int32), 3000)
    m["no_nnz"] = (np.zeros(129, np.int32), np.zeros(0, np.int32), 7)
    B = synth.powerlaw_csr(60000, 80000, 1.8, 1, 20000, 600000, seed=9, force_first_max=True)
    m["powerlaw_60k"] = (B.offsets, B.indices, B.cols)
    return m


MATS = _mats()
TUNINGS = [("merge_path_flat", {}), ("merge_path_flat", {"threads_per_block": 128,
                                                        "items_per_thread": 5}),
           ("merge_path_flat", {"threads_per_block": 256, "items_per_thread": 8}),
           ("work_oriented", {}), ("work_oriented", {"threads_per_block": 256, "grid_size": 77}),
           ("group_mapped", {}), ("group_mapped", {"threads_per_block": 256,
                                                   "threads_per_tile": 32}),
           ("group_mapped", {"threads_per_block": 256, "threads_per_tile": 128}),
           ("thread_mapped", {}), ("original", {})]


def _data(name, kind, seed=3):
    off, idx, cols = MATS[name]
    rng = np.random.default_rng(seed)
    if kind == "f64":
        val = rng.uniform(-1, 1, idx.shape[0])
        x = rng.uniform(-1, 1, cols)
    else:
        val = rng.integers(-2 ** 31, 2 ** 31, idx.shape[0], dtype=np.int64).astype(np.int32)
        x = rng.integers(-2 ** 31, 2 ** 31, cols, dtype=np.int64).astype(np.int32)
    return off, idx, val, x, cols


def _run(schedule, A, x_np, **tuning):
    x = torch.from_numpy(x_np).cuda()
    y = torch.full((A.rows,), -7, dtype=x.dtype, device="cuda")
    loops.run_spmv(schedule, A, x, y, **tuning)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("schedule,tuning", TUNINGS,
                         ids=[f"{s}-{'-'.join(map(str, t.values())) or 'default'}"
                              for s, t in TUNINGS])
@pytest.mark.parametrize("name", list(MATS))
def test_double(schedule, tuning, name):
    off, idx, val, x, cols = _data(name, "f64")
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type="f64")
    y = _run(schedule, A, x, **tuning)
    ref, a = oracle.spmv_d_ref(off, idx, val, x)
    n = np.diff(off).astype(np.float64)
    assert np.all(np.abs(y - ref) <= (n + 1) * 2.0 ** -52 * a), (schedule, tuning, name)
    if schedule in ("thread_mapped", "original"):
        assert np.array_equal(y.view(np.uint64), oracle.spmv_d(off, idx, val, x).view(np.uint64))
    assert np.array_equal(y.view(np.uint64), _run(schedule, A, x, **tuning).view(np.uint64))


@pytest.mark.parametrize("schedule,tuning", TUNINGS,
                         ids=[f"{s}-{'-'.join(map(str, t.values())) or 'default'}"
                              for s, t in TUNINGS])
@pytest.mark.parametrize("name", list(MATS))
def test_int(schedule, tuning, name):
    off, idx, val, x, cols = _data(name, "i32")
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type="i32")
    y = _run(schedule, A, x, **tuning)
    assert np.array_equal(y, oracle.spmv_i32(off, idx, val, x)), (schedule, tuning, name)


def test_sample_known_answer_all_types():
    off, idx, _ = MATS["sample4"]
    val = np.array([5, 8, 3, 6])
    for vt, dt in (("f64", np.float64), ("i32", np.int32)):
        A = loops.csr_t.from_host(4, 4, off, idx, val.astype(dt), value_type=vt)
        for s in SCHEDULES:
            y = _run(s, A, np.array([1, 2, 3, 4], dt))
            assert y.tolist() == [0, 21, 9, 12], (vt, s)


def test_typed_rejects_mismatched_vectors():
    off, idx, val, x, cols = _data("skewed", "f64")
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type="f64")
    with pytest.raises(ValueError):
        loops.run_spmv("merge_path_flat", A, torch.zeros(cols, device="cuda"),
                       torch.zeros(A.rows, dtype=torch.float64, device="cuda"))
    with pytest.raises(loops.LoopsError):  # a tile the typed path does not compile
        loops.run_spmv("merge_path_flat", A, torch.zeros(cols, dtype=torch.float64, device="cuda"),
                       torch.zeros(A.rows, dtype=torch.float64, device="cuda"),
                       threads_per_block=512, items_per_thread=2)


@pytest.mark.parametrize("vt", ["double", "int"])
@pytest.mark.parametrize("alg", ["original", "thread_mapped", "group_mapped", "work_oriented",
                                 "merge_path"])
def test_cli_types(alg, vt):
    """The C++ drop-in with type_t = double / int: the reference's README
    check on chesapeake (Dimensions 39 x 39 (340), Errors 0)."""
    r = subprocess.run([os.path.join(BIN, f"loops.spmv.{alg}"), "-m",
                        os.path.join(ROOT, "tests", "golden", "chesapeake.mtx"), "--validate",
                        "-t", vt], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Dimensions:\t39 x 39 (340)" in r.stdout and "Errors:\t\t0" in r.stdout, r.stdout
