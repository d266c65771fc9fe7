"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

Each check is chosen so that a plausible mistake in the restatement fails
it; the planted-mistake tests below show that for the main ones (a mutated
copy of the oracle's logic, on the CPU, must fail the same check).
"""
import os
import zlib

import numpy as np
import pytest

import oracle
from conftest import random_csr


# ---------------------------------------------------------------------------
# Known answers from the reference's own fixtures
# ---------------------------------------------------------------------------
def test_sample_4x4_known_answer(golden_dir):
    g = np.load(os.path.join(golden_dir, "sample_4x4.npz"))
    y = oracle.spmv_f32(g["offsets"], g["indices"], g["values"], g["x"])
    assert np.array_equal(y, g["y"])  # [0, 21, 9, 12], exact integers


def test_chesapeake_known_answer(golden_dir):
    rows, cols, off, idx, val = oracle.mtx_read(os.path.join(golden_dir, "chesapeake.mtx"))
    # README.md:86 "Dimensions: 39 x 39 (340)"
    assert (rows, cols, int(off[-1])) == (39, 39, 340)
    assert np.all(val == 1.0)
    # (row, col) order: rows non-decreasing by construction, columns sorted
    for r in range(rows):
        seg = idx[off[r]:off[r + 1]]
        assert np.all(np.diff(seg) > 0)
    deg = np.load(os.path.join(golden_dir, "chesapeake_degrees.npy"))
    y = oracle.spmv_f32(off, idx, val, np.ones(cols, np.float32))
    assert np.array_equal(y, deg)
    # symmetric: A == A^T
    dense = np.zeros((rows, cols))
    for r in range(rows):
        dense[r, idx[off[r]:off[r + 1]]] = val[off[r]:off[r + 1]]
    assert np.array_equal(dense, dense.T)


def test_c1_golden_regression(golden_dir):
    from edge_loops_amd import synth
    g = np.load(os.path.join(golden_dir, "c1_4096.npz"))
    A = synth.c1_uniform_4k()
    x = synth.uniform(A.cols, seed=11)
    crc = [zlib.crc32(A.offsets.tobytes()), zlib.crc32(A.indices.tobytes()),
           zlib.crc32(A.values.tobytes()), zlib.crc32(x.tobytes())]
    assert list(g["crc"]) == crc
    assert np.array_equal(oracle.spmv_f32(A.offsets, A.indices, A.values, x), g["y"])
    assert np.array_equal(oracle.mp_tile_coords(A.offsets, 640), g["coords_128_5"])
    assert np.array_equal(oracle.mp_tile_coords(A.offsets, 2048), g["coords_256_8"])
    # merge_path_flat.cuh:116-117 with 4096 rows + 32768 nnz: T = 58
    assert g["coords_128_5"].shape == (59, 2)


# ---------------------------------------------------------------------------
# Values: closed forms and the FP64 bound
# ---------------------------------------------------------------------------
def test_spmv_f32_matches_dense_on_integers():
    # Integer-valued inputs small enough that every fp32 sum is exact, so the
    # oracle must equal a dense matrix-vector product.
    off, idx, _ = random_csr(200, 150, 30, seed=3)
    rng = np.random.default_rng(4)
    val = rng.integers(-8, 9, idx.shape[0]).astype(np.float32)
    x = rng.integers(-8, 9, 150).astype(np.float32)
    dense = np.zeros((200, 150))
    for r in range(200):
        dense[r, idx[off[r]:off[r + 1]]] = val[off[r]:off[r + 1]]
    assert np.array_equal(oracle.spmv_f32(off, idx, val, x), (dense @ x).astype(np.float32))


def test_spmv_f32_is_sequential_fp32():
    # The fp32 result must be the left-to-right fp32 sum (helpers.hxx:106-112):
    # compare against numpy doing the same loop in float32, row by row.
    off, idx, val = random_csr(50, 500, 300, seed=5)
    x = np.random.default_rng(6).uniform(-1, 1, 500).astype(np.float32)
    y = oracle.spmv_f32(off, idx, val, x)
    for r in range(50):
        s = np.float32(0)
        for nz in range(off[r], off[r + 1]):
            s = np.float32(s + np.float32(val[nz] * x[idx[nz]]))
        assert y[r] == s


def test_f64_oracle_bounds_f32():
    off, idx, val = random_csr(300, 4000, 2000, seed=7)
    x = np.random.default_rng(8).uniform(-1, 1, 4000).astype(np.float32)
    y32 = oracle.spmv_f32(off, idx, val, x)
    y64, a = oracle.spmv_f64(off, idx, val, x)
    assert np.all(np.abs(y64) <= a + 1e-300)
    # sequential fp32 summation error <= n * u * sum|a x| (u = 2^-24)
    n = np.diff(off).astype(np.float64)
    assert np.all(np.abs(y32 - y64) <= (n + 1) * 2.0 ** -24 * a + 1e-30)
    bad, worst = oracle.count_rel_errors(y32, y64, a, 1e-5)
    assert bad == 0 and worst < 1e-5


def test_spmm_equals_spmv_columnwise():
    off, idx, val = random_csr(120, 90, 40, seed=9)
    B = np.random.default_rng(10).uniform(-1, 1, (90, 7)).astype(np.float32)
    C = oracle.spmm_f32(off, idx, val, B)
    for c in range(7):
        assert np.array_equal(C[:, c], oracle.spmv_f32(off, idx, val, np.ascontiguousarray(B[:, c])))
    C64, A64 = oracle.spmm_f64(off, idx, val, B)
    assert np.all(np.abs(C - C64) <= 1e-5 * A64 + 1e-30)


def test_validators():
    a = np.array([0, 1, 2, 3], np.float32)
    b = np.array([0, 1.005, 2.02, 3], np.float32)
    assert oracle.count_abs_errors(a, b) == 1  # |2.02-2| > 1e-2 only
    y64 = a.astype(np.float64)
    s = np.array([0, 1, 1, 1], np.float64)
    assert oracle.count_rel_errors(np.array([0, 1, 2, 3], np.float32), y64, s)[0] == 0
    assert oracle.count_rel_errors(np.array([1e-30, 1, 2, 3], np.float32), y64, s)[0] == 1


# ---------------------------------------------------------------------------
# Partitions: brute force, invariants, closed forms
# ---------------------------------------------------------------------------
CASES = [
    (np.array([0], np.int32),),                       # 0 rows
    (np.array([0, 0, 0, 0], np.int32),),              # only empty rows
    (np.array([0, 5], np.int32),),                    # one row
    (np.array([0, 0, 2, 3, 4], np.int32),),           # the 4x4 sample
    (np.array([0, 0, 0, 7, 7, 7, 9, 9], np.int32),),  # clustered empties
]


def _tiny_offsets():
    out = [c[0] for c in CASES]
    for s in range(12):
        off, _, _ = random_csr(int(np.random.default_rng(s).integers(1, 40)), 60, 25,
                               seed=100 + s, long_rows=((0, 55),) if s % 3 == 0 else ())
        out.append(off)
    return out


@pytest.mark.parametrize("off", _tiny_offsets(), ids=lambda o: f"R{o.shape[0]-1}N{o[-1]}")
def test_search_vs_brute_force_walk(off):
    rows, nnz = off.shape[0] - 1, int(off[-1])
    walk = oracle.mp_walk(off)
    for d in range(rows + nnz + 1):
        assert oracle.mp_search(off, d) == tuple(walk[d]), d
        x, y = walk[d]
        assert x + y == d
        assert (x == rows) or (off[x] <= y <= off[x + 1])
    # monotone in d
    assert np.all(np.diff(walk.astype(np.int64), axis=0) >= 0)


@pytest.mark.parametrize("off", _tiny_offsets(), ids=lambda o: f"R{o.shape[0]-1}N{o[-1]}")
def test_tile_and_thread_coords(off):
    rows, nnz = off.shape[0] - 1, int(off[-1])
    total = rows + nnz
    walk = oracle.mp_walk(off)
    for tpb, ipt in ((2, 3), (4, 2), (64, 3), (128, 5)):
        ti = tpb * ipt
        coords = oracle.mp_tile_coords(off, ti)
        assert np.array_equal(coords, oracle.mp_tile_coords(off, ti, clamp=False))
        T = coords.shape[0] - 1
        assert T == -(-total // ti)
        for t in range(T + 1):
            assert tuple(coords[t]) == tuple(walk[min(t * ti, total)])
        for t in range(T):
            tc = oracle.mp_thread_coords(off, tpb, ipt, t)
            for lane in range(tpb):
                d = min(t * ti + lane * ipt, min((t + 1) * ti, total))
                assert tuple(tc[lane]) == tuple(walk[d]), (t, lane)


@pytest.mark.parametrize("off", _tiny_offsets(), ids=lambda o: f"R{o.shape[0]-1}N{o[-1]}")
def test_work_oriented_coords(off):
    rows, nnz = off.shape[0] - 1, int(off[-1])
    total = rows + nnz
    walk = oracle.mp_walk(off)
    for nt in (1, 3, 7, 64):
        c = oracle.wo_coords(off, nt)
        wpt = -(-total // nt) if total else 0
        for t in range(nt):
            up = min(wpt * t, total)
            lo = min(up + wpt, total)
            assert tuple(c[t, :2]) == tuple(walk[up])
            assert tuple(c[t, 2:]) == tuple(walk[lo])


@pytest.mark.parametrize("tpt", [1, 4, 64, 128])
def test_group_mapped_atom_map(tpt):
    off, _, _ = random_csr(300, 50, 12, seed=21, long_rows=((5, 50),))
    g, v = oracle.gm_atom_map(off, tpt)
    rows = off.shape[0] - 1
    for r in range(rows):
        grp = r // tpt
        base = off[grp * tpt]
        for nz in range(off[r], off[r + 1]):
            assert g[nz] == grp and v[nz] == nz - base
    assert np.all(g >= 0)


# ---------------------------------------------------------------------------
# Planted mistakes (in copies of the oracle's logic, on the CPU) must fail
# ---------------------------------------------------------------------------
def _search_mutant(off, d, strict):
    rows, nnz = off.shape[0] - 1, int(off[-1])
    lo, hi = max(d - nnz, 0), min(d, rows)
    first, count = lo, hi - lo
    while count > 0:
        step = count // 2
        idx = first + step
        ok = (off[idx + 1] < d - idx - 1) if strict else (off[idx + 1] <= d - idx - 1)
        if ok:
            first, count = idx + 1, count - step - 1
        else:
            count = step
    return min(first, rows), d - first


def test_planted_search_off_by_one_is_caught():
    off = np.array([0, 0, 2, 3, 4], np.int32)
    walk = oracle.mp_walk(off)
    good = [(_search_mutant(off, d, False) == tuple(walk[d])) for d in range(9)]
    bad = [(_search_mutant(off, d, True) == tuple(walk[d])) for d in range(9)]
    assert all(good) and not all(bad)


def test_planted_dropped_term_is_caught(golden_dir):
    g = np.load(os.path.join(golden_dir, "sample_4x4.npz"))
    off, idx, val, x = g["offsets"], g["indices"], g["values"], g["x"]
    y = np.zeros(4, np.float32)
    for r in range(4):
        for nz in range(off[r], off[r + 1] - 1):  # mutant: last term dropped
            y[r] += val[nz] * x[idx[nz]]
    assert not np.array_equal(y, g["y"])


def test_planted_transposed_operand_is_caught(golden_dir):
    deg = np.load(os.path.join(golden_dir, "chesapeake_degrees.npy"))
    g = np.load(os.path.join(golden_dir, "sample_4x4.npz"))
    off, idx, val, x = g["offsets"], g["indices"], g["values"], g["x"]
    y = np.zeros(4, np.float32)
    for r in range(4):  # mutant: y += A^T x
        for nz in range(off[r], off[r + 1]):
            y[idx[nz]] += val[nz] * x[r]
    assert not np.array_equal(y, g["y"])
    assert deg.shape == (39,)


# ---------------------------------------------------------------------------
# The all-cores CPU baseline must compute what the sequential validator does
# ---------------------------------------------------------------------------
def _skewed_csr():
    return random_csr(3000, 2500, 40, seed=41, long_rows=((0, 2400), (1500, 2000), (2999, 1800)))


@pytest.mark.parametrize("nthreads", [1, 2, 7, 16])
def test_spmv_f32_mt_bit_identical(golden_dir, nthreads):
    from edge_loops_amd import synth
    A = synth.c1_uniform_4k()
    cases = [(A.offsets, A.indices, A.values, synth.uniform(A.cols, seed=11))]
    off, idx, val = _skewed_csr()
    cases.append((off, idx, val, np.random.default_rng(42).uniform(-1, 1, 2500).astype(np.float32)))
    # more threads than rows, and a matrix of empty rows only
    cases.append((np.array([0, 3, 3, 5], np.int32), np.array([0, 1, 2, 0, 2], np.int32),
                  np.array([1, 2, 3, 4, 5], np.float32), np.array([1, 10, 100], np.float32)))
    cases.append((np.zeros(9, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32),
                  np.ones(4, np.float32)))
    for off, idx, val, x in cases:
        ref = oracle.spmv_f32(off, idx, val, x)
        out = np.full(off.shape[0] - 1, np.nan, np.float32)  # a row left unwritten stays NaN
        y = oracle.spmv_f32_mt(off, idx, val, x, nthreads, out=out)
        assert y is out
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_planted_mt_split_mistake_is_caught():
    # Mutant of the nnz-balanced split: each thread stops one row short of
    # its boundary.  The NaN sentinel of the bit-identity check catches it.
    off, idx, val = _skewed_csr()
    x = np.random.default_rng(42).uniform(-1, 1, 2500).astype(np.float32)
    ref = oracle.spmv_f32(off, idx, val, x)
    rows, nnz, nt = off.shape[0] - 1, int(off[-1]), 7
    y = np.full(rows, np.nan, np.float32)
    r = 0
    for t in range(nt):
        r1 = rows if t == nt - 1 else int(np.searchsorted(off[:rows], nnz * (t + 1) // nt, "left"))
        y[r:max(r, r1 - 1)] = ref[r:max(r, r1 - 1)]
        r = r1
    assert not np.array_equal(y.view(np.uint32), ref.view(np.uint32))


# ---------------------------------------------------------------------------
# FP64 oracle: known answers that a float accumulator or product fails
# ---------------------------------------------------------------------------
def _cancel_case():
    # Row 0: 2^25 + 1 - 2^25 with x = 1.  Exact: 1.  In fp32 2^25 + 1 rounds
    # to 2^25 (its ulp is 4), so an fp32 accumulator gives 0.
    # Row 1: (1 + 2^-23) * (1 + 2^-23) = 1 + 2^-22 + 2^-46, exact in double,
    # 1 + 2^-22 when the product is rounded to fp32 first.
    # Row 2: empty.  Row 3: 2^25 - 2^25 + 0.5 * 3 = 1.5.
    e = np.float32(1 + 2.0 ** -23)
    off = np.array([0, 3, 4, 4, 7], np.int32)
    idx = np.array([0, 1, 2, 3, 0, 2, 4], np.int32)
    val = np.array([2.0 ** 25, 1.0, -(2.0 ** 25), e, 2.0 ** 25, -(2.0 ** 25), 0.5], np.float32)
    x = np.array([1, 1, 1, e, 3], np.float32)
    y_exact = np.array([1.0, (1 + 2.0 ** -23) ** 2, 0.0, 1.5])
    abs_exact = np.array([2.0 ** 26 + 1, (1 + 2.0 ** -23) ** 2, 0.0, 2.0 ** 26 + 1.5])
    return off, idx, val, x, y_exact, abs_exact


def test_spmv_f64_known_answers():
    off, idx, val, x, y_exact, abs_exact = _cancel_case()
    y64, a = oracle.spmv_f64(off, idx, val, x)
    assert np.array_equal(y64, y_exact)
    assert np.array_equal(a, abs_exact)
    # the fp32 validator loses both: the cancellation and the product's low bits
    y32 = oracle.spmv_f32(off, idx, val, x)
    assert y32[0] == 0.0 and y32[1] == np.float32(1 + 2.0 ** -22)


def test_spmm_f64_known_answers():
    off, idx, val, x, y_exact, abs_exact = _cancel_case()
    B = np.stack([x, 2 * x, -x], axis=1).astype(np.float32)  # scaling by 2 is exact
    C64, A64 = oracle.spmm_f64(off, idx, val, B)
    assert np.array_equal(C64, np.stack([y_exact, 2 * y_exact, -y_exact], axis=1))
    assert np.array_equal(A64, np.stack([abs_exact, 2 * abs_exact, abs_exact], axis=1))
    # column-by-column equal to the SpMV FP64 oracle on random data
    off, idx, val = random_csr(80, 60, 30, seed=43)
    B = np.random.default_rng(44).uniform(-1, 1, (60, 5)).astype(np.float32)
    C64, A64 = oracle.spmm_f64(off, idx, val, B)
    for c in range(5):
        y64, a = oracle.spmv_f64(off, idx, val, np.ascontiguousarray(B[:, c]))
        assert np.array_equal(C64[:, c], y64) and np.array_equal(A64[:, c], a)


def test_planted_float_accumulator_is_caught():
    # Mutant FP64 oracle accumulating in float32: the known answer fails.
    off, idx, val, x, y_exact, _ = _cancel_case()
    y = np.zeros(4)
    for r in range(4):
        s = np.float32(0)
        for nz in range(off[r], off[r + 1]):
            s = np.float32(s + np.float64(val[nz]) * np.float64(x[idx[nz]]))
        y[r] = s
    assert not np.array_equal(y, y_exact)


# ---------------------------------------------------------------------------
# Matrix Market reader: duplicates stay in input order (stable sort)
# ---------------------------------------------------------------------------
def test_mtx_duplicates_keep_input_order(tmp_path):
    # Entries listed out of order with three copies of (2, 1) and two of
    # (1, 3); each copy has its own value so the CSR order is visible.
    p = tmp_path / "dups.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n"
                 "3 3 8\n"
                 "2 1 5.0\n1 3 7.0\n2 1 -1.0\n3 3 2.0\n1 3 0.25\n2 1 9.0\n1 1 4.0\n2 2 3.0\n")
    rows, cols, off, idx, val = oracle.mtx_read(str(p))
    assert (rows, cols) == (3, 3)
    assert off.tolist() == [0, 3, 7, 8]
    assert idx.tolist() == [0, 2, 2, 0, 0, 0, 1, 2]
    assert val.tolist() == [4.0, 7.0, 0.25, 5.0, -1.0, 9.0, 3.0, 2.0]


def test_mtx_symmetric_expansion_order(tmp_path):
    # symmetric: (r, c) then its mirror (c, r), in entry order; a diagonal
    # entry is not mirrored (market.hxx:187-223).  Duplicate mirrors keep
    # the order in which they were produced.
    p = tmp_path / "sym.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer symmetric\n"
                 "3 3 4\n2 1 1\n3 1 2\n2 1 3\n3 3 4\n")
    rows, cols, off, idx, val = oracle.mtx_read(str(p))
    assert off.tolist() == [0, 3, 5, 7]
    assert idx.tolist() == [1, 1, 2, 0, 0, 0, 2]
    assert val.tolist() == [1.0, 3.0, 2.0, 1.0, 3.0, 2.0, 4.0]


# ---------------------------------------------------------------------------
# double and int value types (the reference's other instantiations)
# ---------------------------------------------------------------------------
def test_spmv_d_known_answers_and_order():
    off, idx, _ = random_csr(200, 150, 30, seed=51)
    rng = np.random.default_rng(52)
    val = rng.integers(-8, 9, idx.shape[0]).astype(np.float64)
    x = rng.integers(-8, 9, 150).astype(np.float64)
    dense = np.zeros((200, 150))
    for r in range(200):
        dense[r, idx[off[r]:off[r + 1]]] = val[off[r]:off[r + 1]]
    assert np.array_equal(oracle.spmv_d(off, idx, val, x), dense @ x)
    # random values: the left-to-right double sum, element by element
    val = rng.uniform(-1, 1, idx.shape[0])
    x = rng.uniform(-1, 1, 150)
    y = oracle.spmv_d(off, idx, val, x)
    for r in range(0, 200, 7):
        s = 0.0
        for nz in range(off[r], off[r + 1]):
            s = s + val[nz] * x[idx[nz]]
        assert y[r] == s


def test_spmv_d_ref_keeps_what_double_loses():
    # 2^54 + 1 - 2^54 with x = 1: exact 1 (long double has a 64-bit
    # mantissa); a double accumulator gives 0 (2^54 + 1 rounds to 2^54)
    off = np.array([0, 3, 3], np.int32)
    idx = np.array([0, 1, 2], np.int32)
    val = np.array([2.0 ** 54, 1.0, -(2.0 ** 54)])
    x = np.ones(3)
    y, a = oracle.spmv_d_ref(off, idx, val, x)
    assert y.tolist() == [1.0, 0.0] and a.tolist() == [2.0 ** 55 + 1, 0.0]
    assert oracle.spmv_d(off, idx, val, x).tolist() == [0.0, 0.0]
    off2, idx2, _ = random_csr(100, 300, 200, seed=53)
    v2 = np.random.default_rng(54).uniform(-1, 1, idx2.shape[0])
    x2 = np.random.default_rng(55).uniform(-1, 1, 300)
    yr, ar = oracle.spmv_d_ref(off2, idx2, v2, x2)
    assert np.all(np.abs(yr) <= ar)
    n = np.diff(off2).astype(np.float64)
    assert np.all(np.abs(oracle.spmv_d(off2, idx2, v2, x2) - yr) <= (n + 1) * 2.0 ** -53 * ar)


def test_spmv_i32_wraps_and_ignores_order():
    off = np.array([0, 2, 5], np.int32)
    idx = np.array([0, 1, 0, 1, 2], np.int32)
    val = np.array([2 ** 16, 2 ** 15, -7, 3, 2 ** 30], np.int32)
    x = np.array([2 ** 15, 2, 4], np.int32)
    y = oracle.spmv_i32(off, idx, val, x)
    # row 0: 2^31 + 2^16 wraps to -2^31 + 2^16; row 1: -7*2^15 + 6 + 2^32 -> wraps
    assert y.tolist() == [np.int32(-(2 ** 31) + 2 ** 16), np.int32(-7 * 2 ** 15 + 6)]
    off2, idx2, _ = random_csr(300, 400, 60, seed=56)
    rng = np.random.default_rng(57)
    v2 = rng.integers(-2 ** 31, 2 ** 31, idx2.shape[0], dtype=np.int64).astype(np.int32)
    x2 = rng.integers(-2 ** 31, 2 ** 31, 400, dtype=np.int64).astype(np.int32)
    y2 = oracle.spmv_i32(off2, idx2, v2, x2)
    # any order: numpy's wrapping int64 products summed per row, mod 2^32
    ref = np.zeros(300, np.int64)
    for r in range(300):
        p = v2[off2[r]:off2[r + 1]].astype(np.int64) * x2[idx2[off2[r]:off2[r + 1]]].astype(np.int64)
        ref[r] = int(p[::-1].sum()) & 0xffffffff
    assert np.array_equal(y2.view(np.uint32).astype(np.int64), ref)
