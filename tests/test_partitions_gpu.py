"""Device partitions of the schedule API, bit for bit against the oracle.

loops.schedule_partitions (edge_loops_amd/examples/schedule_partitions.cpp)
is a user program built against include/loops/: its kernels call
setup<merge_path_flat>::init(), setup<work_oriented>::init() and the
setup<group_mapped> accessors and write what they return.  The oracle
restates the reference's partitions: per-lane merge-path starts
(merge_path_flat.hxx:321-385), work_oriented shares (work_oriented.hxx:
146-167) and the group_mapped (group, virtual atom) map (group_mapped.hxx:
146-213).  Every coordinate and index must be equal.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import random_csr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edge_loops_amd", "bin", "loops.schedule_partitions")
CHES = os.path.join(ROOT, "tests", "golden", "chesapeake.mtx")


def write_mtx(path, rows, cols, off, idx):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n")
        f.write(f"{rows} {cols} {idx.shape[0]}\n")
        r = np.repeat(np.arange(rows), np.diff(off)) + 1
        np.savetxt(f, np.stack([r, idx.astype(np.int64) + 1], axis=1), fmt="%d")


def _cases():
    cases = {"chesapeake": None}
    g = np.load(os.path.join(ROOT, "tests", "golden", "sample_4x4.npz"))
    cases["sample4"] = (g["offsets"], g["indices"], 4)
    from edge_loops_amd import synth
    A = synth.c1_uniform_4k()
    cases["c1"] = (A.offsets, A.indices, A.cols)
    off, idx, _ = random_csr(5000, 4000, 30, seed=31, long_rows=((3, 4000), (4998, 3500)))
    cases["skewed"] = (off, idx, 4000)
    off, idx, _ = random_csr(3000, 900, 3, seed=32, empty_frac=0.8)
    cases["mostly_empty"] = (off, idx, 900)
    B = synth.powerlaw_csr(60000, 80000, 1.8, 1, 20000, 600000, seed=9, force_first_max=True)
    cases["powerlaw_60k"] = (B.offsets, B.indices, B.cols)
    return cases


CASES = _cases()


@pytest.fixture(scope="module")
def dumps(tmp_path_factory):
    out = {}
    for name, case in CASES.items():
        d = tmp_path_factory.mktemp(name)
        if case is None:
            path = CHES
            _, _, off, idx, _ = oracle.mtx_read(CHES)
        else:
            off, idx, cols = case
            path = str(d / f"{name}.mtx")
            write_mtx(path, off.shape[0] - 1, cols, off, idx)
        r = subprocess.run([BIN, "-m", path, "-o", str(d)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        info = dict(line.split() for line in r.stdout.splitlines() if " " in line)
        out[name] = (d, off, info)
    return out


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("tpb,ipt", [(128, 5), (64, 3)])
def test_merge_path_thread_starts(dumps, name, tpb, ipt):
    d, off, info = dumps[name]
    tiles = int(info[f"tiles_{tpb}_{ipt}"])
    rows, nnz = off.shape[0] - 1, int(off[-1])
    assert tiles == -(-(rows + nnz) // (tpb * ipt))
    if tiles == 0:
        return
    ref = np.stack([oracle.mp_thread_coords(off, tpb, ipt, t) for t in range(tiles)])
    for suffix in ("", "_meta"):
        got = np.fromfile(d / f"mp_{tpb}_{ipt}{suffix}.bin", dtype=np.uint32)
        assert np.array_equal(got.reshape(tiles, tpb, 2), ref), suffix


@pytest.mark.parametrize("name", list(CASES))
def test_work_oriented_shares(dumps, name):
    d, off, info = dumps[name]
    n = int(info["wo_threads"])
    got = np.fromfile(d / "wo_128.bin", dtype=np.uint32).reshape(n, 4)
    assert np.array_equal(got, oracle.wo_coords(off, n))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("tpb,tpt", [(128, 128), (256, 32), (256, 64), (256, 128), (256, 8)])
def test_group_mapped_atom_map(dumps, name, tpb, tpt):
    d, off, info = dumps[name]
    got = np.fromfile(d / f"gm_{tpb}_{tpt}.bin", dtype=np.int64).reshape(-1, 2)
    g, v = oracle.gm_atom_map(off, tpt)
    assert got.shape[0] == int(off[-1])
    assert np.array_equal(got[:, 0], g) and np.array_equal(got[:, 1], v)

