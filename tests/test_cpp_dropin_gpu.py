"""The C++ drop-in layer on the GPU: the example programs built from
edge_loops_amd/examples against include/loops/ (host entry points over the C
ABI, and user kernels over the device schedule API).

Known answer from the reference: `loops.spmv.merge_path -m chesapeake.mtx
--validate -v` prints "Dimensions: 39 x 39 (340)" and "Errors: 0"
(README.md:82-88)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import random_csr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edge_loops_amd", "bin")
CHES = os.path.join(ROOT, "tests", "golden", "chesapeake.mtx")
SPMV = ["original", "thread_mapped", "group_mapped", "work_oriented", "merge_path"]


def _run(args, timeout=120):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def write_mtx(path, rows, cols, off, idx, val):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{rows} {cols} {idx.shape[0]}\n")
        r = np.repeat(np.arange(rows), np.diff(off))
        lines = [f"{a + 1} {b + 1} {float(v)!r}\n" for a, b, v in zip(r, idx, val.astype(np.float64))]
        f.writelines(lines)


@pytest.mark.parametrize("alg", SPMV)
def test_readme_sanity_check(alg):
    out = _run([os.path.join(BIN, f"loops.spmv.{alg}"), "-m", CHES, "--validate", "-v"])
    assert "Dimensions:\t39 x 39 (340)" in out, out
    assert "Errors:\t\t0" in out, out
    assert out.splitlines()[0].split(",")[1] == "chesapeake"


@pytest.fixture(scope="module")
def skewed_mtx(tmp_path_factory):
    off, idx, val = random_csr(5000, 4000, 30, seed=31, long_rows=((3, 4000), (4998, 3500)))
    p = str(tmp_path_factory.mktemp("mtx") / "skewed_5k.mtx")
    write_mtx(p, 5000, 4000, off, idx, val)
    return p, off, idx, val


@pytest.mark.parametrize("alg", SPMV)
def test_cli_validate_skewed(alg, skewed_mtx):
    out = _run([os.path.join(BIN, f"loops.spmv.{alg}"), "-m", skewed_mtx[0], "--validate"])
    assert "Dimensions:\t5000 x 4000" in out and "Errors:\t\t0" in out, out


def test_schedule_api_user_kernels(skewed_mtx, tmp_path):
    path, off, idx, val = skewed_mtx
    out = _run([os.path.join(BIN, "loops.schedule_api"), "-m", path, "-o", str(tmp_path)])
    x = np.fromfile(tmp_path / "x.bin", dtype=np.float32)
    y64, a = oracle.spmv_f64(off, idx, val, x)
    # User code is built like a user's build (hipcc defaults, multiply-add
    # contraction allowed), so only the 1e-5 row bar applies.
    for name in ("thread_mapped", "block_mapped", "warp_mapped", "wave_mapped", "group128",
                 "work_oriented", "merge_path_flat"):
        assert f"{name} done" in out
        y = np.fromfile(tmp_path / f"y_{name}.bin", dtype=np.float32)
        bad, worst = oracle.count_rel_errors(y, y64, a, 1e-5)
        assert bad == 0, (name, worst)


@pytest.mark.parametrize("alg", ["thread_mapped", "group_mapped", "merge_path"])
def test_spmm_cli_runs(alg):
    out = _run([os.path.join(BIN, f"loops.spmm.{alg}"), "-m", CHES, "-n", "10"])
    assert "Elapsed (ms):" in out


def test_containers_device():
    """The container utilities on device memory: the sample matrix, the
    conversions, COO / CSC / CSR built into device vectors, the device fill
    kernel (equal to the host fill of the same seed) and random CSR
    matrices in device memory."""
    out = _run([os.path.join(BIN, "loops.containers"), "--device"])
    for group in ("uniform fills (device): ok", "random CSR (device): ok"):
        assert group in out, out
    assert out.count("sample, conversions, COO / CSC / CSR: ok") == 2, out
