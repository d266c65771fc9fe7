"""Container utilities of the C++ drop-in on host memory (no GPU): the
sample matrix (util/sample.hxx), offsets <-> indices conversions
(container/detail/convert.hxx), COO sorts, duplicate removal and the
COO / CSR / CSC conversions, file-name helpers, and the random fills and
random CSR matrices of util/generate.hxx, each against known answers or
closed forms, by the program built from edge_loops_amd/examples/
containers.cpp."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edge_loops_amd", "bin", "loops.containers")


def test_containers_host():
    assert os.path.exists(BIN), f"{BIN} is missing: build with make -C edge_loops_amd/csrc"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for group in ("sample, conversions, COO / CSC / CSR: ok", "conversion edge cases: ok",
                  "COO sorts and duplicate removal: ok", "file names: ok",
                  "uniform fills (host): ok", "random CSR (host): ok"):
        assert group in r.stdout, r.stdout
    assert "checks passed" in r.stdout
