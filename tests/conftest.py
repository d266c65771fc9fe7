import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def random_csr(rows, cols, max_len, seed, empty_frac=0.2, long_rows=()):
    """Small random CSR (distinct sorted columns) with some empty rows and
    optional forced long rows, built with numpy for tests."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(0, max_len + 1, size=rows)
    lengths[rng.random(rows) < empty_frac] = 0
    for r, L in long_rows:
        if r < rows:
            lengths[r] = L
    lengths = np.minimum(lengths, cols)
    offsets = np.zeros(rows + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    indices = np.empty(int(offsets[-1]), np.int32)
    for r in range(rows):
        L = int(lengths[r])
        if L:
            indices[offsets[r]:offsets[r + 1]] = np.sort(rng.choice(cols, L, replace=False))
    values = rng.uniform(-1, 1, int(offsets[-1])).astype(np.float32)
    return offsets.astype(np.int32), indices, values
