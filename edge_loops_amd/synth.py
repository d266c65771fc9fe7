"""Seeded synthetic CSR inputs (host side, numpy output) from
libloops_synth.so, plus the named workloads the benchmark and tests use.

Workloads (sizes of the reference path's benchmark configurations):
  c1_uniform_4k     4096 x 4096, exactly 8 nnz/row, seed 1
  c2_powerlaw_2m    2,000,000 x 2,000,000, ~40M nnz, Zipf(1.8) row lengths
                    on [1, 200000] scaled to 40M, row 0 = 200000, seed 2
  c3_spmm_1m        1,000,000 x 1,000,000, ~16M nnz, Zipf(2.0), k = 64, seed 3
Columns are distinct, uniform and sorted per row; values and x are uniform
[-1, 1) fp32.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._capi import SYNTH_PATH

_lib = None
_NT = min(16, os.cpu_count() or 1)


def _synth():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_PATH):
            raise ImportError(f"{SYNTH_PATH} is missing: run `make -C edge_loops_amd/csrc`")
        L = ctypes.CDLL(SYNTH_PATH)
        i64, f64, u64 = ctypes.c_int64, ctypes.c_double, ctypes.c_uint64
        L.loops_synth_row_lengths.argtypes = [i64, i64, f64, i64, i64, i64, ctypes.c_int,
                                              u64, ctypes.c_void_p,
                                              ctypes.POINTER(i64), ctypes.c_int]
        L.loops_synth_row_lengths.restype = ctypes.c_int
        L.loops_synth_fill.argtypes = [i64, i64, ctypes.c_void_p, u64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int]
        L.loops_synth_fill.restype = ctypes.c_int
        L.loops_synth_uniform.argtypes = [u64, i64, ctypes.c_void_p, ctypes.c_int]
        L.loops_synth_uniform.restype = None
        _lib = L
    return _lib


class HostCSR:
    """Host CSR arrays (numpy): rows, cols, offsets, indices, values."""

    def __init__(self, rows, cols, offsets, indices, values):
        self.rows, self.cols = int(rows), int(cols)
        self.offsets, self.indices, self.values = offsets, indices, values

    @property
    def nnz(self):
        return int(self.offsets[-1])

    def max_row(self):
        return int(np.diff(self.offsets).max()) if self.rows else 0


def powerlaw_csr(rows, cols, zipf_s, min_len, max_len, target_nnz, seed,
                 force_first_max=False, nthreads=_NT):
    """CSR with Zipf(zipf_s) row lengths (zipf_s == 0: every row max_len)."""
    L = _synth()
    lengths = np.empty(rows, np.int32)
    nnz = ctypes.c_int64()
    rc = L.loops_synth_row_lengths(rows, cols, float(zipf_s), min_len, max_len,
                                   int(target_nnz), int(bool(force_first_max)),
                                   seed, lengths.ctypes.data, ctypes.byref(nnz),
                                   nthreads)
    if rc != 0:
        raise ValueError("loops_synth_row_lengths: invalid arguments or overflow")
    offsets = np.zeros(rows + 1, np.int32)
    np.cumsum(lengths, out=offsets[1:])
    indices = np.empty(int(nnz.value), np.int32)
    values = np.empty(int(nnz.value), np.float32)
    rc = L.loops_synth_fill(rows, cols, offsets.ctypes.data, seed,
                            indices.ctypes.data, values.ctypes.data, nthreads)
    if rc != 0:
        raise ValueError("loops_synth_fill: invalid offsets")
    return HostCSR(rows, cols, offsets, indices, values)


def uniform(n, seed, nthreads=_NT):
    out = np.empty(int(n), np.float32)
    _synth().loops_synth_uniform(seed, int(n), out.ctypes.data, nthreads)
    return out


def dense(rows, cols, seed, nthreads=_NT):
    return uniform(rows * cols, seed, nthreads).reshape(rows, cols)


def c1_uniform_4k():
    return powerlaw_csr(4096, 4096, 0.0, 8, 8, 0, seed=1)


def c2_powerlaw_2m():
    return powerlaw_csr(2_000_000, 2_000_000, 1.8, 1, 200_000, 40_000_000, seed=2,
                        force_first_max=True)


def c3_spmm_1m():
    return powerlaw_csr(1_000_000, 1_000_000, 2.0, 1, 200_000, 16_000_000, seed=3)


def spmv_bytes(rows, cols, nnz, value_bytes=4):
    """Algorithmic HBM bytes of one SpMV: values + column ids, offsets, x read
    once, y written once."""
    return nnz * (4 + value_bytes) + 4 * (rows + 1) + 4 * cols + 4 * rows


def spmm_bytes(rows, cols, nnz, k, value_bytes=4):
    return nnz * (4 + value_bytes) + 4 * (rows + 1) + 4 * k * cols + 4 * k * rows


def c4_batch_shapes(count=512, seed=4):
    """Row counts of the batched workload: square matrices with rows
    log-uniform in [1e3, 2e5]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(1e3), np.log(2e5), count)).astype(np.int64)


def c4_batch_matrix(i, rows, seed=4):
    """Matrix i of the batched workload: nnz/row Zipf(1.8) on [1, 4096]
    (unscaled), uniform sorted distinct columns, fp32 uniform[-1, 1]."""
    rows = int(rows)
    return powerlaw_csr(rows, rows, 1.8, 1, min(4096, rows), 0, seed=seed * 100_000 + i)


def c4_batch_512(count=512, seed=4):
    return [c4_batch_matrix(i, r, seed) for i, r in enumerate(c4_batch_shapes(count, seed))]


def c5_pool_shapes(count=64, seed=5):
    """(rows, target nnz, Zipf exponent) of the served pool: rows log-uniform
    in [1e4, 4e6], nnz log-uniform in [1e5, 8e6] (at least one per row),
    exponent uniform in [1.8, 2.2]."""
    rng = np.random.default_rng(seed)
    rows = np.exp(rng.uniform(np.log(1e4), np.log(4e6), count)).astype(np.int64)
    nnz = np.exp(rng.uniform(np.log(1e5), np.log(8e6), count)).astype(np.int64)
    nnz = np.maximum(nnz, rows)
    s = rng.uniform(1.8, 2.2, count)
    return [(int(r), int(n), float(z)) for r, n, z in zip(rows, nnz, s)]


def c5_pool_matrix(i, shape, seed=5):
    rows, nnz, s = shape
    return powerlaw_csr(rows, rows, s, 1, min(100_000, rows), nnz, seed=seed * 100_000 + i)


def c5_pool_64(count=64, seed=5):
    return [c5_pool_matrix(i, sh, seed) for i, sh in enumerate(c5_pool_shapes(count, seed))]
