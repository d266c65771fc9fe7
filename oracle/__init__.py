"""CPU oracle for the load-balanced CSR SpMV/SpMM path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``, always as the checker and
never as the thing measured or shipped.  The product package
(``edge_loops_amd``) does not import this module.

The arithmetic lives in ``oracle.c`` (built to ``liboracle.so`` by the
Makefile next to it); this module is a thin ctypes/numpy wrapper.  Each C
function cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def build() -> None:
    """Compile liboracle.so (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_spmv_f32.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.oracle_spmv_f32.restype = None
        L.oracle_spmv_f64.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, _f64p, _f64p]
        L.oracle_spmv_f64.restype = None
        L.oracle_spmv_d.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.oracle_spmv_d.restype = None
        L.oracle_spmv_d_ref.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.oracle_spmv_d_ref.restype = None
        L.oracle_spmv_i32.argtypes = [_i64, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.oracle_spmv_i32.restype = None
        L.oracle_spmv_f32_mt.argtypes = [ctypes.c_int, _i64, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.oracle_spmv_f32_mt.restype = ctypes.c_int
        L.oracle_count_abs_errors.argtypes = [_i64, _f32p, _f32p, ctypes.c_double]
        L.oracle_count_abs_errors.restype = _i64
        L.oracle_count_rel_errors.argtypes = [_i64, _f32p, _f64p, _f64p, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_double)]
        L.oracle_count_rel_errors.restype = _i64
        L.oracle_mp_search.argtypes = [_i32p, _i64, _i64, _i64,
                                       ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_mp_search.restype = None
        L.oracle_mp_tile_coords.argtypes = [_i32p, _i64, _i64, _i64, _i64, _u32p]
        L.oracle_mp_tile_coords.restype = None
        L.oracle_mp_tile_coords_unclamped.argtypes = [_i32p, _i64, _i64, _i64, _i64, _u32p]
        L.oracle_mp_tile_coords_unclamped.restype = None
        L.oracle_mp_thread_coords.argtypes = [_i32p, _i64, _i64, _i64, _i64, _i64, _u32p]
        L.oracle_mp_thread_coords.restype = None
        L.oracle_mp_walk.argtypes = [_i32p, _i64, _i64, _u32p]
        L.oracle_mp_walk.restype = None
        L.oracle_wo_coords.argtypes = [_i32p, _i64, _i64, _i64, _u32p]
        L.oracle_wo_coords.restype = None
        L.oracle_gm_atom_map.argtypes = [_i32p, _i64, _i64, _i64p, _i64p]
        L.oracle_gm_atom_map.restype = None
        L.oracle_spmm_f32.argtypes = [_i64, _i32p, _i32p, _f32p, _i64, _f32p, _i64, _f32p, _i64]
        L.oracle_spmm_f32.restype = None
        L.oracle_spmm_f64.argtypes = [_i64, _i32p, _i32p, _f32p, _i64, _f32p, _i64, _f64p, _f64p]
        L.oracle_spmm_f64.restype = None
        L.oracle_mtx_read.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(_i64)] * 3 + [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_mtx_read.restype = ctypes.c_int
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --------------------------------------------------------------------------
# SpMV / SpMM values
# --------------------------------------------------------------------------
def spmv_f32(offsets, indices, values, x):
    """examples/spmv/helpers.hxx:106-112 (fp32, CSR order, no contraction)."""
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, x = _c(values, np.float32), _c(x, np.float32)
    rows = offsets.shape[0] - 1
    y = np.empty(rows, np.float32)
    lib().oracle_spmv_f32(rows, offsets, indices, values, x, y)
    return y


def spmv_f32_mt(offsets, indices, values, x, nthreads, out=None):
    """Row-partitioned threads over the validator loop (each row in CSR
    order); ``out`` (float32, rows) receives y when given."""
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, x = _c(values, np.float32), _c(x, np.float32)
    rows = offsets.shape[0] - 1
    y = np.empty(rows, np.float32) if out is None else out
    if y.dtype != np.float32 or y.shape != (rows,) or not y.flags["C_CONTIGUOUS"]:
        raise ValueError("out must be a contiguous float32 array of rows entries")
    if lib().oracle_spmv_f32_mt(int(nthreads), rows, offsets, indices, values, x, y) != 0:
        raise RuntimeError("oracle_spmv_f32_mt: thread creation failed")
    return y


def spmv_d(offsets, indices, values, x):
    """The validator loop for type_t = double (CSR order, double)."""
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, x = _c(values, np.float64), _c(x, np.float64)
    y = np.empty(offsets.shape[0] - 1, np.float64)
    lib().oracle_spmv_d(y.shape[0], offsets, indices, values, x, y)
    return y


def spmv_d_ref(offsets, indices, values, x):
    """Long-double reference for double SpMV: (y, sum_j |a_ij x_j|)."""
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, x = _c(values, np.float64), _c(x, np.float64)
    y = np.empty(offsets.shape[0] - 1, np.float64)
    a = np.empty_like(y)
    lib().oracle_spmv_d_ref(y.shape[0], offsets, indices, values, x, y, a)
    return y, a


def spmv_i32(offsets, indices, values, x):
    """The validator loop for type_t = int (32-bit wrap-around)."""
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, x = _c(values, np.int32), _c(x, np.int32)
    y = np.empty(offsets.shape[0] - 1, np.int32)
    lib().oracle_spmv_i32(y.shape[0], offsets, indices, values, x, y)
    return y


def spmv_f64(offsets, indices, values, x):
    """FP64 host oracle: returns (y64, sum_j |a_ij x_j|)."""
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, x = _c(values, np.float32), _c(x, np.float32)
    rows = offsets.shape[0] - 1
    y = np.empty(rows, np.float64)
    a = np.empty(rows, np.float64)
    lib().oracle_spmv_f64(rows, offsets, indices, values, x, y, a)
    return y, a


def spmm_f32(offsets, indices, values, B):
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, B = _c(values, np.float32), _c(B, np.float32)
    rows = offsets.shape[0] - 1
    k = B.shape[1]
    C = np.empty((rows, k), np.float32)
    lib().oracle_spmm_f32(rows, offsets, indices, values, k, B, k, C, k)
    return C


def spmm_f64(offsets, indices, values, B):
    offsets, indices = _c(offsets, np.int32), _c(indices, np.int32)
    values, B = _c(values, np.float32), _c(B, np.float32)
    rows = offsets.shape[0] - 1
    k = B.shape[1]
    C = np.empty((rows, k), np.float64)
    A = np.empty((rows, k), np.float64)
    lib().oracle_spmm_f64(rows, offsets, indices, values, k, B, k, C, A)
    return C, A


def count_abs_errors(a, b, tol=1e-2):
    """helpers.hxx:136-139: number of entries with |a - b| > tol."""
    a, b = _c(a, np.float32).ravel(), _c(b, np.float32).ravel()
    return int(lib().oracle_count_abs_errors(a.shape[0], a, b, tol))


def count_rel_errors(y, y64, abs_sum, rel=1e-5):
    """Rows failing |y - y64| <= rel * sum|a_ij x_j|; returns (count, worst)."""
    y = _c(y, np.float32).ravel()
    y64, abs_sum = _c(y64, np.float64).ravel(), _c(abs_sum, np.float64).ravel()
    worst = ctypes.c_double(0.0)
    n = lib().oracle_count_rel_errors(y.shape[0], y, y64, abs_sum, rel, ctypes.byref(worst))
    return int(n), float(worst.value)


# --------------------------------------------------------------------------
# Partitions
# --------------------------------------------------------------------------
def mp_search(offsets, diagonal):
    """util/search.hxx:34-60: merge-path coordinate (x, y) at a diagonal."""
    offsets = _c(offsets, np.int32)
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    x, y = ctypes.c_uint32(), ctypes.c_uint32()
    lib().oracle_mp_search(offsets, rows, nnz, int(diagonal), ctypes.byref(x), ctypes.byref(y))
    return int(x.value), int(y.value)


def mp_tile_coords(offsets, items_per_tile, clamp=True):
    """merge_path_flat.hxx:108-138: T+1 tile coordinates, shape (T+1, 2)."""
    offsets = _c(offsets, np.int32)
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    T = -(-(rows + nnz) // items_per_tile)
    xy = np.empty((T + 1, 2), np.uint32)
    fn = lib().oracle_mp_tile_coords if clamp else lib().oracle_mp_tile_coords_unclamped
    fn(offsets, rows, nnz, items_per_tile, T + 1, xy)
    return xy


def mp_thread_coords(offsets, threads_per_block, items_per_thread, tile):
    """merge_path_flat.hxx:321-385: per-thread start coordinates of a tile."""
    offsets = _c(offsets, np.int32)
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    xy = np.empty((threads_per_block, 2), np.uint32)
    lib().oracle_mp_thread_coords(offsets, rows, nnz, threads_per_block,
                                  items_per_thread, tile, xy)
    return xy


def mp_walk(offsets):
    """Brute-force merge walk: MP(d) for every d in [0, rows+nnz]."""
    offsets = _c(offsets, np.int32)
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    xy = np.empty((rows + nnz + 1, 2), np.uint32)
    lib().oracle_mp_walk(offsets, rows, nnz, xy)
    return xy


def wo_coords(offsets, num_threads):
    """work_oriented.hxx:146-167: (start_x, start_y, end_x, end_y) per thread."""
    offsets = _c(offsets, np.int32)
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    xy = np.empty((num_threads, 4), np.uint32)
    lib().oracle_wo_coords(offsets, rows, nnz, num_threads, xy)
    return xy


def gm_atom_map(offsets, threads_per_tile):
    """group_mapped.hxx:146-213: (group, virtual atom) of every nonzero."""
    offsets = _c(offsets, np.int32)
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    g = np.full(nnz, -1, np.int64)
    v = np.full(nnz, -1, np.int64)
    lib().oracle_gm_atom_map(offsets, rows, threads_per_tile, g, v)
    return g, v


# --------------------------------------------------------------------------
# Matrix Market
# --------------------------------------------------------------------------
def mtx_read(path):
    """container/market.hxx:99-227 + csr.hxx:86-94: (rows, cols, offsets,
    indices, values) with symmetric expansion and (row, col) order."""
    L = lib()
    r, c, n = _i64(), _i64(), _i64()
    p = str(path).encode()
    rc = L.oracle_mtx_read(p, ctypes.byref(r), ctypes.byref(c), ctypes.byref(n), None, None, None)
    if rc != 0:
        raise ValueError(f"oracle_mtx_read({path}) failed with code {rc}")
    offsets = np.empty(r.value + 1, np.int32)
    indices = np.empty(n.value, np.int32)
    values = np.empty(n.value, np.float32)
    rc = L.oracle_mtx_read(p, ctypes.byref(r), ctypes.byref(c), ctypes.byref(n),
                           offsets.ctypes.data, indices.ctypes.data, values.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_mtx_read({path}) failed with code {rc}")
    return int(r.value), int(c.value), offsets, indices, values
