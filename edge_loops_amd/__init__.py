"""edge_loops_amd -- MI355X (gfx950) load-balanced CSR SpMV/SpMM.

Python mirror of the reference's host API (loops::algorithms::spmv / spmm,
csr_t / vector_t / matrix_t) over the C ABI of libloops_mi355x.so
(include/loops_capi.h).  Device memory and streams come from torch; every
result is computed by the library's HIP kernels.
"""
from . import _capi, algorithms
from ._capi import LoopsError, lib
from .algorithms import (csr_check, merge_path_partition, run_spmm, run_spmv, run_spmv_typed, spmm,
                         spmm_workspace_size, spmv, spmv_plan, spmv_typed_workspace_size,
                         spmv_workspace_size, work_oriented_partition, device_cu_count)
from .containers import csr_t, matrix_t, vector_t

__all__ = [
    "LoopsError", "lib", "csr_t", "vector_t", "matrix_t", "spmv", "spmm",
    "run_spmv", "run_spmm", "run_spmv_typed", "spmv_workspace_size", "spmm_workspace_size",
    "spmv_typed_workspace_size",
    "spmv_plan", "merge_path_partition", "work_oriented_partition", "device_cu_count",
    "csr_check",
]
