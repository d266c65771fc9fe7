"""Device containers mirroring the reference's csr_t / vector_t / matrix_t
(reference include/loops/container/csr.hxx:32-95, vector.hxx:31-42,
matrix.cuh:8-51).

Device memory is held in torch tensors (plumbing only: allocation and
streams); all arithmetic runs in the library's HIP kernels.
"""
from __future__ import annotations

import numpy as np

from . import _capi

_VTYPES = {"f32": _capi.VALUE_F32, "bf16": _capi.VALUE_BF16, "f16": _capi.VALUE_F16,
           "f64": _capi.VALUE_F64, "i32": _capi.VALUE_I32}


def _torch():
    import torch
    return torch


class csr_t:
    """CSR matrix on the device: rows, cols, nnzs and the arrays offsets
    (int32, rows+1), indices (int32, nnzs) and values (nnzs; fp32 by default,
    or bf16 / fp16 storage with fp32 accumulation, or float64 / int32 values
    for the typed SpMV, run_spmv_typed)."""

    def __init__(self, rows, cols, offsets, indices, values, value_type="f32"):
        self.rows = int(rows)
        self.cols = int(cols)
        self.offsets = offsets
        self.indices = indices
        self.values = values
        self.nnzs = int(indices.numel())
        if value_type not in _VTYPES:
            raise ValueError(f"value_type must be one of {sorted(_VTYPES)}")
        self.value_type = value_type

    @classmethod
    def from_host(cls, rows, cols, offsets, indices, values, device="cuda",
                  value_type="f32"):
        """Upload host (numpy) CSR arrays.  values are given in fp32 and
        converted to the storage type on the way (round to nearest even)."""
        torch = _torch()
        off = torch.from_numpy(np.ascontiguousarray(offsets, dtype=np.int32)).to(device)
        idx = torch.from_numpy(np.ascontiguousarray(indices, dtype=np.int32)).to(device)
        if value_type in ("f64", "i32"):
            dt = np.float64 if value_type == "f64" else np.int32
            val = torch.from_numpy(np.ascontiguousarray(values, dtype=dt)).to(device)
            return cls(rows, cols, off, idx, val, value_type)
        val = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float32)).to(device)
        if value_type == "bf16":
            val = val.to(torch.bfloat16)
        elif value_type == "f16":
            val = val.to(torch.float16)
        return cls(rows, cols, off, idx, val, value_type)

    def view(self):
        """The loops_csr_view_t handed to the C ABI (non-owning)."""
        return _capi.CsrView(
            self.rows, self.cols, self.nnzs,
            self.offsets.data_ptr() if self.offsets.numel() else None,
            self.indices.data_ptr() if self.indices.numel() else None,
            self.values.data_ptr() if self.values.numel() else None,
            _VTYPES[self.value_type])

    def to_host(self):
        vals = self.values.float().cpu().numpy()
        return (self.offsets.cpu().numpy(), self.indices.cpu().numpy(), vals)


def vector_t(n, device="cuda", dtype=None):
    """Zero-initialised fp32 device vector (a thrust::device_vector<float>)."""
    torch = _torch()
    return torch.zeros(int(n), dtype=dtype or torch.float32, device=device)


def matrix_t(rows, cols, device="cuda"):
    """Zero-initialised row-major fp32 device matrix; element (r, c) is at
    r * cols + c."""
    torch = _torch()
    return torch.zeros((int(rows), int(cols)), dtype=torch.float32, device=device)
