"""Argument checks of the Python layer, made before any pointer reaches the
C ABI.  The library receives raw device addresses; a host tensor, a wrong
dtype, a transposed operand or a short vector would otherwise become an
out-of-bounds device access or a silently wrong result.  Every wrapper
raises ValueError here instead.

The checks look at tensor metadata only (device, dtype, shape, strides), so
they cost nothing on the device.  Every wrapper checks dtypes, shapes and
strides of all its arguments first and their devices last (check_devices),
so each rejection can be exercised with host tensors.
"""
from __future__ import annotations

_VALUE_DTYPES = {"f32": "float32", "bf16": "bfloat16", "f16": "float16", "f64": "float64",
                 "i32": "int32"}


def _dtype_name(t):
    return str(t.dtype).replace("torch.", "")


def _device_ok(t, device=None, name="tensor"):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) device tensor, got {t.device}")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, the matrix on {device}")


def check_csr(csr):
    """offsets int32 [rows + 1], indices int32 [nnzs], values [nnzs] of the
    storage type, all contiguous (devices: check_devices)."""
    if csr.rows < 0 or csr.cols < 0:
        raise ValueError("rows and cols must be non-negative")
    for name, t in (("offsets", csr.offsets), ("indices", csr.indices)):
        if _dtype_name(t) != "int32":
            raise ValueError(f"{name} must be int32, got {_dtype_name(t)}")
    want = _VALUE_DTYPES[csr.value_type]
    if _dtype_name(csr.values) != want:
        raise ValueError(f"values must be {want} for value_type {csr.value_type!r}, "
                         f"got {_dtype_name(csr.values)}")
    if csr.offsets.dim() != 1 or csr.offsets.numel() != csr.rows + 1:
        raise ValueError(f"offsets must hold rows + 1 = {csr.rows + 1} entries")
    if csr.indices.dim() != 1 or csr.values.dim() != 1 or \
            csr.values.numel() != csr.indices.numel():
        raise ValueError("indices and values must be 1-D with nnzs entries each")
    for name, t in (("offsets", csr.offsets), ("indices", csr.indices),
                    ("values", csr.values)):
        if t.numel() > 1 and t.stride(0) != 1:
            raise ValueError(f"{name} must be contiguous")


def check_devices(csr, **tensors):
    """The matrix's arrays and every named tensor (None skipped) on one HIP
    device."""
    _device_ok(csr.offsets, None, "offsets")
    _device_ok(csr.indices, csr.offsets.device, "indices")
    _device_ok(csr.values, csr.offsets.device, "values")
    for name, t in tensors.items():
        if t is not None:
            _device_ok(t, csr.offsets.device, name)


def check_on(device, **tensors):
    """Every named tensor on the given HIP device."""
    for name, t in tensors.items():
        _device_ok(t, device, name)


def check_vector(t, n, name, dtype="float32"):
    """1-D of the dtype (fp32 unless stated), unit stride, at least n
    entries."""
    if _dtype_name(t) != dtype:
        raise ValueError(f"{name} must be {dtype}, got {_dtype_name(t)}")
    if t.dim() != 1:
        raise ValueError(f"{name} must be 1-D")
    if t.numel() < n:
        raise ValueError(f"{name} has {t.numel()} entries, needs {n}")
    if t.numel() > 1 and t.stride(0) != 1:
        raise ValueError(f"{name} must have unit stride")


def check_dense(t, rows, k, name):
    """fp32 row-major (rows x k view with unit column stride, row stride >=
    k), at least `rows` rows."""
    if _dtype_name(t) != "float32":
        raise ValueError(f"{name} must be float32, got {_dtype_name(t)}")
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D (rows x k)")
    if t.shape[1] != k:
        raise ValueError(f"{name} has {t.shape[1]} columns, expected {k}")
    if t.shape[0] < rows:
        raise ValueError(f"{name} has {t.shape[0]} rows, needs {rows}")
    if k > 0 and t.shape[0] > 0 and (t.stride(1) != 1 or t.stride(0) < k):
        raise ValueError(f"{name} must be row-major with unit column stride "
                         f"(got strides {tuple(t.stride())})")


def check_workspace(ws):
    if ws is not None and ws.numel() > 1 and not ws.is_contiguous():
        raise ValueError("workspace must be contiguous")


def check_rhs(csr, x, name="x"):
    """A right-hand side: an fp32 vector of cols entries (k = 1) or a
    cols x k row-major block; returns k."""
    if x.dim() == 1:
        check_vector(x, csr.cols, name)
        return 1
    if x.dim() != 2:
        raise ValueError(f"{name} must be 2-D (cols x k) or a vector")
    check_dense(x, csr.cols, int(x.shape[1]), name)
    return int(x.shape[1])


def check_out(csr, y, k, name="y"):
    if k == 1 and y.dim() == 1:
        check_vector(y, csr.rows, name)
    else:
        check_dense(y, csr.rows, k, name)
