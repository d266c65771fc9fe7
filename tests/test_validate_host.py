"""The Python layer rejects malformed arguments with ValueError before any
pointer reaches the C ABI (host tensors only: no GPU is needed, and no
library call is made).  Checks run dtype, then shape and strides, then
device, so each rejection below is the one named."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import _validate as v  # noqa: E402
from edge_loops_amd.batch import BatchItem  # noqa: E402


def _csr(rows=4, cols=5, off_dtype=torch.int32, idx_dtype=torch.int32,
         val_dtype=torch.float32, value_type="f32", off_len=None):
    off = torch.tensor([0, 1, 3, 3, 4], dtype=off_dtype)
    if off_len is not None:
        off = off[:off_len]
    idx = torch.tensor([0, 1, 4, 2], dtype=idx_dtype)
    val = torch.tensor([1.0, 2.0, 3.0, 4.0]).to(val_dtype)
    return loops.csr_t(rows, cols, off, idx, val, value_type=value_type)


def _raises(fn, match):
    with pytest.raises(ValueError, match=match):
        fn()


def test_csr_rejections():
    x, y = torch.zeros(5), torch.zeros(4)
    _raises(lambda: loops.run_spmv("merge_path_flat", _csr(off_dtype=torch.int64), x, y),
            "offsets must be int32")
    _raises(lambda: loops.run_spmv("merge_path_flat", _csr(idx_dtype=torch.int64), x, y),
            "indices must be int32")
    _raises(lambda: loops.run_spmv("merge_path_flat", _csr(val_dtype=torch.float64), x, y),
            "values must be float32")
    _raises(lambda: loops.run_spmv("merge_path_flat",
                                   _csr(val_dtype=torch.float32, value_type="bf16"), x, y),
            "values must be bfloat16")
    _raises(lambda: loops.run_spmv("merge_path_flat", _csr(off_len=4), x, y),
            "rows \\+ 1")
    # everything well formed, but in host memory
    _raises(lambda: loops.run_spmv("merge_path_flat", _csr(), x, y), "device tensor")


def test_vector_rejections(monkeypatch):
    A = _csr()
    # let the CSR pass so that the vector checks are reached
    monkeypatch.setattr(v, "check_csr", lambda csr: None)
    _raises(lambda: loops.run_spmv("merge_path_flat", A, torch.zeros(5, dtype=torch.float64),
                                   torch.zeros(4)), "x must be float32")
    _raises(lambda: loops.run_spmv("merge_path_flat", A, torch.zeros(4), torch.zeros(4)),
            "x has 4 entries, needs 5")
    _raises(lambda: loops.run_spmv("merge_path_flat", A, torch.zeros(10)[::2], torch.zeros(4)),
            "x must have unit stride")
    _raises(lambda: loops.run_spmv("merge_path_flat", A, torch.zeros(5), torch.zeros(3)),
            "y has 3 entries, needs 4")
    _raises(lambda: loops.run_spmv("merge_path_flat", A, torch.zeros(5), torch.zeros(2, 2)),
            "y must be 1-D")
    _raises(lambda: loops.run_spmv("merge_path_flat", A, torch.zeros(5), torch.zeros(4)),
            "device tensor")


def test_dense_rejections(monkeypatch):
    A = _csr()
    monkeypatch.setattr(v, "check_csr", lambda csr: None)
    B, C = torch.zeros(5, 3), torch.zeros(4, 3)
    _raises(lambda: loops.run_spmm("merge_path_flat", A, torch.zeros(3, 5).t(), C),
            "B must be row-major")
    _raises(lambda: loops.run_spmm("merge_path_flat", A, torch.zeros(4, 3), C),
            "B has 4 rows, needs 5")
    _raises(lambda: loops.run_spmm("merge_path_flat", A, B, torch.zeros(4, 2)),
            "C has 2 columns, expected 3")
    _raises(lambda: loops.run_spmm("merge_path_flat", A, B, torch.zeros(4, 3).double()),
            "C must be float32")
    _raises(lambda: loops.run_spmm("merge_path_flat", A, torch.zeros(5), C), "B must be 2-D")
    _raises(lambda: loops.run_spmm("merge_path_flat", A, B, C), "device tensor")
    # a padded leading dimension is fine up to the device check
    Bp = torch.zeros(5, 8)[:, :3]
    _raises(lambda: loops.run_spmm("merge_path_flat", A, Bp, C), "device tensor")


def test_batch_and_rhs_rejections(monkeypatch):
    A = _csr()
    monkeypatch.setattr(v, "check_csr", lambda csr: None)
    _raises(lambda: BatchItem(A, torch.zeros(5, dtype=torch.float16), torch.zeros(4)).to_c(),
            "x must be float32")
    _raises(lambda: BatchItem(A, torch.zeros(5, 2), torch.zeros(4)).to_c(),
            "y must be 2-D")
    _raises(lambda: v.check_rhs(A, torch.zeros(2, 5).t()), "x must be row-major")
    _raises(lambda: v.check_rhs(A, torch.zeros(5, 2, 1)), "x must be 2-D")


def test_plan_and_partition_reject_host_csr():
    A = _csr()
    _raises(lambda: loops.spmv_plan(A), "device tensor")
    _raises(lambda: loops.merge_path_partition(A, 640), "device tensor")
    _raises(lambda: loops.work_oriented_partition(A, 64), "device tensor")
    assert np.array_equal(A.offsets.numpy(), [0, 1, 3, 3, 4])
