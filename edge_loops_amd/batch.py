"""Batches of independent SpMV / SpMM products in one persistent launch
(the loops_batch_* C ABI).  No counterpart in the reference.

    items = [BatchItem(A0, x0, y0), BatchItem(A1, X1, Y1), ...]
    b = Batch(items)          # host table -> device, workspace
    b.launch()                # asynchronous; allocates nothing
    g = b.capture()           # the launch as a CUDA/HIP graph; g.replay()

x / y are 1-D fp32 device tensors (SpMV) or 2-D row-major (SpMM, k =
columns).  An item's result does not depend on the rest of the batch.

An SpMV item may carry a row-block plan of its matrix
(``BatchItem(A, x, y, "blocks", plan=spmv_plan(A, layout="blocks"))``):
it then runs the plan's streams and gives what ``plan.execute`` gives, bit
for bit.  A plan may appear in one item of a batch only, and a batch's
launches must not overlap other executions of its plans (they share the
plan's partial sums).
"""
from __future__ import annotations

import ctypes

from . import _capi
from . import _validate as _v


class BatchItem:
    def __init__(self, csr, x, y, schedule="merge_path_flat", items_per_thread=0, plan=None):
        if schedule == "blocks" and plan is None:
            raise ValueError("a 'blocks' batch item needs a row-block plan (plan=)")
        self.csr, self.x, self.y = csr, x, y
        self.schedule = schedule
        self.items_per_thread = int(items_per_thread)
        self.plan = plan  # kept alive with the item

    def to_c(self):
        _v.check_csr(self.csr)
        _v.check_out(self.csr, self.y, _v.check_rhs(self.csr, self.x, "x"), "y")
        _v.check_devices(self.csr, x=self.x, y=self.y)
        k = 1 if self.x.dim() == 1 else int(self.x.shape[1])
        ldx = 1 if self.x.dim() == 1 else int(self.x.stride(0))
        ldy = 1 if self.y.dim() == 1 else int(self.y.stride(0))
        plan = None
        if self.plan is not None:
            if self.plan.layout != "blocks":
                raise ValueError(f"batch items take row-block plans, not {self.plan.layout!r}")
            if k != 1:
                raise ValueError("a plan item is an SpMV (1-D x and y)")
            if (self.plan.rows, self.plan.cols) != (self.csr.rows, self.csr.cols):
                raise ValueError("the plan was built for a matrix of another shape")
            if not getattr(self.plan, "_h", None):
                raise ValueError("the plan has been destroyed")
            plan = self.plan._h.value
        sched = self.schedule
        if sched == "blocks":
            sched = "merge_path_flat"  # not read for plan items
        sched = _capi.SCHEDULES[sched] if isinstance(sched, str) else sched
        return _capi.BatchItem(self.csr.view(), k,
                               self.x.data_ptr() if self.x.numel() else None, ldx,
                               self.y.data_ptr() if self.y.numel() else None, ldy,
                               sched, self.items_per_thread, plan)


def query(items):
    arr = (_capi.BatchItem * max(len(items), 1))(*[it.to_c() for it in items])
    info = _capi.BatchInfo()
    _capi.check(_capi.lib().loops_batch_query(arr, len(items), ctypes.byref(info)),
                "loops_batch_query")
    return arr, info


class Batch:
    """Prepared batch.  Buffers may be supplied (host_table must be pinned
    for an asynchronous upload); otherwise they are allocated here."""

    def __init__(self, items, stream=None, host_table=None, device_table=None,
                 workspace=None, device="cuda"):
        import torch
        self.items = list(items)
        arr, info = query(self.items)
        self.info = info
        tb, wb = int(info.table_bytes), int(info.workspace_bytes)
        if host_table is None:
            host_table = torch.empty(tb, dtype=torch.uint8, pin_memory=True)
        if device_table is None:
            device_table = torch.empty(tb, dtype=torch.uint8, device=device)
        if workspace is None:
            workspace = torch.empty(max(wb, 256), dtype=torch.uint8, device=device)
        if host_table.numel() < tb or device_table.numel() < tb or workspace.numel() < wb:
            raise ValueError("batch buffers too small")
        self.host_table, self.device_table, self.workspace = host_table, device_table, workspace
        _capi.check(_capi.lib().loops_batch_prepare(
            arr, len(self.items), workspace.data_ptr(), workspace.numel(),
            host_table.data_ptr(), host_table.numel(), ctypes.byref(self.info)),
            "loops_batch_prepare")
        s = torch.cuda.current_stream() if stream is None else stream
        with torch.cuda.stream(s):
            device_table[:tb].copy_(host_table[:tb], non_blocking=True)

    def launch(self, stream=None, grid=0):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else (
            stream.cuda_stream if hasattr(stream, "cuda_stream") else stream)
        _capi.check(_capi.lib().loops_batch_launch(
            self.device_table.data_ptr(), ctypes.byref(self.info), int(grid), s),
            "loops_batch_launch")

    def capture(self, grid=0):
        """The launch captured as a graph (torch.cuda.CUDAGraph)."""
        import torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.launch(stream=s, grid=grid)
        return g
