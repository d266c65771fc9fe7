"""Host entry points mirroring the reference's loops::algorithms namespace.

    spmv.original / thread_mapped / group_mapped / work_oriented (csr, x, y)
        algorithms/spmv/{original,thread_mapped,group_mapped,work_oriented}.cuh
    spmv.merge_path_flat(csr, x, y) -> timer_t
        algorithms/spmv/merge_path_flat.cuh:96-139
    spmm.thread_mapped(csr, B, C)      algorithms/spmm/thread_mapped.cuh:67-90
    spmm.group_mapped / merge_path     (new: the reference has no such SpMM)

Like the reference, each entry point launches on a stream (default: torch's
current stream) and synchronises it before returning; ``run_spmv`` /
``run_spmm`` are the asynchronous forms with explicit tuning and workspace.
Unlike the reference, y and C are overwritten rather than accumulated into,
so they need not be zeroed first.
"""
from __future__ import annotations

import ctypes

from . import _capi
from . import _validate as _v
from .hip import Event, stream_synchronize, timer_t


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream


def _ptr(t):
    return t.data_ptr() if t is not None and t.numel() else None


def make_params(threads_per_block=0, items_per_thread=0, threads_per_tile=0,
                grid_size=0, tile_coordinates=None, start_event=None,
                stop_event=None):
    return _capi.Params(
        int(threads_per_block), int(items_per_thread), int(threads_per_tile),
        int(grid_size),
        _ptr(tile_coordinates) if tile_coordinates is not None else None,
        start_event.ptr if start_event is not None else None,
        stop_event.ptr if stop_event is not None else None, 0)


def workspace(schedule, csr, stream=None, **tuning):
    """A device workspace sized for loops_spmv (loops_spmv_typed for float64
    / int32 matrices); contents need no initialisation."""
    import torch
    size = spmv_typed_workspace_size if csr.value_type in ("f64", "i32") else spmv_workspace_size
    n = max(size(schedule, csr, **tuning), 4)
    return torch.empty((n + 3) // 4 * 4, dtype=torch.uint8, device=csr.offsets.device)


def spmv_workspace_size(schedule, csr, **tuning):
    sched = _capi.SCHEDULES[schedule] if isinstance(schedule, str) else schedule
    n = ctypes.c_size_t()
    v = csr.view()
    p = make_params(**tuning)
    _capi.check(_capi.lib().loops_spmv_workspace_size(
        _capi.handle(), sched, ctypes.byref(v), ctypes.byref(p), ctypes.byref(n)),
        "loops_spmv_workspace_size")
    return int(n.value)


def run_spmv(schedule, csr, x, y, stream=None, workspace=None, **tuning):
    """Asynchronous y = A x through loops_spmv (no synchronisation); a
    matrix of float64 or int32 values goes through loops_spmv_typed."""
    if csr.value_type in ("f64", "i32"):
        return run_spmv_typed(schedule, csr, x, y, stream, workspace, **tuning)
    sched = _capi.SCHEDULES[schedule] if isinstance(schedule, str) else schedule
    _v.check_csr(csr)
    _v.check_vector(x, csr.cols, "x")
    _v.check_vector(y, csr.rows, "y")
    _v.check_workspace(workspace)
    _v.check_devices(csr, x=x, y=y, workspace=workspace)
    v = csr.view()
    p = make_params(**tuning)
    ws = _ptr(workspace) if workspace is not None else None
    wsb = workspace.numel() * workspace.element_size() if workspace is not None else 0
    _capi.check(_capi.lib().loops_spmv(
        _capi.handle(), sched, ctypes.byref(v), _ptr(x), _ptr(y), ctypes.byref(p),
        ws, wsb, _stream(stream)), f"loops_spmv({schedule})")


def spmv_typed_workspace_size(schedule, csr, **tuning):
    sched = _capi.SCHEDULES[schedule] if isinstance(schedule, str) else schedule
    n = ctypes.c_size_t()
    v = csr.view()
    p = make_params(**tuning)
    _capi.check(_capi.lib().loops_spmv_typed_workspace_size(
        _capi.handle(), sched, ctypes.byref(v), ctypes.byref(p), ctypes.byref(n)),
        "loops_spmv_typed_workspace_size")
    return int(n.value)


def run_spmv_typed(schedule, csr, x, y, stream=None, workspace=None, **tuning):
    """Asynchronous y = A x through loops_spmv_typed: A's values, x and y
    all float64 (value_type "f64") or all int32 ("i32", wrap-around)."""
    sched = _capi.SCHEDULES[schedule] if isinstance(schedule, str) else schedule
    if csr.value_type not in ("f64", "i32"):
        raise ValueError("run_spmv_typed takes value_type 'f64' or 'i32' (use run_spmv)")
    dt = "float64" if csr.value_type == "f64" else "int32"
    _v.check_csr(csr)
    _v.check_vector(x, csr.cols, "x", dt)
    _v.check_vector(y, csr.rows, "y", dt)
    _v.check_workspace(workspace)
    _v.check_devices(csr, x=x, y=y, workspace=workspace)
    v = csr.view()
    p = make_params(**tuning)
    ws = _ptr(workspace) if workspace is not None else None
    wsb = workspace.numel() * workspace.element_size() if workspace is not None else 0
    _capi.check(_capi.lib().loops_spmv_typed(
        _capi.handle(), sched, ctypes.byref(v), _ptr(x), _ptr(y), ctypes.byref(p),
        ws, wsb, _stream(stream)), f"loops_spmv_typed({schedule})")


def spmm_workspace_size(schedule, csr, k, **tuning):
    sched = _capi.SCHEDULES[schedule] if isinstance(schedule, str) else schedule
    n = ctypes.c_size_t()
    v = csr.view()
    p = make_params(**tuning)
    _capi.check(_capi.lib().loops_spmm_workspace_size(
        _capi.handle(), sched, ctypes.byref(v), int(k), ctypes.byref(p),
        ctypes.byref(n)), "loops_spmm_workspace_size")
    return int(n.value)


def run_spmm(schedule, csr, B, C, stream=None, workspace=None, **tuning):
    """Asynchronous C = A B (row-major fp32 B: cols x k, C: rows x k)."""
    sched = _capi.SCHEDULES[schedule] if isinstance(schedule, str) else schedule
    _v.check_csr(csr)
    if B.dim() != 2:
        raise ValueError("B must be 2-D (cols x k)")
    k = int(B.shape[1])
    _v.check_dense(B, csr.cols, k, "B")
    _v.check_dense(C, csr.rows, k, "C")
    _v.check_workspace(workspace)
    _v.check_devices(csr, B=B, C=C, workspace=workspace)
    v = csr.view()
    p = make_params(**tuning)
    ws = _ptr(workspace) if workspace is not None else None
    wsb = workspace.numel() * workspace.element_size() if workspace is not None else 0
    _capi.check(_capi.lib().loops_spmm(
        _capi.handle(), sched, ctypes.byref(v), int(k), _ptr(B), B.stride(0),
        _ptr(C), C.stride(0), ctypes.byref(p), ws, wsb, _stream(stream)),
        f"loops_spmm({schedule})")


class spmv:
    """loops::algorithms::spmv."""

    @staticmethod
    def original(csr, x, y, stream=None, **tuning):
        s = _stream(stream)
        run_spmv("original", csr, x, y, s, **tuning)
        stream_synchronize(s)

    @staticmethod
    def thread_mapped(csr, x, y, stream=None, **tuning):
        s = _stream(stream)
        run_spmv("thread_mapped", csr, x, y, s, **tuning)
        stream_synchronize(s)

    @staticmethod
    def group_mapped(csr, x, y, stream=None, **tuning):
        s = _stream(stream)
        run_spmv("group_mapped", csr, x, y, s, **tuning)
        stream_synchronize(s)

    @staticmethod
    def work_oriented(csr, x, y, stream=None, **tuning):
        s = _stream(stream)
        run_spmv("work_oriented", csr, x, y, s, **tuning)
        stream_synchronize(s)

    @staticmethod
    def merge_path_flat(csr, x, y, stream=None, **tuning):
        """Returns a timer_t measuring the merge-path kernel, like the
        reference (merge_path_flat.cuh:121-138)."""
        s = _stream(stream)
        t = timer_t(s)
        run_spmv("merge_path_flat", csr, x, y, s, start_event=t.start_ev,
                 stop_event=t.stop_ev, **tuning)
        stream_synchronize(s)
        t.time = t.start_ev.elapsed_ms(t.stop_ev)
        return t

    merge_path = merge_path_flat


class spmm:
    """loops::algorithms::spmm."""

    @staticmethod
    def thread_mapped(csr, B, C, stream=None, **tuning):
        s = _stream(stream)
        run_spmm("thread_mapped", csr, B, C, s, **tuning)
        stream_synchronize(s)

    @staticmethod
    def group_mapped(csr, B, C, stream=None, **tuning):
        s = _stream(stream)
        run_spmm("group_mapped", csr, B, C, s, **tuning)
        stream_synchronize(s)

    @staticmethod
    def merge_path(csr, B, C, stream=None, **tuning):
        s = _stream(stream)
        run_spmm("merge_path_flat", csr, B, C, s, **tuning)
        stream_synchronize(s)


class spmv_plan:
    """An SpMV plan for one matrix (loops_spmv_plan_create).  layout
    "panels": merge-path tile coordinates computed once and, with panels >
    1, A copied into column panels that run on separate XCDs; "blocks": A
    copied into row blocks x column segments with y accumulated in LDS and
    x read in column order; "auto": blocks for matrices with at least 2^12
    columns and 2^19 nonzeros when no panels / threads_per_block /
    items_per_thread is given, else panels.  Both are deterministic.  ``execute(x, y)`` is
    asynchronous and allocates nothing.  ``for_batches=True`` cuts a
    blocks plan for batch items (edge_loops_amd.batch): one column
    segment when cols <= 2^20 and the fewest row blocks the LDS holds, so
    each workgroup streams many full rounds; alone such a plan underfills
    the GPU."""

    def __init__(self, csr, panels=0, threads_per_block=0, items_per_thread=0,
                 stream=None, layout="auto", for_batches=False):
        _v.check_csr(csr)
        _v.check_devices(csr)
        self.csr = csr  # keeps A alive: with one panel the plan reads it in place
        self.rows, self.cols = csr.rows, csr.cols
        self.device = csr.offsets.device
        v = csr.view()
        p = _capi.PlanParams(int(threads_per_block), int(items_per_thread), int(panels),
                             _capi.PLAN_LAYOUTS[layout], int(bool(for_batches)))
        h = ctypes.c_void_p()
        _capi.check(_capi.lib().loops_spmv_plan_create(
            _capi.handle(), ctypes.byref(v), ctypes.byref(p), _stream(stream),
            ctypes.byref(h)), "loops_spmv_plan_create")
        self._h = h
        n, b = ctypes.c_int32(), ctypes.c_int64()
        _capi.check(_capi.lib().loops_spmv_plan_info(h, ctypes.byref(n), ctypes.byref(b)),
                    "loops_spmv_plan_info")
        self.panels, self.device_bytes = int(n.value), int(b.value)
        lay = ctypes.c_int32()
        _capi.check(_capi.lib().loops_spmv_plan_layout(h, ctypes.byref(lay)),
                    "loops_spmv_plan_layout")
        self.layout = {v: k for k, v in _capi.PLAN_LAYOUTS.items()}[int(lay.value)]
        nb, comp, ns = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _capi.check(_capi.lib().loops_spmv_plan_blocks(h, ctypes.byref(nb), ctypes.byref(comp),
                                                       ctypes.byref(ns)),
                    "loops_spmv_plan_blocks")
        # row blocks layout: blocks, compressed stream rows, rows cut into pieces
        self.blocks, self.compressed, self.split_rows = int(nb.value), bool(comp.value), int(ns.value)

    def execute(self, x, y, stream=None, start_event=None, stop_event=None):
        """y = A x (asynchronous).  Executions of one plan must not overlap
        in time (another stream, or graph replays on several streams): they
        share the plan's partial sums."""
        _v.check_vector(x, self.cols, "x")
        _v.check_vector(y, self.rows, "y")
        _v.check_on(self.device, x=x, y=y)
        _capi.check(_capi.lib().loops_spmv_plan_execute(
            self._h, _ptr(x), _ptr(y),
            start_event.ptr if start_event is not None else None,
            stop_event.ptr if stop_event is not None else None, _stream(stream)),
            "loops_spmv_plan_execute")

    def destroy(self):
        if getattr(self, "_h", None):
            _capi.lib().loops_spmv_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def merge_path_partition(csr, items_per_tile, stream=None):
    """Tile coordinates (num_tiles + 1, 2) as int32 (x = row, y = nonzero):
    the reference's generate_search_coordinates
    (schedule/merge_path_flat.hxx:108-138)."""
    import torch
    _v.check_csr(csr)
    _v.check_devices(csr)
    T = ctypes.c_int64()
    _capi.check(_capi.lib().loops_merge_path_num_tiles(
        csr.rows, csr.nnzs, int(items_per_tile), ctypes.byref(T)),
        "loops_merge_path_num_tiles")
    coords = torch.empty((T.value + 1, 2), dtype=torch.int32, device=csr.offsets.device)
    _capi.check(_capi.lib().loops_merge_path_partition(
        _ptr(csr.offsets), csr.rows, csr.nnzs, int(items_per_tile), _ptr(coords),
        _stream(stream)), "loops_merge_path_partition")
    return coords


def work_oriented_partition(csr, num_threads, stream=None):
    """Per-thread (start, end) coordinates, shape (num_threads, 4)
    (schedule/work_oriented.hxx:150-167)."""
    import torch
    _v.check_csr(csr)
    _v.check_devices(csr)
    coords = torch.empty((int(num_threads), 4), dtype=torch.int32, device=csr.offsets.device)
    _capi.check(_capi.lib().loops_work_oriented_partition(
        _ptr(csr.offsets), csr.rows, csr.nnzs, int(num_threads), _ptr(coords),
        _stream(stream)), "loops_work_oriented_partition")
    return coords


CSR_FLAGS = {"bad_offsets": 1, "column_range": 2, "unsorted": 4}


def csr_check(csr, stream=None):
    """Structure of a device CSR matrix (loops_csr_check): the names of the
    LOOPS_CSR_* problems found, an empty list when it is well formed
    (offsets from 0 to nnz without decreasing, column ids in [0, cols) and
    ascending inside each row).  Plans need all three; loops_spmv_plan_create
    rejects a matrix that fails."""
    _v.check_csr(csr)
    _v.check_devices(csr)
    v = csr.view()
    f = ctypes.c_uint32()
    _capi.check(_capi.lib().loops_csr_check(ctypes.byref(v), ctypes.byref(f), _stream(stream)),
                "loops_csr_check")
    return [name for name, bit in CSR_FLAGS.items() if f.value & bit]


def device_cu_count():
    n = ctypes.c_int32()
    _capi.check(_capi.lib().loops_device_cu_count(_capi.handle(), ctypes.byref(n)),
                "loops_device_cu_count")
    return int(n.value)
