"""Serving: continuous batching of SpMV / SpMM requests against CSR matrices
resident in HBM (the loops_server_* C ABI).  No counterpart in the
reference.

    srv = Server()                       # batched; Server(mode="plain") = per-call
    mid = srv.add_matrix(A)              # tunings from the autotuner's table
    t = srv.submit(mid, X)               # X: fp32 device vector (k = 1) or cols x k
    Y = srv.result(t)                    # waits, copies the output out, releases
    lat, span = srv.replay(requests, arrivals)   # open-loop load, seconds

The server thread admits requests in arrival order as device memory for
their outputs frees up, and launches everything admitted as one batch on a
free stream while earlier batches still run.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi, hip, tune
from . import _validate as _v

MODES = {"batched": _capi.SERVER_BATCHED, "plain": _capi.SERVER_PLAIN}


def matrix_tunings(csr, table=None, max_row=None):
    """loops_server_matrix_t for a device csr_t: SpMV schedule and tile, and
    SpMM steps per unit for each k bucket, from the tuning table."""
    table = table or tune.Table.load()
    if max_row is None:
        import torch
        max_row = int(torch.diff(csr.offsets).max().item()) if csr.rows else 0
    sched, ipt = table.pick(csr.rows, csr.nnzs, max_row, 1, allow_blocks=False)  # no plans here
    m = _capi.ServerMatrix()
    m.A = csr.view()
    m.spmv_schedule = _capi.SCHEDULES[sched]
    m.spmv_items_per_thread = ipt
    for b, k in enumerate((2, 4, 8, 16, 32, 64, 128, 256)):
        m.spmm_items_per_unit[b] = table.pick(csr.rows, csr.nnzs, max_row, k)[1]
    return m


class Server:
    def __init__(self, mode="batched", arena_bytes=0, scratch_bytes=0, slots=0,
                 max_batch_requests=0, reference_schedules=False, batch_all=False):
        """reference_schedules (plain mode only): every request through the
        reference's own entry points (merge_path_flat SpMV, thread_mapped
        SpMM) instead of the tuned merge-path kernels.  batch_all (batched
        mode only): every admitted request in the slot's batch, none launched
        alone."""
        flags = _capi.SERVER_FLAG_REFERENCE_SCHEDULES if reference_schedules else 0
        flags |= _capi.SERVER_FLAG_BATCH_ALL if batch_all else 0
        cfg = _capi.ServerConfig(int(arena_bytes), int(scratch_bytes), int(slots),
                                 int(max_batch_requests), MODES[mode], flags)
        self.ptr = ctypes.c_void_p()
        _capi.check(_capi.lib().loops_server_create(ctypes.byref(cfg), ctypes.byref(self.ptr)),
                    "loops_server_create")
        self.mode = mode
        self.matrices = []   # keeps the device CSR alive while the server runs
        self.inputs = {}     # ticket -> x (alive until the request completes)

    def add_matrix(self, csr, table=None, max_row=None, tunings=None):
        _v.check_csr(csr)
        _v.check_devices(csr)
        m = tunings if tunings is not None else matrix_tunings(csr, table, max_row)
        mid = ctypes.c_int32()
        _capi.check(_capi.lib().loops_server_add_matrix(self.ptr, ctypes.byref(m),
                                                        ctypes.byref(mid)),
                    "loops_server_add_matrix")
        self.matrices.append((csr, m))
        return mid.value

    def submit(self, matrix, x, stream=None):
        """Queue Y = A x.  x (fp32 device vector, or cols x k with unit
        column stride) may still be in production on `stream` (default:
        torch's current stream): an event recorded there now orders the
        server's read of x after it."""
        import torch
        if 0 <= matrix < len(self.matrices):  # an unknown id is the library's to reject
            _v.check_rhs(self.matrices[matrix][0], x)
            _v.check_devices(self.matrices[matrix][0], x=x)
        k = 1 if x.dim() == 1 else int(x.shape[1])
        ldx = 1 if x.dim() == 1 else int(x.stride(0))
        s = torch.cuda.current_stream(x.device) if stream is None else stream
        ready = hip.Event()
        ready.record(s.cuda_stream if hasattr(s, "cuda_stream") else s)
        t = ctypes.c_int64()
        _capi.check(_capi.lib().loops_server_submit(self.ptr, int(matrix), k, x.data_ptr(),
                                                    ldx, ready.ptr, ctypes.byref(t)),
                    "loops_server_submit")
        self.inputs[t.value] = (matrix, k, x, ready)
        return t.value

    def wait(self, ticket):
        """(device address of the output, submit time, completion time)."""
        y = ctypes.c_void_p()
        s, d = ctypes.c_double(), ctypes.c_double()
        _capi.check(_capi.lib().loops_server_wait(self.ptr, int(ticket), ctypes.byref(y),
                                                  ctypes.byref(s), ctypes.byref(d)),
                    "loops_server_wait")
        return y.value, s.value, d.value

    def release(self, ticket):
        _capi.check(_capi.lib().loops_server_release(self.ptr, int(ticket)),
                    "loops_server_release")
        self.inputs.pop(ticket, None)

    def result(self, ticket):
        """Wait for a request, copy its output to a new device tensor, release."""
        import torch
        matrix, k = self.inputs[ticket][:2]
        rows = self.matrices[matrix][0].rows
        yp, _, _ = self.wait(ticket)
        out = torch.empty((rows,) if k == 1 else (rows, k), device="cuda")
        if out.numel():
            hip.memcpy(out.data_ptr(), yp, out.numel() * 4)
            # a device-to-device hipMemcpy may return before the copy is
            # done; the arena region must not be released (and handed to
            # the next admitted request) while it is still being read
            hip.stream_synchronize(None)
        self.release(ticket)
        return out

    def replay(self, requests, arrivals):
        """Open-loop load: requests = [(matrix, x)], arrivals in seconds
        (non-decreasing).  Returns (latencies in s, makespan in s)."""
        import torch
        n = len(requests)
        arr = (_capi.ServerRequest * max(n, 1))()
        for m, x in requests:
            if 0 <= m < len(self.matrices):
                _v.check_rhs(self.matrices[m][0], x)
                _v.check_devices(self.matrices[m][0], x=x)
        torch.cuda.synchronize()  # the replayed inputs are complete on the device
        for i, (m, x) in enumerate(requests):
            k = 1 if x.dim() == 1 else int(x.shape[1])
            arr[i] = _capi.ServerRequest(int(m), k, x.data_ptr(),
                                         1 if x.dim() == 1 else int(x.stride(0)))
        a = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n, np.float64)
        span = ctypes.c_double()
        _capi.check(_capi.lib().loops_server_replay(
            self.ptr, arr, n, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(span)),
            "loops_server_replay")
        return lat, span.value

    def stats(self):
        st = _capi.ServerStats()
        _capi.check(_capi.lib().loops_server_stats(self.ptr, ctypes.byref(st)),
                    "loops_server_stats")
        return {n: getattr(st, n) for n, _ in st._fields_}

    def close(self):
        if self.ptr:
            _capi.lib().loops_server_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(n, rate, seed):
    """Arrival times (s) of a seeded Poisson process with the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n)) - 0.0
