"""Minimal HIP runtime access (events, device sync) through ctypes, used for
timing on the exact stream the library launches on -- the analogue of the
reference's util::timer_t (include/loops/util/timer.hxx:19-52)."""
from __future__ import annotations

import ctypes
import ctypes.util

_hip = None


def runtime():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                             ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
    return _hip


def _ok(err, what):
    if err != 0:
        raise RuntimeError(f"{what} failed with hipError {err}")


class Event:
    """A hipEvent_t owned by Python."""

    def __init__(self):
        self.ptr = ctypes.c_void_p()
        _ok(runtime().hipEventCreate(ctypes.byref(self.ptr)), "hipEventCreate")

    def record(self, stream):
        _ok(runtime().hipEventRecord(self.ptr, stream), "hipEventRecord")

    def synchronize(self):
        _ok(runtime().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        t = ctypes.c_float()
        _ok(runtime().hipEventElapsedTime(ctypes.byref(t), self.ptr, end.ptr),
            "hipEventElapsedTime")
        return float(t.value)

    def __del__(self):
        try:
            if self.ptr:
                runtime().hipEventDestroy(self.ptr)
        except Exception:
            pass


class timer_t:
    """start()/stop() around work on a stream; milliseconds() afterwards."""

    def __init__(self, stream=None):
        self.stream = stream
        self.start_ev = Event()
        self.stop_ev = Event()
        self.time = 0.0

    def start(self):
        self.start_ev.record(self.stream)

    def stop(self):
        self.stop_ev.record(self.stream)
        self.stop_ev.synchronize()
        self.time = self.start_ev.elapsed_ms(self.stop_ev)
        return self.time

    def milliseconds(self):
        return self.time

    def seconds(self):
        return self.time * 1e-3


def stream_synchronize(stream):
    _ok(runtime().hipStreamSynchronize(stream), "hipStreamSynchronize")


def memcpy(dst, src, nbytes):
    """Synchronous hipMemcpy between device (or host) addresses."""
    _ok(runtime().hipMemcpy(dst, src, int(nbytes), 4), "hipMemcpy")  # hipMemcpyDefault
