"""ctypes binding of the C ABI in include/loops_capi.h (libloops_mi355x.so).

The library is loaded from ``edge_loops_amd/lib``; if it is missing the
import fails loudly -- there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libloops_mi355x.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libloops_synth.so")

# loops_status_t
SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_OVERFLOW = 2
ERROR_WORKSPACE = 3
ERROR_NOT_SUPPORTED = 4
ERROR_HIP = 5
ERROR_ALLOCATION = 6

# loops_schedule_t (schedule.hxx:23-29 numbering, ORIGINAL appended)
MERGE_PATH_FLAT = 0
WORK_ORIENTED = 1
THREAD_MAPPED = 2
GROUP_MAPPED = 3
BUCKETING = 4
ORIGINAL = 5

SCHEDULES = {
    "merge_path_flat": MERGE_PATH_FLAT,
    "merge_path": MERGE_PATH_FLAT,
    "work_oriented": WORK_ORIENTED,
    "thread_mapped": THREAD_MAPPED,
    "group_mapped": GROUP_MAPPED,
    "bucketing": BUCKETING,
    "original": ORIGINAL,
}

# loops_value_type_t
VALUE_F32 = 0
VALUE_BF16 = 1
VALUE_F16 = 2
VALUE_F64 = 3
VALUE_I32 = 4


class CsrView(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("nnzs", ctypes.c_int64),
        ("offsets", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
        ("value_type", ctypes.c_int),
    ]


class Coord(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32)]


class Params(ctypes.Structure):
    _fields_ = [
        ("threads_per_block", ctypes.c_int32),
        ("items_per_thread", ctypes.c_int32),
        ("threads_per_tile", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
        ("tile_coordinates", ctypes.c_void_p),
        ("start_event", ctypes.c_void_p),
        ("stop_event", ctypes.c_void_p),
        ("flags", ctypes.c_uint32),
    ]




class PlanParams(ctypes.Structure):
    _fields_ = [
        ("threads_per_block", ctypes.c_int32),
        ("items_per_thread", ctypes.c_int32),
        ("panels", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("for_batches", ctypes.c_int32),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("A", CsrView),
        ("k", ctypes.c_int64),
        ("x", ctypes.c_void_p),
        ("ldx", ctypes.c_int64),
        ("y", ctypes.c_void_p),
        ("ldy", ctypes.c_int64),
        ("schedule", ctypes.c_int),
        ("items_per_thread", ctypes.c_int32),
        ("plan", ctypes.c_void_p),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("table_bytes", ctypes.c_int64),
        ("workspace_bytes", ctypes.c_int64),
        ("work_items", ctypes.c_int64),
        ("spmv_tiles", ctypes.c_int64),
        ("spmm_units", ctypes.c_int64),
        ("spmm_work_items", ctypes.c_int64),
        ("coordinates", ctypes.c_int64),
        ("max_items_per_thread", ctypes.c_int64),
        ("spmm_variant_mask", ctypes.c_int64),
        ("spmv_class_mask", ctypes.c_int64),
        ("blocks_work_items", ctypes.c_int64),
        ("blocks_row_pieces", ctypes.c_int64),
    ]


# loops_plan_layout_t
PLAN_LAYOUTS = {"auto": 0, "panels": 1, "blocks": 2}

# loops_server_mode_t
SERVER_BATCHED = 0
SERVER_PLAIN = 1
SERVER_FLAG_REFERENCE_SCHEDULES = 1
SERVER_FLAG_BATCH_ALL = 2


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("arena_bytes", ctypes.c_int64),
        ("scratch_bytes", ctypes.c_int64),
        ("slots", ctypes.c_int32),
        ("max_batch_requests", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
    ]


class ServerMatrix(ctypes.Structure):
    _fields_ = [
        ("A", CsrView),
        ("spmv_schedule", ctypes.c_int32),
        ("spmv_items_per_thread", ctypes.c_int32),
        ("spmm_items_per_unit", ctypes.c_int32 * 8),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "submitted", "completed", "batches", "batched_requests", "arena_bytes",
        "arena_high_water")]


class ServerRequest(ctypes.Structure):
    _fields_ = [
        ("matrix", ctypes.c_int32),
        ("k", ctypes.c_int64),
        ("x", ctypes.c_void_p),
        ("ldx", ctypes.c_int64),
    ]


class LoopsError(RuntimeError):
    """Raised for a non-success loops_status_t (the C++ layer's
    error::exception_t, reference include/loops/error.hxx:21-45)."""

    def __init__(self, status, where=""):
        self.status = int(status)
        msg = lib().loops_status_string(self.status).decode()
        if self.status == ERROR_HIP:
            msg += f" (hipError {lib().loops_last_hip_error()})"
        super().__init__(f"{where}: {msg}" if where else msg)


# Every symbol include/loops_capi.h declares, with its ctypes signature.
_vp, _i64, _i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
_st = ctypes.c_int
SIGNATURES = {
    "loops_create": (_st, [ctypes.POINTER(ctypes.c_void_p)]),
    "loops_destroy": (_st, [_vp]),
    "loops_status_string": (ctypes.c_char_p, [_st]),
    "loops_schedule_name": (ctypes.c_char_p, [_st]),
    "loops_last_hip_error": (ctypes.c_int, []),
    "loops_device_cu_count": (_st, [_vp, ctypes.POINTER(_i32)]),
    "loops_spmv_workspace_size": (_st, [_vp, _st, ctypes.POINTER(CsrView),
                                        ctypes.POINTER(Params),
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "loops_spmv": (_st, [_vp, _st, ctypes.POINTER(CsrView), _vp, _vp,
                         ctypes.POINTER(Params), _vp, ctypes.c_size_t, _vp]),
    "loops_spmv_typed_workspace_size": (_st, [_vp, _st, ctypes.POINTER(CsrView),
                                              ctypes.POINTER(Params),
                                              ctypes.POINTER(ctypes.c_size_t)]),
    "loops_spmv_typed": (_st, [_vp, _st, ctypes.POINTER(CsrView), _vp, _vp,
                               ctypes.POINTER(Params), _vp, ctypes.c_size_t, _vp]),
    "loops_csr_check": (_st, [ctypes.POINTER(CsrView), ctypes.POINTER(ctypes.c_uint32), _vp]),
    "loops_spmv_noalloc": (_st, [_vp, _st, ctypes.POINTER(CsrView), _vp, _vp,
                                 ctypes.POINTER(Params), _vp]),
    "loops_merge_path_num_tiles": (_st, [_i64, _i64, _i64, ctypes.POINTER(_i64)]),
    "loops_merge_path_partition": (_st, [_vp, _i64, _i64, _i64, _vp, _vp]),
    "loops_work_oriented_partition": (_st, [_vp, _i64, _i64, _i64, _vp, _vp]),
    "loops_spmv_plan_create": (_st, [_vp, ctypes.POINTER(CsrView), ctypes.POINTER(PlanParams),
                                     _vp, ctypes.POINTER(ctypes.c_void_p)]),
    "loops_spmv_plan_execute": (_st, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "loops_spmv_plan_info": (_st, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i64)]),
    "loops_spmv_plan_layout": (_st, [_vp, ctypes.POINTER(_i32)]),
    "loops_spmv_plan_blocks": (_st, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                     ctypes.POINTER(_i32)]),
    "loops_spmv_plan_destroy": (_st, [_vp]),
    "loops_batch_query": (_st, [ctypes.POINTER(BatchItem), _i64, ctypes.POINTER(BatchInfo)]),
    "loops_batch_prepare": (_st, [ctypes.POINTER(BatchItem), _i64, _vp, ctypes.c_size_t, _vp,
                                  ctypes.c_size_t, ctypes.POINTER(BatchInfo)]),
    "loops_batch_launch": (_st, [_vp, ctypes.POINTER(BatchInfo), _i32, _vp]),
    "loops_server_create": (_st, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    "loops_server_destroy": (_st, [_vp]),
    "loops_server_add_matrix": (_st, [_vp, ctypes.POINTER(ServerMatrix),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "loops_server_submit": (_st, [_vp, ctypes.c_int32, _i64, _vp, _i64, _vp,
                                  ctypes.POINTER(ctypes.c_int64)]),
    "loops_server_test": (_st, [_vp, _i64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_vp),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "loops_server_wait": (_st, [_vp, _i64, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]),
    "loops_server_release": (_st, [_vp, _i64]),
    "loops_server_now": (_st, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "loops_server_stats": (_st, [_vp, ctypes.POINTER(ServerStats)]),
    "loops_server_replay": (_st, [_vp, ctypes.POINTER(ServerRequest), _i64,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]),
    "loops_spmm_workspace_size": (_st, [_vp, _st, ctypes.POINTER(CsrView), _i64,
                                        ctypes.POINTER(Params),
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "loops_spmm": (_st, [_vp, _st, ctypes.POINTER(CsrView), _i64, _vp, _i64, _vp,
                         _i64, ctypes.POINTER(Params), _vp, ctypes.c_size_t, _vp]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """Load libloops_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with "
                        "`make -C edge_loops_amd/csrc` (hipcc, gfx950)")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(status, where=""):
    if status != SUCCESS:
        raise LoopsError(status, where)


_handles = threading.local()


def handle():
    """The calling thread's loops_handle_t (created on first use; needs a
    GPU)."""
    h = getattr(_handles, "h", None)
    if h is None:
        p = ctypes.c_void_p()
        check(lib().loops_create(ctypes.byref(p)), "loops_create")
        h = p
        _handles.h = h
    return h
