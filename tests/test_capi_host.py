"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/loops_capi.h declares, the host-only entry points validate
their arguments, and the synthetic generator keeps its contract.  No kernel
is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import edge_loops_amd as loops
from edge_loops_amd import _capi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "loops_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(loops_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = _declared_symbols()
    assert len(names) >= 12
    L = _capi.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_capi.SIGNATURES)


def test_status_and_schedule_strings():
    L = _capi.lib()
    assert L.loops_status_string(0) == b"success"
    assert b"overflow" in L.loops_status_string(2)
    assert L.loops_schedule_name(0) == b"merge_path_flat"
    assert L.loops_schedule_name(4) == b"bucketing"
    assert L.loops_schedule_name(5) == b"original"


def test_num_tiles_and_validation():
    L = _capi.lib()
    T = ctypes.c_int64()
    assert L.loops_merge_path_num_tiles(4096, 32768, 640, ctypes.byref(T)) == 0
    assert T.value == 58  # merge_path_flat.cuh:116-117 for config 1
    assert L.loops_merge_path_num_tiles(2_000_000, 40_000_002, 2048, ctypes.byref(T)) == 0
    assert T.value == -(-42_000_002 // 2048)
    assert L.loops_merge_path_num_tiles(-1, 0, 640, ctypes.byref(T)) == _capi.ERROR_INVALID_VALUE
    assert L.loops_merge_path_num_tiles(1, 0, 0, ctypes.byref(T)) == _capi.ERROR_INVALID_VALUE
    assert L.loops_merge_path_num_tiles(2**30, 2**30, 640, ctypes.byref(T)) == _capi.ERROR_OVERFLOW
    # the limit, 2^31 - 1 - 2^20 rows + nonzeros, is accepted; one more is not
    limit = 2**31 - 1 - 2**20
    assert L.loops_merge_path_num_tiles(4, limit - 4, 2048, ctypes.byref(T)) == 0
    assert T.value == -(-limit // 2048)
    assert L.loops_merge_path_num_tiles(4, limit - 3, 2048, ctypes.byref(T)) == _capi.ERROR_OVERFLOW
    # (non-null stand-in pointers: the size check comes before any launch)
    assert L.loops_work_oriented_partition(1, 4, limit - 3, 64, 1, None) == _capi.ERROR_OVERFLOW
    # partition / workspace queries reject bad views before touching a GPU
    assert L.loops_merge_path_partition(None, 10, 0, 640, None, None) == _capi.ERROR_INVALID_VALUE
    v = _capi.CsrView(-1, 4, 0, None, None, None, 0)
    n = ctypes.c_size_t()
    assert L.loops_spmv_workspace_size(None, 0, ctypes.byref(v), None, ctypes.byref(n)) == _capi.ERROR_INVALID_VALUE
    v = _capi.CsrView(2**30, 4, 2**30, 1, 1, 1, 0)
    assert L.loops_spmv_workspace_size(None, 0, ctypes.byref(v), None, ctypes.byref(n)) == _capi.ERROR_OVERFLOW
    v = _capi.CsrView(4, 4, limit - 3, 1, 1, 1, 0)
    assert L.loops_spmv_workspace_size(None, 0, ctypes.byref(v), None, ctypes.byref(n)) == _capi.ERROR_OVERFLOW
    # plans are refused above 2^26 rows + nonzeros, before any device work
    plan = ctypes.c_void_p()
    v = _capi.CsrView(4, 4, 2**26 - 3, 1, 1, 1, 0)
    assert L.loops_spmv_plan_create(None, ctypes.byref(v), None, None,
                                    ctypes.byref(plan)) == _capi.ERROR_NOT_SUPPORTED
    v = _capi.CsrView(4, 4, 4, 1, 1, 1, 0)
    assert L.loops_spmv_workspace_size(None, 4, ctypes.byref(v), None, ctypes.byref(n)) == _capi.ERROR_NOT_SUPPORTED
    p = _capi.Params(128, 7, 0, 0, None, None, None)  # no (128, 7) tuning
    assert L.loops_spmv_workspace_size(None, 0, ctypes.byref(v), ctypes.byref(p), ctypes.byref(n)) == _capi.ERROR_NOT_SUPPORTED
    p = _capi.Params(128, 5, 0, 0, None, None, None)
    assert L.loops_spmv_workspace_size(None, 0, ctypes.byref(v), ctypes.byref(p), ctypes.byref(n)) == 0
    assert n.value >= 12  # one merge tile: carry row + two partial sums
    p = _capi.Params(0, 0, 0, 0, None, None, None)
    assert L.loops_spmv_workspace_size(None, 2, ctypes.byref(v), ctypes.byref(p), ctypes.byref(n)) == 0
    assert n.value == 0  # thread_mapped needs none
    # null vectors with work to do are rejected before any launch
    assert L.loops_spmv(None, 2, ctypes.byref(v), None, None, None, None, 0, None) == _capi.ERROR_INVALID_VALUE


def test_synth_contract():
    A = synth.powerlaw_csr(5000, 3000, 1.8, 1, 2000, 40000, seed=5, force_first_max=True)
    L = np.diff(A.offsets)
    assert L[0] == 2000 and L.min() >= 1 and L.max() <= 2000
    assert abs(A.nnz - 40000) < 40000 * 0.01
    for r in range(0, 5000, 37):
        seg = A.indices[A.offsets[r]:A.offsets[r + 1]]
        assert np.all(np.diff(seg) > 0) and seg.min() >= 0 and seg.max() < 3000
    assert A.values.min() >= -1 and A.values.max() < 1
    B = synth.powerlaw_csr(5000, 3000, 1.8, 1, 2000, 40000, seed=5, force_first_max=True, nthreads=1)
    assert np.array_equal(A.indices, B.indices) and np.array_equal(A.values, B.values)
    # dense rows (selection sampling path)
    D = synth.powerlaw_csr(10, 16, 0.0, 16, 16, 0, seed=1)
    assert np.array_equal(D.indices.reshape(10, 16), np.tile(np.arange(16), (10, 1)))


def test_python_mirror_surface():
    for name in ("original", "thread_mapped", "group_mapped", "work_oriented",
                 "merge_path_flat", "merge_path"):
        assert callable(getattr(loops.spmv, name))
    for name in ("thread_mapped", "group_mapped", "merge_path"):
        assert callable(getattr(loops.spmm, name))


def test_plan_parameter_validation_without_a_device():
    # Invalid plan parameters are rejected before anything touches a device.
    import ctypes
    base = 1 << 20
    v = _capi.CsrView(1000, 1000, 9000, base, base + 4096, base + 8192, _capi.VALUE_F32)
    h = ctypes.c_void_p()
    L = _capi.lib()
    for bad in (_capi.PlanParams(0, 0, 0, 3), _capi.PlanParams(0, 0, 0, -1),
                _capi.PlanParams(0, 0, 9, 1), _capi.PlanParams(-1, 0, 0, 1),
                # a blocks cut for batches with PANELS, or with a PANELS parameter
                _capi.PlanParams(0, 0, 0, 1, 1), _capi.PlanParams(0, 0, 2, 0, 1),
                _capi.PlanParams(256, 0, 0, 2, 1)):
        st = L.loops_spmv_plan_create(None, ctypes.byref(v), ctypes.byref(bad), None,
                                      ctypes.byref(h))
        assert st == _capi.ERROR_INVALID_VALUE
    assert L.loops_spmv_plan_layout(None, ctypes.byref(ctypes.c_int32())) == \
        _capi.ERROR_INVALID_VALUE
    assert _capi.PLAN_LAYOUTS == {"auto": 0, "panels": 1, "blocks": 2}


STRUCTS = {  # ctypes mirror -> C typedef of include/loops_capi.h
    "CsrView": "loops_csr_view_t", "Coord": "loops_coord_t", "Params": "loops_params_t",
    "PlanParams": "loops_plan_params_t", "BatchItem": "loops_batch_item_t",
    "BatchInfo": "loops_batch_info_t", "ServerConfig": "loops_server_config_t",
    "ServerMatrix": "loops_server_matrix_t", "ServerStats": "loops_server_stats_t",
    "ServerRequest": "loops_server_request_t",
}


def test_ctypes_mirrors_match_the_header(tmp_path):
    """Size and every field offset of each ctypes structure equal the C
    header's, as gcc lays them out (a field added on one side only shifts
    what the library reads)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("gcc or the HIP headers are absent")
    lines = ["#include <stdio.h>", "#include <stddef.h>", "#include <loops_capi.h>",
             "int main(void) {"]
    expect = []
    for py, c in STRUCTS.items():
        cls = getattr(_capi, py)
        lines.append(f'  printf("%zu\\n", sizeof({c}));')
        expect.append((f"sizeof {c}", ctypes.sizeof(cls)))
        for f in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({c}, {f[0]}));')
            expect.append((f"{c}.{f[0]}", getattr(cls, f[0]).offset))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert len(got) == len(expect)
    bad = [(name, want, g) for (name, want), g in zip(expect, got) if want != g]
    assert not bad, bad
