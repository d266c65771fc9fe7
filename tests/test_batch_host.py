"""Host-side checks of the batch layout (loops_batch_query needs no GPU) and
of the autotuner's bucketing and table lookup."""
import ctypes

import pytest

from edge_loops_amd import _capi, tune

BASE = 1 << 20  # fake, 256-byte aligned device addresses (never dereferenced)


def item(rows, nnz, k=1, schedule="merge_path_flat", ipt=0, vt=_capi.VALUE_F32,
         idx_off=0, ldx=None, ldy=None):
    v = _capi.CsrView(rows, rows, nnz, BASE, BASE + 4096 + idx_off, BASE + 8192, vt)
    return _capi.BatchItem(v, k, BASE + 16384, ldx if ldx is not None else (k if k > 1 else 1),
                           BASE + 32768, ldy if ldy is not None else (k if k > 1 else 1),
                           _capi.SCHEDULES[schedule], ipt)


def query(items):
    arr = (_capi.BatchItem * len(items))(*items)
    info = _capi.BatchInfo()
    st = _capi.lib().loops_batch_query(arr, len(items), ctypes.byref(info))
    return st, info


def test_layout_counts():
    st, info = query([item(1000, 9000, ipt=8)])
    assert st == _capi.SUCCESS
    assert (info.spmv_tiles, info.coordinates, info.work_items) == (5, 6, 5)
    assert info.max_items_per_thread == 8 and info.spmm_work_items == 0
    st, info = query([item(1000, 9000, schedule="thread_mapped")])
    assert st == _capi.SUCCESS and info.work_items == 4 and info.spmv_tiles == 0
    st, info = query([item(1000, 9000, schedule="group_mapped", ipt=16)])
    assert st == _capi.SUCCESS and info.work_items == 63
    # k = 8 with 16-byte B rows: 4 columns per lane, ranges of 2 lanes,
    # 128 ranges x 16 steps (the default) = 2048 steps per tile
    st, info = query([item(1000, 9000, k=8)])
    assert st == _capi.SUCCESS
    assert (info.spmm_units, info.spmm_work_items, info.work_items) == (5, 5, 5)
    assert info.coordinates == 6
    # k = 7: 1 column per lane, ranges of 8 lanes, 32 x 16 = 512 steps per tile
    st, info = query([item(1000, 9000, k=7, ldx=8)])
    assert st == _capi.SUCCESS and info.spmm_units == 20
    # k = 32 with 16-byte rows: 4 columns per lane, 8 lanes -> 512-step tiles at ipt 16
    st, info = query([item(1000, 9000, k=32, ipt=16)])
    assert st == _capi.SUCCESS and info.spmm_units == 20
    # k = 32 at 32 steps per range (ranges of 8 lanes): 1024-step tiles
    st, info = query([item(1000, 9000, k=32, ipt=32)])
    assert st == _capi.SUCCESS and info.spmm_units == 10
    # k = 72: chunks of 64 and 8 columns (as loops_spmm launches them): 20
    # units of 512 steps (16 lanes x 4 columns, 32 steps) and 5 of 2048
    # steps for the 8 columns (2 lanes x 4 columns, 16 steps)
    st, info = query([item(1000, 9000, k=72, ldx=72, ldy=72)])
    assert st == _capi.SUCCESS
    assert (info.spmm_units, info.coordinates) == (25, 27)
    # k = 33: one chunk, ranges of 16 lanes x 4 columns, 32 steps
    st, info = query([item(1000, 9000, k=33, ldx=36, ldy=36)])
    assert st == _capi.SUCCESS and info.spmm_units == 20 and info.spmv_tiles == 0
    # k = 65: 60 + 5 columns, never a one-column chunk (that would be an
    # SpMV item): 20 units of 512 steps, then 5 columns (8 lanes x 1 column,
    # 16 steps: 512 steps per tile)
    st, info = query([item(1000, 9000, k=65, ldx=68, ldy=68)])
    assert st == _capi.SUCCESS and info.spmm_units == 40 and info.spmv_tiles == 0
    # a wide item is validated whole: ldx < k is rejected even though each
    # chunk would fit
    assert query([item(1000, 9000, k=72, ldx=40)])[0] == _capi.ERROR_INVALID_VALUE
    st, info = query([item(1000, 9000, ipt=4), item(1000, 9000, k=8), item(50, 0, ipt=16)])
    assert st == _capi.SUCCESS
    assert info.spmv_tiles == 10 + 1 and info.coordinates == 11 + 6 + 2
    assert info.work_items == 16 and info.max_items_per_thread == 16
    assert info.table_bytes > 0 and info.workspace_bytes > 0


def test_layout_rejections():
    NS, IV = _capi.ERROR_NOT_SUPPORTED, _capi.ERROR_INVALID_VALUE
    assert query([item(1000, 9000, idx_off=4)])[0] == NS          # unaligned indices
    assert query([item(1000, 9000, schedule="group_mapped", ipt=3)])[0] == NS
    assert query([item(1000, 9000, schedule="work_oriented")])[0] == NS
    assert query([item(1000, 9000, ipt=5)])[0] == NS
    assert query([item(1000, 9000, k=4, ipt=32)])[0] == NS         # 32 steps: 8+ lanes
    assert query([item(1000, 9000, k=16, ipt=32)])[0] == NS        # 4 lanes x 4 columns
    assert query([item(1000, 9000, k=16, ipt=64)])[0] == NS
    assert query([item(1000, 9000, k=0)])[0] == IV
    assert query([item(1000, 9000, k=8, ldx=4)])[0] == IV          # ldx < k
    assert query([item(-1, 0)])[0] == IV
    assert query([item(1000, 9000, vt=_capi.VALUE_F16)])[0] == _capi.SUCCESS
    assert query([item(2**30, 2**30)])[0] == _capi.ERROR_OVERFLOW


def test_tune_features_and_lookup(tmp_path):
    assert tune.features(1000, 12000, 400, 1) == "k0/s6/d3/w2"
    assert tune.features(1000, 12000, 400, 32) == "k5/s6/d3/w2"
    assert tune.features(0, 0, 0, 1) == "k0/s0/d0/w0"
    t = tune.Table()
    assert t.pick(1000, 12000, 400) == tune.DEFAULT_SPMV
    assert t.pick(1000, 12000, 400, 8) == tune.DEFAULT_SPMM
    t.data["spmv"]["k0/s6/d3/w2"] = {"schedule": "merge_path_flat", "tile": "16"}
    t.data["spmv"]["k0/s9/d0/w0"] = {"schedule": "group_mapped", "tile": "x4"}
    assert t.pick(1000, 12000, 400) == ("merge_path_flat", 16)
    assert t.pick(1100, 13000, 400) == ("merge_path_flat", 16)     # nearest bucket
    assert t.pick(200000, 200000, 1) == ("group_mapped", 8)        # 2 lanes x 4
    p = tmp_path / "t.json"
    t.save(str(p))
    assert tune.Table.load(str(p)).pick(1000, 12000, 400) == ("merge_path_flat", 16)


def test_tune_batch_folds_small_classes():
    # a bucket whose measured winner is 16 steps per lane
    t = tune.Table()
    t.data["spmv"]["k0/s6/d3/w2"] = {"schedule": "merge_path_flat", "tile": "16"}
    t.data["spmv"]["k0/s9/d0/w0"] = {"schedule": "merge_path_flat", "tile": "8"}
    t.data["spmm"]["k1/s6/d3/w2"] = {"schedule": "merge_path_flat", "tile": "4"}
    one = (1000, 12000, 400, 1)        # 16-step bucket, 13,000 work
    other = (200000, 200000, 1, 1)     # 8-step bucket, 400,000 work
    # one 16-step product among 20 8-step ones: 0.16 % of the work
    picks = t.pick_batch([one] + [other] * 20)
    assert picks == [tune.DEFAULT_SPMV] * 21
    # as much work in the 16-step class as in the default's: kept
    picks = t.pick_batch([one] * 31 + [other])
    assert picks[0] == ("merge_path_flat", 16) and picks[-1] == tune.DEFAULT_SPMV
    # SpMM classes are counted with their k: 2 x 13,000 of 426,000 folds,
    # 8 x 13,000 of 504,000 does not
    assert t.pick(1000, 12000, 400, 2) == ("merge_path_flat", 4)
    assert t.pick_batch([(1000, 12000, 400, 2), other])[0] == tune.DEFAULT_SPMM
    assert t.pick_batch([(1000, 12000, 400, 8), other])[0] == ("merge_path_flat", 4)
    # a small class that is the default's own class stays as picked
    assert t.pick_batch([other, one])[0] == tune.DEFAULT_SPMV
    assert tune.kernel_class("group_mapped", 8) == ("spmv", 4)
    assert tune.kernel_class("merge_path_flat", 16, 32) == ("spmm", 16)


def test_committed_table_is_complete():
    t = tune.Table.load()
    assert t.data["spmv"] and t.data["spmm"]
    for sec in ("spmv", "spmm"):
        for key, e in t.data[sec].items():
            assert e["schedule"] in _capi.SCHEDULES or (sec == "spmv" and e["schedule"] == "blocks")
            sched, ipt = t.pick(1000, 5000, 50, 1 if sec == "spmv" else 8)
            assert sched in _capi.SCHEDULES and ipt >= 0


def test_table_size_grows_with_items_not_with_their_sizes():
    # One descriptor and a few prefix counts per item: a batch of large
    # matrices has the same table as a batch of small ones.
    small = query([item(1000, 9000, k=16, ipt=8) for _ in range(5)])[1].table_bytes
    big = query([item(4_000_000, 80_000_000, k=16, ipt=8) for _ in range(5)])[1].table_bytes
    assert small == big and big < 8192
    n5 = query([item(1000, 9000) for _ in range(5)])[1].table_bytes
    n50 = query([item(1000, 9000) for _ in range(50)])[1].table_bytes
    assert n50 > n5 and n50 < 50 * 1024


def test_spmm_variant_mask():
    # Variant bit = (value type * 2 + (4 columns per lane)) * 4 + (steps 4/8/16/32 -> 0..3).
    # With the fake addresses B rows are 16-byte aligned, so k >= 8 with ldx a
    # multiple of 4 takes 4 columns per lane.
    def mask(items):
        st, info = query(items)
        assert st == _capi.SUCCESS
        return info.spmm_variant_mask
    assert mask([item(1000, 9000)]) == 0                                   # SpMV only
    assert mask([item(1000, 9000, k=7, ipt=8)]) == 1 << 1                  # f32, 1 col, 8
    assert mask([item(1000, 9000, k=8, ipt=8)]) == 1 << 5                  # f32, 4 cols, 8
    assert mask([item(1000, 9000, k=12, ipt=8, ldx=14)]) == 1 << 1         # ldx % 4: 1 col
    assert mask([item(1000, 9000, k=16, ipt=4)]) == 1 << 4                 # f32, 4 cols, 4
    assert mask([item(1000, 9000, k=16, ipt=16, vt=_capi.VALUE_BF16)]) == 1 << 14
    assert mask([item(1000, 9000, k=5, ipt=16, vt=_capi.VALUE_F16)]) == 1 << 18
    assert mask([item(1000, 9000, k=32, ipt=32)]) == 1 << 7                # f32, 4 cols, 32
    assert mask([item(1000, 9000, k=64)]) == 1 << 7                        # default: 32
    assert mask([item(1000, 9000, k=16)]) == 1 << 6                        # default: 16
    assert mask([item(1000, 9000, k=7, ipt=8), item(1000, 9000, k=7, ipt=8),
                 item(0, 0, k=7, ipt=4)]) == 1 << 1                        # no pieces: no bit
    assert mask([item(10, 0, k=7, ipt=4)]) == 1 << 0                       # empty rows are work


def test_pick_margin_keeps_the_default_for_small_gains():
    t = tune.Table()
    key = tune.features(1000, 12000, 400, 1)
    us = {"merge_path_flat:4": 101.0, "merge_path_flat:8": 100.0, "merge_path_flat:16": 97.0}
    t.data["spmv"][key] = {"schedule": "merge_path_flat", "tile": "16", "us": us}
    assert t.pick(1000, 12000, 400) == tune.DEFAULT_SPMV          # 3 % < margin
    us["merge_path_flat:16"] = 90.0
    assert t.pick(1000, 12000, 400) == ("merge_path_flat", 16)    # 10 % > margin
    kkey = tune.features(1000, 12000, 400, 16)
    t.data["spmm"][kkey] = {"schedule": "merge_path_flat", "tile": "4",
                            "us": {"merge_path_flat:4": 99.0, "merge_path_flat:0": 100.0}}
    assert t.pick(1000, 12000, 400, 16) == tune.DEFAULT_SPMM


def test_blocks_picks():
    """A bucket won by row-block plans: picked when allowed and the matrix
    holds enough nonzeros; otherwise the best other candidate (subject to
    the margin); folded into the default like any small kernel class."""
    t = tune.Table()
    key = tune.features(100000, 1400000, 4096, 1)
    us = {"merge_path_flat:4": 120.0, "merge_path_flat:8": 100.0, "merge_path_flat:16": 93.0,
          "blocks:0": 70.0}
    t.data["spmv"][key] = {"schedule": "blocks", "tile": "0", "us": us}
    assert t.pick(100000, 1400000, 4096) == ("blocks", 0)
    assert t.pick(100000, 1400000, 4096, allow_blocks=False) == ("merge_path_flat", 16)
    us["merge_path_flat:16"] = 97.0  # the best other one within the margin: the default
    assert t.pick(100000, 1400000, 4096, allow_blocks=False) == tune.DEFAULT_SPMV
    # nearest bucket for a matrix too small for a plan
    assert t.pick(2000, tune.BLOCKS_MIN_NNZ - 1, 100)[0] != "blocks"
    assert tune.kernel_class("blocks", 0) == ("blocks",)
    other = (200000, 200000, 1, 1)
    t.data["spmv"][tune.features(*other)] = {"schedule": "merge_path_flat", "tile": "8"}
    big = (100000, 1400000, 4096, 1)
    assert t.pick_batch([big, other])[0] == ("blocks", 0)
    assert t.pick_batch([big] + [other] * 40)[0] == tune.DEFAULT_SPMV  # 3.5 % of the work
    assert tune.resolve("blocks", "0", 10, 10) == 0
