"""Autotuner for batch items: picks the schedule and tile size per
(size, mean row length, row skew, k) bucket from measurements on the GPU.

No counterpart in the reference, whose schedule picks exist only as data
(plots/data/heuristics.csv: per-matrix best of merge-path / thread-mapped /
group-mapped).  Here the candidates are the batch kernel's piece kinds:

  k == 1: merge_path_flat with 4, 8 or 16 steps per lane; thread_mapped;
          group_mapped with G lanes per row (G from the mean row length, and
          4x that); "blocks": a row-block plan cut for batches
          (spmv_plan(..., layout="blocks", for_batches=True)), measured for
          buckets whose matrices hold at least BLOCKS_MIN_NNZ nonzeros on
          average;
  k >  1: merge_path_flat SpMM tiles of 4, 8 or 16 merge steps per range,
          or 0: the library's default (16 steps, 32 for ranges of 16 lanes
          or more).

For each bucket the tuning matrices that fall in it are run as one batch per
candidate (items repeated until the batch holds enough work to fill the
GPU), captured in a graph and replayed; the fastest candidate wins.  The
table is plain JSON: bucket key -> {"schedule", "tile", "us": {candidate:
microseconds}}.  Lookup falls back to the nearest tuned bucket, then to
merge_path_flat with 8 steps per lane (SpMM: the library's default); a bucket's
measured winner is used only when it beats that default by PICK_MARGIN.
A "blocks" pick asks the caller for a plan per matrix (built once; a copy
of A in the row-block layout); callers that take no plans ask for picks
without it (allow_blocks=False) and get the best of the other candidates,
as do matrices of fewer than BLOCKS_MIN_NNZ nonzeros.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

SPMV_MERGE = [4, 8, 16]
SPMM_IPT = [0, 4, 8, 16]
DEFAULT_SPMV = ("merge_path_flat", 8)
DEFAULT_SPMM = ("merge_path_flat", 0)
# A measured pick replaces the default only when it is at least this much
# faster: smaller differences are within run-to-run noise, and in a mixed
# batch every extra tile class costs a persistent kernel of its own.
PICK_MARGIN = 0.05
# In one batch launch every kernel class (SpMV: merge tiles of 4 steps with
# the row / lane-group pieces, 8, 16; SpMM: steps per range) is a persistent
# kernel of its own, run after the others.  A class other than the default's
# holding less than this share of the batch's work runs with the default
# instead: one 16-step matrix among 511 8-step ones cost 13-17 us of a
# 1.28 ms pass.
MIN_CLASS_SHARE = 0.1
TABLE_PATH = os.path.join(os.path.dirname(__file__), "tuning", "mi355x_batch.json")
# Below this many nonzeros per matrix a plan is a round or two per stream
# and is not measured.
BLOCKS_MIN_NNZ = 1 << 14


def features(rows, nnz, max_row, k=1):
    """Bucket key of a product: 'k{kb}/s{size}/d{density}/w{skew}'.
    size = floor(log2(rows + nnz)) // 2 (two-octave bins), density =
    floor(log2(mean row length)) clamped to 0..8, skew = floor(log2(max row /
    mean row)) // 2 clamped to 0..6, kb = 0 for k == 1 else ceil(log2 k)."""
    rows, nnz, max_row, k = int(rows), int(nnz), int(max_row), int(k)
    size = int(math.log2(max(rows + nnz, 1))) // 2
    mean = nnz / rows if rows else 0.0
    dens = min(max(int(math.floor(math.log2(mean))) if mean >= 1 else 0, 0), 8)
    skew_r = max_row / mean if mean > 0 else 1.0
    skew = min(max(int(math.floor(math.log2(max(skew_r, 1.0)))) // 2, 0), 6)
    kb = 0 if k <= 1 else min(int(math.ceil(math.log2(k))), 7)
    return f"k{kb}/s{size}/d{dens}/w{skew}"


def _key_vec(key):
    return np.array([int(p[1:]) for p in key.split("/")])


def group_lanes(rows, nnz):
    mean = (nnz + rows - 1) // rows if rows else 1
    g = 2
    while g < 64 and g < mean:
        g *= 2
    return g


def spmv_candidates():
    """(schedule, tile) names; group_mapped tiles are lane-count factors
    over group_lanes() of the matrix ("x1", "x4")."""
    return ([("merge_path_flat", i) for i in SPMV_MERGE]
            + [("thread_mapped", 0), ("group_mapped", "x1"), ("group_mapped", "x4"),
               ("blocks", 0)])


def kernel_class(sched, ipt, k=1):
    """The batch kernel an item runs in: ("spmv", 4 | 8 | 16), ("blocks",)
    or ("spmm", steps per range)."""
    if k > 1:
        return ("spmm", ipt)
    if sched == "blocks":
        return ("blocks",)
    if sched == "merge_path_flat" and ipt in (8, 16):
        return ("spmv", ipt)
    return ("spmv", 4)


def resolve(sched, tile, rows, nnz):
    """items_per_thread for the batch C ABI from a (schedule, tile) name."""
    if sched == "group_mapped":
        f = int(str(tile).lstrip("x"))
        return min(64, group_lanes(rows, nnz) * f)
    if sched == "blocks":
        return 0
    return int(tile)


def spmm_candidates():
    return [("merge_path_flat", u) for u in SPMM_IPT]


class Table:
    """Tuning table with nearest-bucket lookup."""

    def __init__(self, data=None):
        self.data = data or {"spmv": {}, "spmm": {}}

    @classmethod
    def load(cls, path=TABLE_PATH):
        if not os.path.exists(path):
            return cls()
        with open(path) as f:
            return cls(json.load(f))

    def save(self, path=TABLE_PATH):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)

    def pick(self, rows, nnz, max_row, k=1, allow_blocks=True):
        """(schedule, items_per_thread) for one product."""
        sched, tile = self.pick_name(rows, nnz, max_row, k, allow_blocks)
        return sched, resolve(sched, tile, rows, nnz)

    def pick_batch(self, products, allow_blocks=True):
        """Picks for the products of one batch launch, products = [(rows,
        nnz, max_row, k)]: each product's pick, except where its kernel
        class holds less than MIN_CLASS_SHARE of the batch's work ((rows +
        nnz) * k) and is not the default's class; those take the default."""
        picks = [self.pick(r, n, mr, k, allow_blocks) for r, n, mr, k in products]
        work = [(r + n) * max(k, 1) for r, n, _, k in products]
        share = {}
        for (sched, ipt), w, (_, _, _, k) in zip(picks, work, products):
            c = kernel_class(sched, ipt, k)
            share[c] = share.get(c, 0) + w
        total = sum(work) or 1
        out = []
        for (sched, ipt), (r, n, _, k) in zip(picks, products):
            d = DEFAULT_SPMV if k == 1 else DEFAULT_SPMM
            d = (d[0], resolve(d[0], d[1], r, n))
            c = kernel_class(sched, ipt, k)
            if c != kernel_class(d[0], d[1], k) and share[c] < MIN_CLASS_SHARE * total:
                sched, ipt = d
            out.append((sched, ipt))
        return out

    def pick_name(self, rows, nnz, max_row, k=1, allow_blocks=True):
        key = features(rows, nnz, max_row, k)
        tab = self.data["spmv" if k == 1 else "spmm"]
        if key in tab:
            e = tab[key]
        elif tab:
            kv = _key_vec(key)
            near = min(tab, key=lambda q: (int(np.abs(_key_vec(q) - kv).sum()), q))
            e = tab[near]
        else:
            return DEFAULT_SPMV if k == 1 else DEFAULT_SPMM
        default = DEFAULT_SPMV if k == 1 else DEFAULT_SPMM
        us = e.get("us", {})
        sched, tile = e["schedule"], e["tile"]
        if sched == "blocks" and (not allow_blocks or nnz < BLOCKS_MIN_NNZ):
            rest = {c: v for c, v in us.items() if not c.startswith("blocks:")}
            if not rest:
                return default
            sched, tile = min(rest, key=rest.get).split(":")
        # JSON keeps tiles as strings ("8", "x4"); numeric ones become int.
        t = int(tile) if str(tile).isdigit() else tile
        if k > 1 and t not in SPMM_IPT:
            return default  # an entry from another tile design
        best = us.get(f"{sched}:{tile}")
        base = us.get(f"{default[0]}:{default[1]}")
        if best is not None and base is not None and best > (1.0 - PICK_MARGIN) * base:
            return default
        return sched, t


def tuning_set(count=192, seed=77):
    """Host matrices covering the buckets the batched and served workloads
    hit: rows log-uniform in [1e3, 4e6], mean row length log-uniform in
    [1, 48] (half unscaled Zipf), Zipf exponent uniform in [1.5, 2.5],
    at most 8M nonzeros each."""
    from . import synth
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        rows = int(math.exp(rng.uniform(math.log(1e3), math.log(4e6))))
        s = float(rng.uniform(1.5, 2.5))
        cap = min(4096, rows)
        if i % 2 == 0:
            target = 0
        else:
            mean = math.exp(rng.uniform(0, math.log(48)))
            target = int(min(max(rows * mean, rows), 8e6))
        m = synth.powerlaw_csr(rows, rows, s, 1, cap, target, seed=seed * 1000 + i)
        if m.nnz > 8e6:
            continue
        out.append(m)
    return out


def _time_batch(items, reps=5):
    """Microseconds per replay of a graph holding one batch launch."""
    import torch
    from .batch import Batch
    b = Batch(items)
    g = b.capture()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1000.0 / reps


def tune(mats, ks=(2, 4, 8, 16, 32), min_work=30_000_000, log=print):
    """Measure every candidate per bucket on the GPU; returns a Table."""
    import torch
    from . import containers
    from .algorithms import spmv_plan
    from .batch import BatchItem
    dev = []
    for m in mats:
        A = containers.csr_t.from_host(m.rows, m.cols, m.offsets, m.indices, m.values)
        dev.append((m, A))
    table = Table()
    # SpMV buckets
    buckets = {}
    for m, A in dev:
        buckets.setdefault(features(m.rows, m.nnz, m.max_row(), 1), []).append((m, A))
    xs = {}
    for key in sorted(buckets):
        group = buckets[key]
        work = sum(m.rows + m.nnz for m, _ in group)
        rep = max(1, math.ceil(min_work / max(work, 1)))
        rep = min(rep, 64)
        times = {}
        for sched, tile in spmv_candidates():
            if sched == "blocks" and work - sum(m.rows for m, _ in group) < \
                    BLOCKS_MIN_NNZ * len(group):
                continue
            items = []
            for r in range(rep):
                for m, A in group:
                    if id(A) not in xs:
                        xs[id(A)] = torch.rand(A.cols, device="cuda") * 2 - 1
                    y = torch.empty(A.rows, device="cuda")
                    if sched == "blocks":  # a plan per item: items of one plan would race
                        plan = spmv_plan(A, layout="blocks", for_batches=True)
                        items.append(BatchItem(A, xs[id(A)], y, "blocks", plan=plan))
                    else:
                        items.append(BatchItem(A, xs[id(A)], y, sched,
                                               resolve(sched, tile, m.rows, m.nnz)))
            times[f"{sched}:{tile}"] = _time_batch(items)
            del items
        best = min(times, key=times.get)
        sched, tile = best.split(":")
        table.data["spmv"][key] = {"schedule": sched, "tile": tile, "us": times,
                                   "matrices": len(group), "repeat": rep}
        log(f"spmv {key}: n={len(group)} rep={rep} best={best} "
            + " ".join(f"{c}={t:.1f}" for c, t in sorted(times.items())))
    # SpMM buckets (a subset of matrices per bucket keeps the run short)
    for k in ks:
        buckets = {}
        for m, A in dev:
            buckets.setdefault(features(m.rows, m.nnz, m.max_row(), k), []).append((m, A))
        for key in sorted(buckets):
            group = buckets[key][:6]
            work = sum((m.rows + m.nnz) * k for m, _ in group)
            rep = min(max(1, math.ceil(min_work * 4 / max(work, 1))), 32)
            times = {}
            X = {}
            for sched, ipt in spmm_candidates():
                items = []
                for r in range(rep):
                    for m, A in group:
                        if id(A) not in X:
                            X[id(A)] = torch.rand(A.cols, k, device="cuda") * 2 - 1
                        Y = torch.empty(A.rows, k, device="cuda")
                        items.append(BatchItem(A, X[id(A)], Y, sched, ipt))
                times[f"{sched}:{ipt}"] = _time_batch(items)
            best = min(times, key=times.get)
            sched, tile = best.split(":")
            table.data["spmm"][key] = {"schedule": sched, "tile": tile, "us": times,
                                       "matrices": len(group), "repeat": rep}
            log(f"spmm {key}: n={len(group)} rep={rep} best={best} "
                + " ".join(f"{c}={t:.1f}" for c, t in sorted(times.items())))
            del X
            torch.cuda.empty_cache()
    return table


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Tune batch schedules on this GPU")
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--count", type=int, default=192)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--spmv-only", action="store_true",
                    help="tune the SpMV buckets only; keep the SpMM entries of --out")
    a = ap.parse_args(argv)
    mats = tuning_set(24 if a.quick else a.count)
    print(f"tuning set: {len(mats)} matrices, {sum(m.nnz for m in mats)} nnz", flush=True)
    ks = () if a.spmv_only else (4, 32) if a.quick else (2, 4, 8, 16, 32)
    t = tune(mats, ks=ks, log=lambda s: print(s, flush=True))
    if a.spmv_only:
        t.data["spmm"] = Table.load(a.out).data.get("spmm", {})
    t.data["about"] = {
        "device": "MI355X (gfx950)",
        "features": "k{ceil log2 k}/s{floor log2(rows+nnz) // 2}/d{floor log2 mean row}"
                    "/w{floor log2(max row / mean row) // 2}",
        "tuning_set": f"{len(mats)} matrices from edge_loops_amd.tune.tuning_set",
        "command": "python -m edge_loops_amd.tune --out <path>",
    }
    t.save(a.out)
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
