"""Measurements of the batched form against per-matrix plain launches
(used by bench.py; runs on the GPU).

Batched workload: 512 square CSR matrices, rows log-uniform in [1e3, 2e5],
nnz/row Zipf(1.8) capped at 4096, fp32 (edge_loops_amd.synth.c4_batch_512).
A "request" is one SpMV on one matrix.  Legs, each timed with events over
whole passes of all 512 products, inputs resident in HBM:
  batched_tuned   one persistent batch launch, schedule per matrix from the
                  tuning table, replayed from a captured graph;
  batched_merge8  the same with merge_path_flat 256 x 8 for every matrix;
  plain_eager     512 loops_spmv calls (merge_path_flat, caller workspace,
                  precomputed nothing), launched back to back on one stream;
  plain_graph     the same 512 calls captured once in a graph and replayed.
"""
from __future__ import annotations

import time

import numpy as np

from . import synth
from .algorithms import run_spmv, spmv_workspace_size
from .batch import Batch, BatchItem
from .containers import csr_t

HBM_PEAK_GBS = 8000.0


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _time(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps  # ms


def batched(count=512, reps=10, table=None, log=print):
    import torch
    from . import tune
    t0 = time.time()
    mats = synth.c4_batch_512(count)
    log(f"batched: generated {count} matrices in {time.time() - t0:.1f}s")
    dev, xs, ys = [], [], []
    for m in mats:
        A = csr_t.from_host(m.rows, m.cols, m.offsets, m.indices, m.values)
        dev.append(A)
        xs.append(torch.rand(m.cols, device="cuda", generator=None) * 2 - 1)
        ys.append(torch.empty(m.rows, device="cuda"))
    rows = sum(m.rows for m in mats)
    nnz = sum(m.nnz for m in mats)
    byts = sum(synth.spmv_bytes(m.rows, m.cols, m.nnz) for m in mats)
    flops = 2 * nnz
    table = table or tune.Table.load()
    picks = [table.pick(m.rows, m.nnz, m.max_row(), 1) for m in mats]
    hist = {}
    for s, i in picks:
        hist[f"{s}:{i}"] = hist.get(f"{s}:{i}", 0) + 1

    out = {"workload": f"{count} CSR matrices, rows log-uniform [1e3, 2e5], "
                       "nnz/row Zipf(1.8) capped at 4096, fp32",
           "matrices": count, "rows": rows, "nnz": nnz, "algorithmic_bytes": byts,
           "tuned_picks": hist}

    def leg(ms):
        return {"ms_per_pass": ms, "requests_per_s": count / (ms * 1e-3),
                "gflops": flops / (ms * 1e-3) / 1e9,
                "gbs": byts / (ms * 1e-3) / 1e9,
                "roofline_frac": byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    # batched, tuned and fixed
    results = {}
    for name, pk in (("batched_tuned", picks),
                     ("batched_merge8", [("merge_path_flat", 8)] * count)):
        b = Batch([BatchItem(A, x, y, s, i) for A, x, y, (s, i) in zip(dev, xs, ys, pk)])
        g = b.capture()
        ms = _time(g.replay, reps)
        results[name] = [y.clone() for y in ys] if name == "batched_merge8" else None
        out[name] = leg(ms)
        out[name]["work_items"] = int(b.info.work_items)
        log(f"{name}: {ms:.3f} ms/pass")
        del g, b

    # plain per-call
    wss = [torch.empty(max(spmv_workspace_size("merge_path_flat", A, threads_per_block=256,
                                               items_per_thread=8), 256),
                       dtype=torch.uint8, device="cuda") for A in dev]

    def plain():
        for A, x, y, w in zip(dev, xs, ys, wss):
            run_spmv("merge_path_flat", A, x, y, workspace=w, threads_per_block=256,
                     items_per_thread=8)

    ms = _time(plain, max(2, reps // 3))
    out["plain_eager"] = leg(ms)
    log(f"plain_eager: {ms:.3f} ms/pass")
    same = all(torch.equal(a, b) for a, b in zip(results["batched_merge8"], ys))
    out["batched_merge8_bitwise_equal_plain"] = bool(same)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plain()
    ms = _time(g.replay, reps)
    out["plain_graph"] = leg(ms)
    log(f"plain_graph: {ms:.3f} ms/pass")
    del g
    for k in ("plain_eager", "plain_graph"):
        out[f"speedup_tuned_vs_{k}"] = out[k]["ms_per_pass"] / out["batched_tuned"]["ms_per_pass"]
    return out
