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
  plain_blocking  the same calls, each waited for before the next (the
                  reference's host entries synchronise before returning);
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

    def plain_blocking():
        for A, x, y, w in zip(dev, xs, ys, wss):
            run_spmv("merge_path_flat", A, x, y, workspace=w, threads_per_block=256,
                     items_per_thread=8)
            torch.cuda.current_stream().synchronize()

    out["plain_blocking"] = leg(_time(plain_blocking, 2))
    log(f"plain_blocking: {out['plain_blocking']['ms_per_pass']:.3f} ms/pass")
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
    for k in ("plain_eager", "plain_graph", "plain_blocking"):
        out[f"speedup_tuned_vs_{k}"] = out[k]["ms_per_pass"] / out["batched_tuned"]["ms_per_pass"]
    return out


def _pcts(lat):
    lat = np.asarray(lat) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


def served(n_requests=3000, seed=55, slots=3, log=print, table=None):
    """Served workload: 64 HBM-resident matrices (synth.c5_pool_64), requests
    with a uniform random matrix and k uniform in [1, 32] fp32 right-hand
    sides (k == 1: a vector), seeded Poisson arrivals.  Batched server vs the
    plain per-call mode on the same loop and clock."""
    import torch
    from . import tune
    from .serve import Server, matrix_tunings, poisson_arrivals
    t0 = time.time()
    mats = synth.c5_pool_64()
    log(f"served: generated 64 matrices in {time.time() - t0:.1f}s")
    table = table or tune.Table.load()
    pools = {}
    for vt in ("f32", "bf16"):
        dev = [csr_t.from_host(m.rows, m.cols, m.offsets, m.indices, m.values, value_type=vt)
               for m in mats]
        tun = [matrix_tunings(A, table, m.max_row()) for A, m in zip(dev, mats)]
        pools[vt] = (dev, tun)
    x1 = [torch.rand(m.cols, device="cuda") * 2 - 1 for m in mats]
    X32 = [torch.rand(m.cols, 32, device="cuda") * 2 - 1 for m in mats]
    rng = np.random.default_rng(seed)
    mid = rng.integers(0, len(mats), n_requests)
    ks = rng.integers(1, 33, n_requests)
    flops = float(sum(2 * mats[m].nnz * k for m, k in zip(mid, ks)))
    out = {"workload": "64 CSR matrices, rows log-uniform [1e4, 4e6], nnz log-uniform "
                       "[1e5, 8e6], Zipf 1.8-2.2; requests: uniform matrix, k uniform "
                       "[1, 32] fp32 RHS, seeded Poisson arrivals",
           "matrices": len(mats), "pool_nnz": int(sum(m.nnz for m in mats)),
           "pool_rows": int(sum(m.rows for m in mats)), "requests": n_requests,
           "mean_k": float(ks.mean()), "slots": slots}

    def server(mode, vt, plain_slots=64):
        srv = Server(mode=mode, arena_bytes=32 << 30, scratch_bytes=1 << 30,
                     slots=slots if mode == "batched" else plain_slots)
        dev, tun = pools[vt]
        ids = [srv.add_matrix(A, tunings=t) for A, t in zip(dev, tun)]
        reqs = [(ids[m], x1[m] if k == 1 else X32[m][:, :k]) for m, k in zip(mid, ks)]
        srv.replay(reqs[:200], np.zeros(200))  # warm-up
        return srv, reqs

    def run(srv, reqs, rate):
        arr = np.zeros(len(reqs)) if rate is None else poisson_arrivals(len(reqs), rate, seed + 1)
        lat, span = srv.replay(reqs, arr)
        r = _pcts(lat)
        r["requests_per_s"] = len(reqs) / span
        r["gflops"] = flops / span / 1e9
        r["offered_per_s"] = rate
        return r

    res = {}
    # Blocking per-call: one request on the GPU at a time, as with the
    # reference's host entry points, which synchronise before returning.
    srv, reqs = server("plain", "f32", plain_slots=1)
    res["plain_blocking_f32_burst"] = run(srv, reqs, None)
    srv.close()
    srv, reqs = server("plain", "f32")
    res["plain_f32_burst"] = run(srv, reqs, None)
    cap_plain = res["plain_f32_burst"]["requests_per_s"]
    lam0 = 0.7 * cap_plain
    res["plain_f32_at_load"] = run(srv, reqs, lam0)
    srv.close()
    log(f"plain: capacity {cap_plain:.0f} req/s; at {lam0:.0f}/s "
        f"p99 {res['plain_f32_at_load']['p99_ms']:.2f} ms")
    for vt in ("f32", "bf16"):
        srv, reqs = server("batched", vt)
        res[f"batched_{vt}_burst"] = b = run(srv, reqs, None)
        cap = b["requests_per_s"]
        res[f"batched_{vt}_at_load"] = run(srv, reqs, lam0)
        hi = min(3.0 * lam0, 0.9 * cap)
        res[f"batched_{vt}_at_3x_load"] = run(srv, reqs, hi)
        st = srv.stats()
        res[f"batched_{vt}_mean_batch"] = st["batched_requests"] / max(st["batches"], 1)
        srv.close()
        log(f"batched {vt}: capacity {cap:.0f} req/s; at {lam0:.0f}/s p99 "
            f"{res[f'batched_{vt}_at_load']['p99_ms']:.2f} ms; at {hi:.0f}/s p99 "
            f"{res[f'batched_{vt}_at_3x_load']['p99_ms']:.2f} ms")
    out.update(res)
    out["capacity_ratio_f32"] = res["batched_f32_burst"]["requests_per_s"] / cap_plain
    out["capacity_ratio_f32_vs_blocking"] = (res["batched_f32_burst"]["requests_per_s"] /
                                             res["plain_blocking_f32_burst"]["requests_per_s"])
    out["offered_load_per_s"] = lam0
    return out
