"""Measurement legs of bench.py beside the headline (run on the GPU): the
batched form against per-matrix plain launches, the served form, the SpMM
workload and the bf16 storage plan.  Measurement harness, not part of the
edge_loops_amd package.

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

from edge_loops_amd import synth
from edge_loops_amd.algorithms import run_spmv, spmv_plan, spmv_workspace_size
from edge_loops_amd.batch import Batch, BatchItem
from edge_loops_amd.containers import csr_t

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


def batched(count=512, reps=10, table=None, log=print, plain=True):
    """The batched form (config 4).  plain=False skips the per-call legs (a
    profile of the batch kernels only)."""
    import torch
    from edge_loops_amd import tune
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
    picks = table.pick_batch([(m.rows, m.nnz, m.max_row(), 1) for m in mats])
    hist = {}
    for s, i in picks:
        hist[f"{s}:{i}"] = hist.get(f"{s}:{i}", 0) + 1
    # row-block plans (cut for batches) of the matrices picked for them:
    # built once, like the matrices resident before the passes are timed
    t0 = time.time()
    plans = {j: spmv_plan(dev[j], layout="blocks", for_batches=True)
             for j, (s, _) in enumerate(picks) if s == "blocks"}
    torch.cuda.synchronize()
    plan_s = time.time() - t0

    out = {"workload": f"{count} CSR matrices, rows log-uniform [1e3, 2e5], "
                       "nnz/row Zipf(1.8) capped at 4096, fp32",
           "matrices": count, "rows": rows, "nnz": nnz, "algorithmic_bytes": byts,
           "tuned_picks": hist,
           "plans": {"count": len(plans), "build_s": plan_s,
                     "device_bytes": int(sum(p.device_bytes for p in plans.values())),
                     "csr_bytes_of_those": int(sum(mats[j].nnz * 8 + (mats[j].rows + 1) * 4
                                                   for j in plans))}}
    if plans:
        log(f"batched: {len(plans)} row-block plans built in {plan_s:.2f}s")

    def leg(ms):
        return {"ms_per_pass": ms, "requests_per_s": count / (ms * 1e-3),
                "gflops": flops / (ms * 1e-3) / 1e9,
                "gbs": byts / (ms * 1e-3) / 1e9,
                "roofline_frac": byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    # batched, tuned and fixed: both captured, then timed alternately three
    # times (the median of each), so neither leg pays for the clocks ramping
    legs = {}
    for name, pk in (("batched_tuned", picks),
                     ("batched_merge8", [("merge_path_flat", 8)] * count)):
        b = Batch([BatchItem(A, x, y, s, i, plan=plans.get(j) if s == "blocks" else None)
                   for j, (A, x, y, (s, i)) in enumerate(zip(dev, xs, ys, pk))])
        legs[name] = (b, b.capture())
    times = {name: [] for name in legs}
    for _ in range(3):
        for name, (b, g) in legs.items():
            times[name].append(_time(g.replay, reps))
    for name, (b, g) in legs.items():
        ms = float(np.median(times[name]))
        out[name] = leg(ms)
        out[name]["work_items"] = int(b.info.work_items)
        log(f"{name}: {ms:.3f} ms/pass")
    # error of the tuned batch against the FP64 host product, 16 matrices
    legs["batched_tuned"][1].replay()
    torch.cuda.synchronize()
    bad, worst, checked = 0, 0.0, 0
    for i in np.linspace(0, count - 1, 16).astype(int):
        m = mats[i]
        r64, a64 = _f64_product(m, m.values, xs[i].cpu().numpy())
        got = ys[i].cpu().numpy().astype(np.float64)
        bad += int(np.sum(np.abs(got - r64) > 1e-5 * a64 + 1e-30))
        worst = max(worst, float((np.abs(got - r64) / (a64 + 1e-30)).max()))
        checked += m.rows
    out["batched_tuned_check"] = {"matrices": 16, "rows_checked": checked,
                                  "rows_beyond_1e-5": bad, "worst_rel_err_vs_fp64": worst}
    log(f"batched_tuned: {bad} of {checked} rows of 16 matrices beyond 1e-5 vs FP64, "
        f"worst {worst:.2e}")
    results = {}
    legs["batched_merge8"][1].replay()
    torch.cuda.synchronize()
    results["batched_merge8"] = [y.clone() for y in ys]
    del legs

    if not plain:
        return out
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


def _served_setup(n_requests, seed, vts, table, log):
    """The served pool (synth.c5_pool_64) on the device per value type with
    its tunings, the right-hand sides, and the seeded request mix."""
    import torch
    from edge_loops_amd import tune
    from edge_loops_amd.serve import matrix_tunings
    t0 = time.time()
    mats = synth.c5_pool_64()
    log(f"served: generated 64 matrices in {time.time() - t0:.1f}s")
    table = table or tune.Table.load()
    pools = {}
    for vt in vts:
        dev = [csr_t.from_host(m.rows, m.cols, m.offsets, m.indices, m.values, value_type=vt)
               for m in mats]
        tun = [matrix_tunings(A, table, m.max_row()) for A, m in zip(dev, mats)]
        pools[vt] = (dev, tun)
    x1 = [torch.rand(m.cols, device="cuda") * 2 - 1 for m in mats]
    X32 = [torch.rand(m.cols, 32, device="cuda") * 2 - 1 for m in mats]
    rng = np.random.default_rng(seed)
    mid = rng.integers(0, len(mats), n_requests)
    ks = rng.integers(1, 33, n_requests)
    return mats, pools, x1, X32, mid, ks


def served_burst(mode="batched", n_requests=3000, seed=55, slots=3, log=print, table=None):
    """One burst of the served workload (all requests at t = 0) through the
    server in `mode`, fp32: requests per second.  The leg a kernel trace of
    the served form runs."""
    from edge_loops_amd.serve import Server
    mats, pools, x1, X32, mid, ks = _served_setup(n_requests, seed, ("f32",), table, log)
    dev, tun = pools["f32"]
    srv = Server(mode=mode, arena_bytes=32 << 30, scratch_bytes=1 << 30,
                 slots=slots if mode == "batched" else 64)
    ids = [srv.add_matrix(A, tunings=t) for A, t in zip(dev, tun)]
    reqs = [(ids[m], x1[m] if k == 1 else X32[m][:, :k]) for m, k in zip(mid, ks)]
    lat, span = srv.replay(reqs, np.zeros(len(reqs)))
    srv.close()
    rate = len(reqs) / span
    log(f"served burst, {mode}: {rate:.0f} req/s")
    return rate


def served(n_requests=3000, seed=55, slots=3, log=print, table=None):
    """Served workload: 64 HBM-resident matrices (synth.c5_pool_64), requests
    with a uniform random matrix and k uniform in [1, 32] fp32 right-hand
    sides (k == 1: a vector), seeded Poisson arrivals.  Batched server vs the
    plain per-call mode on the same loop and clock."""
    from edge_loops_amd.serve import Server, poisson_arrivals
    mats, pools, x1, X32, mid, ks = _served_setup(n_requests, seed, ("f32", "bf16"), table, log)
    flops = float(sum(2 * mats[m].nnz * k for m, k in zip(mid, ks)))
    out = {"workload": "64 CSR matrices, rows log-uniform [1e4, 4e6], nnz log-uniform "
                       "[1e5, 8e6], Zipf 1.8-2.2; requests: uniform matrix, k uniform "
                       "[1, 32] fp32 RHS, seeded Poisson arrivals",
           "matrices": len(mats), "pool_nnz": int(sum(m.nnz for m in mats)),
           "pool_rows": int(sum(m.rows for m in mats)), "requests": n_requests,
           "mean_k": float(ks.mean()), "slots": slots}

    def server(mode, vt, plain_slots=64, reference=False, warm=200, batch_all=False):
        srv = Server(mode=mode, arena_bytes=32 << 30, scratch_bytes=1 << 30,
                     slots=slots if mode == "batched" else plain_slots,
                     reference_schedules=reference, batch_all=batch_all)
        dev, tun = pools[vt]
        ids = [srv.add_matrix(A, tunings=t) for A, t in zip(dev, tun)]
        reqs = [(ids[m], x1[m] if k == 1 else X32[m][:, :k]) for m, k in zip(mid, ks)]
        srv.replay(reqs[:warm], np.zeros(warm))  # warm-up
        return srv, reqs

    def run(srv, reqs, rate, work=None):
        arr = np.zeros(len(reqs)) if rate is None else poisson_arrivals(len(reqs), rate, seed + 1)
        lat, span = srv.replay(reqs, arr)
        r = _pcts(lat)
        r["requests_per_s"] = len(reqs) / span
        r["gflops"] = (flops if work is None else work) / span / 1e9
        r["offered_per_s"] = rate
        return r

    res = {}
    # Per-call launches of the reference's own entry points (merge_path_flat
    # SpMV, thread_mapped SpMM -- its only SpMM): burst capacity on the first
    # n_ref requests (thread_mapped walks each long row with one lane).
    n_ref = min(300, n_requests)
    srv, reqs = server("plain", "f32", reference=True, warm=20)
    work_ref = float(sum(2 * mats[m].nnz * k for m, k in zip(mid[:n_ref], ks[:n_ref])))
    res["plain_reference_f32_burst"] = run(srv, reqs[:n_ref], None, work_ref)
    res["plain_reference_f32_burst"]["requests"] = n_ref
    srv.close()
    log(f"plain (reference entry points): capacity "
        f"{res['plain_reference_f32_burst']['requests_per_s']:.0f} req/s")
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
    def check(srv, reqs, vt, n_check=24, n_rows=400):
        """Error of a mixed burst of n_check requests against the FP64 host
        product, on n_rows sampled rows of each (bf16: on the stored
        values): rows beyond the 1e-5 per-row bar and the worst relative
        error."""
        picks = np.linspace(0, len(reqs) - 1, n_check).astype(int)
        tickets = [srv.submit(*reqs[i]) for i in picks]
        bad, worst, checked = 0, 0.0, 0
        srng = np.random.default_rng(seed + 7)
        for i, t in zip(picks, tickets):
            out = srv.result(t)
            A_h, A = mats[mid[i]], pools[vt][0][mid[i]]
            rows = np.sort(srng.choice(A_h.rows, min(n_rows, A_h.rows), replace=False))
            got = out[torch.from_numpy(rows).cuda()].cpu().numpy().astype(np.float64)
            vals = A.values.float().cpu().numpy() if vt != "f32" else A_h.values
            lens = A_h.offsets[rows + 1] - A_h.offsets[rows]
            sub_off = np.zeros(rows.shape[0] + 1, np.int64)
            np.cumsum(lens, out=sub_off[1:])
            idx = np.concatenate([A_h.indices[A_h.offsets[r]:A_h.offsets[r + 1]] for r in rows])
            val = np.concatenate([vals[A_h.offsets[r]:A_h.offsets[r + 1]] for r in rows])
            X = reqs[i][1].cpu().numpy()
            r64, a64 = _f64_product(synth.HostCSR(rows.shape[0], A_h.cols, sub_off, idx, val),
                                    val, X)
            r64, a64 = r64.reshape(got.shape), a64.reshape(got.shape)
            err = np.abs(got - r64) / (a64 + 1e-30)
            bad += int(np.sum(np.any((np.abs(got - r64) > 1e-5 * a64 + 1e-30).reshape(
                rows.shape[0], -1), axis=1)))
            worst = max(worst, float(err.max()) if err.size else 0.0)
            checked += rows.shape[0]
        return {"requests": int(n_check), "rows_checked": checked, "rows_beyond_1e-5": bad,
                "worst_rel_err_vs_fp64": worst}

    import torch
    for vt in ("f32", "bf16"):
        srv, reqs = server("batched", vt)
        res[f"batched_{vt}_check"] = check(srv, reqs, vt)
        log(f"batched {vt}: {res[f'batched_{vt}_check']['rows_beyond_1e-5']} of "
            f"{res[f'batched_{vt}_check']['rows_checked']} sampled rows beyond 1e-5 vs FP64, "
            f"worst {res[f'batched_{vt}_check']['worst_rel_err_vs_fp64']:.2e}")
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
    # every request in a batch (none launched alone): the batch kernels'
    # own capacity on these large requests
    srv, reqs = server("batched", "f32", batch_all=True)
    res["batched_all_f32_burst"] = run(srv, reqs, None)
    srv.close()
    log(f"batched f32, every request in a batch: capacity "
        f"{res['batched_all_f32_burst']['requests_per_s']:.0f} req/s")
    out.update(res)
    out["capacity_ratio_f32"] = res["batched_f32_burst"]["requests_per_s"] / cap_plain
    out["capacity_ratio_f32_vs_blocking"] = (res["batched_f32_burst"]["requests_per_s"] /
                                             res["plain_blocking_f32_burst"]["requests_per_s"])
    out["capacity_ratio_f32_vs_reference_entry_points"] = (
        res["batched_f32_burst"]["requests_per_s"] /
        res["plain_reference_f32_burst"]["requests_per_s"])
    out["offered_load_per_s"] = lam0
    return out


def _f64_product(A_h, values, X):
    """FP64 product A X (X a vector or a matrix) and sum |a_ij x_j| per row,
    with scipy.sparse on the host: the error reference of these legs."""
    import scipy.sparse as sp
    M = sp.csr_matrix((np.asarray(values, np.float64), A_h.indices, A_h.offsets),
                      shape=(A_h.rows, A_h.cols))
    Xd = np.asarray(X, np.float64)
    return M @ Xd, abs(M) @ np.abs(Xd)


def spmv_bf16_plan(A_h, x_h, reps=20, log=print):
    """The 2M-row matrix's SpMV plan with bf16 value storage (fp32
    accumulation): step time, GFLOP/s, algorithmic bytes with 2-byte values,
    and the worst per-row error against an FP64 host product on the stored values
    relative to sum |a x| (the 1e-5 bar)."""
    import torch
    from edge_loops_amd.algorithms import spmv_plan
    A = csr_t.from_host(A_h.rows, A_h.cols, A_h.offsets, A_h.indices, A_h.values,
                        value_type="bf16")
    x = torch.from_numpy(x_h).cuda()
    y = torch.empty(A_h.rows, device="cuda")
    p = spmv_plan(A)
    ms = _time(lambda: p.execute(x, y), reps)
    torch.cuda.synchronize()
    stored = A.values.float().cpu().numpy()
    yg = y.cpu().numpy().astype(np.float64)
    y64, a = _f64_product(A_h, stored, x_h)
    rel = np.abs(yg - y64) / np.maximum(a, 1e-30)
    bad, worst = int(np.sum(np.abs(yg - y64) > 1e-5 * a)), float(rel.max())
    y32v64, a32 = _f64_product(A_h, A_h.values, x_h)
    dev = float(np.max(np.abs(yg - y32v64) / np.maximum(a32, 1e-30)))
    byts = synth.spmv_bytes(A_h.rows, A_h.cols, A_h.nnz, 2)
    out = {"layout": p.layout, "ms_per_step": ms, "gflops": 2 * A_h.nnz / (ms * 1e-3) / 1e9,
           "algorithmic_bytes": byts, "gbs": byts / (ms * 1e-3) / 1e9,
           "rows_beyond_1e-5_vs_fp64_on_stored": int(bad), "worst_rel_err_vs_fp64_on_stored": worst,
           "worst_rel_dev_vs_fp32_values": dev}
    log(f"bf16 plan: {ms*1e3:.1f} us/step, {out['gflops']:.1f} GFLOP/s, worst rel err {worst:.2g}")
    p.destroy()
    return out


def _rows_gather_ceiling(A, B, passes, reps=5):
    """Time of the least reading any CSR-order SpMM does for the given
    column passes ((first column, columns) each): every nonzero's column id
    and value, and its B row's columns of the pass (hbm_probe_gather_rows),
    summed over the passes, each the best of 1-16 workgroups per CU and 2-8
    nonzeros in flight per lane group (ms)."""
    import ctypes
    import os

    import torch
    root = os.path.dirname(os.path.abspath(__file__))
    lib = ctypes.CDLL(os.path.join(root, "edge_loops_amd", "lib", "libloops_hbm.so"))
    lib.hbm_probe_gather_rows.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_longlong,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    assert A.indices.dtype == torch.int32 and A.values.dtype == torch.float32
    assert B.dtype == torch.float32 and B.is_contiguous() and B.shape[0] >= A.cols
    assert B.data_ptr() % 16 == 0 and B.shape[1] % 4 == 0
    sink = torch.zeros(256, dtype=torch.int32, device="cuda")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    stream = torch.cuda.current_stream().cuda_stream
    total = 0.0
    for c0, k in passes:
        assert c0 % 4 == 0 and c0 + k <= B.shape[1]

        def run(g, d):
            rc = lib.hbm_probe_gather_rows(A.indices.data_ptr(), A.values.data_ptr(), A.nnzs,
                                           B.data_ptr(), B.shape[1], c0, k, g, d,
                                           sink.data_ptr(), stream)
            if rc != 0:
                raise RuntimeError(f"row gather probe failed: hipError {rc}")
        total += min(_time(lambda: run(cus * per_cu, d), reps)
                     for per_cu in (1, 2, 4, 8, 16) for d in (2, 4, 8))
    return total


def spmm_c3(reps=5, log=print):
    """SpMM C = A B on the 1M-row power-law matrix (Zipf 2.0, 16M nonzeros)
    with a dense 1M x 64 fp32 B: group_mapped and merge_path, warm, with the
    per-row 1e-5 check of a sample of rows against an FP64 host product."""
    import torch
    from edge_loops_amd.algorithms import run_spmm, spmm_workspace_size
    t0 = time.time()
    A_h = synth.c3_spmm_1m()
    k = 64
    B_h = synth.dense(A_h.cols, k, seed=31)
    log(f"spmm c3: generated rows={A_h.rows} nnz={A_h.nnz} in {time.time() - t0:.1f}s")
    A = csr_t.from_host(A_h.rows, A_h.cols, A_h.offsets, A_h.indices, A_h.values)
    B = torch.from_numpy(B_h).cuda()
    C = torch.empty(A_h.rows, k, device="cuda")
    byts = synth.spmm_bytes(A_h.rows, A_h.cols, A_h.nnz, k)
    flops = 2 * A_h.nnz * k
    out = {"workload": "1M x 1M CSR, Zipf(2.0) rows scaled to 16M nonzeros, B 1M x 64 fp32",
           "rows": A_h.rows, "nnz": A_h.nnz, "k": k, "algorithmic_bytes": byts}
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(A_h.rows, 2000, replace=False))
    for sched in ("group_mapped", "merge_path_flat"):
        ws = torch.empty(max(spmm_workspace_size(sched, A, k), 256), dtype=torch.uint8,
                         device="cuda")
        ms = _time(lambda: run_spmm(sched, A, B, C, workspace=ws), reps)
        torch.cuda.synchronize()
        Cs = C.cpu().numpy()[sample]
        # FP64 host product of the sampled rows only
        offs = A_h.offsets
        sub_off = np.zeros(sample.shape[0] + 1, np.int64)
        lens = offs[sample + 1] - offs[sample]
        np.cumsum(lens, out=sub_off[1:])
        idx = np.concatenate([A_h.indices[offs[r]:offs[r + 1]] for r in sample])
        val = np.concatenate([A_h.values[offs[r]:offs[r + 1]] for r in sample])
        r64, a64 = _f64_product(synth.HostCSR(sample.shape[0], A_h.cols, sub_off, idx, val), val,
                                B_h)
        bad = int(np.sum(np.abs(Cs - r64) > 1e-5 * a64 + 1e-30))
        out[sched] = {"ms": ms, "gflops": flops / (ms * 1e-3) / 1e9,
                      "gbs": byts / (ms * 1e-3) / 1e9,
                      "roofline_frac": byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "sampled_rows_beyond_1e-5": bad}
        log(f"spmm c3 {sched}: {ms:.3f} ms, {out[sched]['gflops']:.0f} GFLOP/s, "
            f"{out[sched]['gbs']:.0f} GB/s, bad rows {bad}")
    # the B-row gathers any CSR-order product of A and B makes: two 32-column
    # passes (how merge_path runs k = 64) and one 64-column pass
    g32 = _rows_gather_ceiling(A, B, [(0, 32), (32, 32)], reps)
    g64 = _rows_gather_ceiling(A, B, [(0, 64)], reps)
    best = min(g32, g64)
    out["gather_ceiling"] = {
        "two_32_column_passes_ms": g32, "one_64_column_pass_ms": g64,
        "what": "hbm_probe_gather_rows: each nonzero's column id, value and B-row columns "
                "of the pass (16-byte loads), no row walk or C; best of 1-16 workgroups per "
                "CU and 2-8 nonzeros in flight per lane group, warm like the timed legs"}
    for sched in ("group_mapped", "merge_path_flat"):
        out[sched]["frac_of_gather_ceiling"] = best / out[sched]["ms"]
    log(f"spmm c3 gather ceiling: {g32:.3f} ms (2 x 32 columns), {g64:.3f} ms (64 columns)")
    return out
