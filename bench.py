#!/usr/bin/env python3
"""Benchmark: merge-path SpMV on the 2M-row power-law matrix, one MI355X.

Workload ("c2_powerlaw_2m"): y = A x, A 2,000,000 x 2,000,000 CSR with ~40M
nonzeros, Zipf(1.8) row lengths on [1, 200000] scaled to 40M, row 0 forced
to 200,000; columns uniform and sorted per row; int32 indices; fp32 values
and x uniform [-1, 1).  Synthetic data (seeded generator), resident in HBM
before timing.

Unless --quick / --no-forms, the line also carries the batched form (512
matrices in one persistent launch vs per-matrix plain launches) and the
served form (continuous batching over 64 resident matrices vs plain
per-call launches: req/s, p50/p99), see bench_forms.py.

A step is one execution of the matrix's SpMV plan (loops_spmv_plan_execute;
the automatic layout for this matrix is row blocks: one workgroup per (row
block, column segment) stream accumulating y in LDS, then the combine of the
segments' partial rows), built once per matrix beforehand (its build time is
reported; --layout panels selects the column-panel merge-path plan).  The timed steps rotate over several identical
copies of A (working set > the 256 MB MALL) so no step is served from the
previous step's cache lines; the same-copy (warm) figure and the plain
per-call drop-in (loops_spmv, no plan) are reported beside it.

Prints one JSON line (value = GFLOP/s = 2 nnz / step time) with a roofline
object for the main kernel (algorithmic bytes / its average duration from
HIP events recorded around it on its stream), the CPU baseline (the
oracle's sequential validator loop on the host), the device's properties,
measured HBM ceilings (2 GiB streamed and copied by probe kernels) and the
CSR-order gather ceiling of this matrix beside the per-call merge-path leg.
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--copies", type=int, default=4)
    ap.add_argument("--tpb", type=int, default=0)
    ap.add_argument("--ipt", type=int, default=0)
    ap.add_argument("--panels", type=int, default=0, help="column panels (0: automatic)")
    ap.add_argument("--layout", default="auto", choices=["auto", "panels", "blocks"],
                    help="SpMV plan layout (auto: blocks for this matrix)")
    ap.add_argument("--schedules", default="all",
                    help="comma list of extra schedules to time, 'all' or 'none'")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="skip extras (profiling runs)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched (512 matrices) and served (64 matrices) legs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes (--steps times); "
                         "used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    return ap.parse_args()


def copy_calibration(mb, steps):
    """A plain copy kernel of known size: reads and writes mb * 2^20 bytes
    per launch, the reference point for the PMC byte counters."""
    import torch
    n = mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": n * 4, "launches": steps}), flush=True)


def traffic_from_profile(layout):
    """HBM bytes per launch of the main kernel of a layout ("blocks",
    "panels" or "merge_path_plain") on this workload, measured by rocprofv3
    PMC runs (FETCH_SIZE, WRITE_SIZE, calibrated against a copy of known
    size) and recorded in profiles/; None if absent."""
    name = {"blocks": "spmv_blocks_c2_pmc.md", "panels": "merge_path_plan_c2_pmc.md",
            "merge_path_plain": "merge_path_plain_c2_pmc.md"}[layout]
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


def hbm_ceilings(stream):
    """Read and copy rates of 2 GiB of device memory through the library's
    probe kernels (edge_loops_amd/csrc/hbm_probe.hip), the best over grids
    of 1-16 workgroups per CU, and torch's copy_ / sum for comparison."""
    import ctypes

    import torch

    from edge_loops_amd.hip import Event
    lib = ctypes.CDLL(os.path.join(ROOT, "edge_loops_amd", "lib", "libloops_hbm.so"))
    lib.hbm_probe_read.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p]
    lib.hbm_probe_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_void_p]
    nbytes = 2 << 30
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    sink = torch.zeros(256, dtype=torch.int32, device="cuda")
    cus = torch.cuda.get_device_properties(0).multi_processor_count

    def timed(fn, reps=8):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_ms(e1) / reps

    def ok(rc):
        if rc != 0:
            raise RuntimeError(f"hbm probe failed: hipError {rc}")

    best_read, best_copy = (0.0, 0), (0.0, 0)
    for per_cu in (1, 2, 4, 8, 16):
        g = cus * per_cu
        ms = timed(lambda: ok(lib.hbm_probe_read(src.data_ptr(), nbytes, g, sink.data_ptr(),
                                                 stream)))
        best_read = max(best_read, (nbytes / (ms * 1e-3) / 1e9, g))
        ms = timed(lambda: ok(lib.hbm_probe_copy(dst.data_ptr(), src.data_ptr(), nbytes, g,
                                                 stream)))
        best_copy = max(best_copy, (2 * nbytes / (ms * 1e-3) / 1e9, g))
    t_copy = timed(lambda: dst.copy_(src))
    t_sum = timed(lambda: torch.sum(src))
    out = {"read_gbs": round(best_read[0], 1), "read_grid": best_read[1],
           "copy_gbs": round(best_copy[0], 1), "copy_grid": best_copy[1],
           "torch_copy_gbs": round(2 * nbytes / (t_copy * 1e-3) / 1e9, 1),
           "torch_sum_gbs": round(nbytes / (t_sum * 1e-3) / 1e9, 1),
           "what": "2 GiB streamed by hbm_probe.hip (16-byte non-temporal loads, 8 in flight "
                   "per lane; copy: bytes read + written), best grid of 1-16 workgroups "
                   "per CU; torch copy_ and sum over the same 2 GiB beside it; HIP events "
                   "on the bench stream"}
    del src, dst
    torch.cuda.empty_cache()
    return out


def csr_gather_ceiling(copies, x, stream, reps=10):
    """How fast this GPU reads a CSR matrix's nonzeros in CSR order: the
    column ids and values streamed (16-byte non-temporal loads) and x
    gathered at every id, without the row walk or y (hbm_probe_gather),
    rotating over the copies of A like the timed legs; the best of 1-16
    workgroups per CU and 1-8 pairs of loads in flight per lane.  Any SpMV
    that reads A in CSR order does at least this per nonzero."""
    import ctypes

    import torch

    from edge_loops_amd.hip import Event
    lib = ctypes.CDLL(os.path.join(ROOT, "edge_loops_amd", "lib", "libloops_hbm.so"))
    lib.hbm_probe_gather.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]
    sink = torch.zeros(256, dtype=torch.int32, device="cuda")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    for A in copies:
        assert A.indices.dtype == torch.int32 and A.values.dtype == torch.float32
        assert A.indices.data_ptr() % 16 == 0 and A.values.data_ptr() % 16 == 0
        assert x.numel() >= A.cols

    def run(i, g, depth):
        A = copies[i % len(copies)]
        rc = lib.hbm_probe_gather(A.indices.data_ptr(), A.values.data_ptr(), A.nnzs,
                                  x.data_ptr(), g, depth, sink.data_ptr(), stream)
        if rc != 0:
            raise RuntimeError(f"gather probe failed: hipError {rc}")

    best = (float("inf"), 0, 0)
    for depth in (1, 2, 4, 8):
        for per_cu in (1, 2, 4, 8, 16):
            g = cus * per_cu
            for i in range(2):
                run(i, g, depth)
            torch.cuda.synchronize()
            e0, e1 = Event(), Event()
            e0.record(stream)
            for i in range(reps):
                run(i, g, depth)
            e1.record(stream)
            torch.cuda.synchronize()
            best = min(best, (e0.elapsed_ms(e1) / reps, g, depth))
    return {"ms": round(best[0], 5), "grid": best[1], "depth": best[2],
            "what": "hbm_probe_gather: A's column ids and values streamed, x gathered at "
                    "each id, no row walk or y; best of 1-16 workgroups per CU and 1-8 "
                    "pairs of 16-byte loads (4-32 gathers) in flight per lane, rotating "
                    "over the copies of A"}


def merge_path_legs(copies, x, y, stream, args, nbytes, flops):
    """The reference's named path on the same matrix: per-call
    loops_spmv(merge_path_flat) with HIP events around its main kernel
    (rotating over the copies of A), and the column-panel plan."""
    import torch

    import edge_loops_amd as loops
    from edge_loops_amd.hip import Event
    out = {}
    ws = loops.algorithms.workspace("merge_path_flat", copies[0])
    evs = [(Event(), Event()) for _ in range(args.steps)]
    for i in range(args.warmup):
        loops.run_spmv("merge_path_flat", copies[i % len(copies)], x, y, stream=stream,
                       workspace=ws)
    torch.cuda.synchronize()
    p0, p1 = Event(), Event()
    p0.record(stream)
    for i in range(args.steps):
        loops.run_spmv("merge_path_flat", copies[i % len(copies)], x, y, stream=stream,
                       workspace=ws, start_event=evs[i][0], stop_event=evs[i][1])
    p1.record(stream)
    torch.cuda.synchronize()
    step = p0.elapsed_ms(p1) / args.steps
    kern = sum(a.elapsed_ms(b) for a, b in evs) / len(evs)
    out["per_call"] = {
        "kernel": "spmv_merge_path_kernel<256, 8, float>", "kernel_avg_ms": round(kern, 5),
        "ms_per_step": round(step, 5), "gflops": round(flops / (step * 1e-3) / 1e9, 2),
        "frac": round(nbytes / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "frac_step": round(nbytes / (step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile("merge_path_plain")}
    out["csr_gather_ceiling"] = csr_gather_ceiling(copies, x, stream)
    out["per_call"]["frac_of_csr_gather_ceiling"] = round(
        out["csr_gather_ceiling"]["ms"] / kern, 4)
    plan = loops.spmv_plan(copies[0], layout="panels", stream=stream)
    for _ in range(args.warmup):
        plan.execute(x, y, stream)
    evs = [(Event(), Event()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    p0.record(stream)
    for i in range(args.steps):
        plan.execute(x, y, stream, evs[i][0], evs[i][1])
    p1.record(stream)
    torch.cuda.synchronize()
    step = p0.elapsed_ms(p1) / args.steps
    kern = sum(a.elapsed_ms(b) for a, b in evs) / len(evs)
    out["panels_plan"] = {
        "kernel": "spmv_panels_kernel", "panels": plan.panels, "kernel_avg_ms": round(kern, 5),
        "ms_per_step": round(step, 5), "gflops": round(flops / (step * 1e-3) / 1e9, 2),
        "frac": round(nbytes / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "frac_step": round(nbytes / (step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile("panels")}
    plan.destroy()
    log(f"merge_path per call: kernel {out['per_call']['kernel_avg_ms']*1e3:.1f} us "
        f"(frac {out['per_call']['frac']}); CSR-order gather ceiling "
        f"{out['csr_gather_ceiling']['ms']*1e3:.1f} us; panels plan "
        f"{out['panels_plan']['kernel_avg_ms']*1e3:.1f} us")
    return out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, args.steps)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build "
              "does not run multi-GPU jobs.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import edge_loops_amd as loops
    from edge_loops_amd import synth
    from edge_loops_amd.hip import Event

    log("generating c2_powerlaw_2m")
    t0 = time.time()
    A_h = synth.c2_powerlaw_2m()
    x_h = synth.uniform(A_h.cols, seed=21)
    R, C, N = A_h.rows, A_h.cols, A_h.nnz
    log(f"rows={R} cols={C} nnz={N} max_row={A_h.max_row()} ({time.time()-t0:.1f}s)")

    torch.cuda.init()
    stream = torch.cuda.current_stream().cuda_stream
    copies = [loops.csr_t.from_host(R, C, A_h.offsets, A_h.indices, A_h.values)
              for _ in range(max(1, args.copies))]
    x = torch.from_numpy(x_h).cuda()
    y = torch.empty(R, dtype=torch.float32, device="cuda")
    tuning = {}
    if args.tpb:
        tuning["threads_per_block"] = args.tpb
    if args.ipt:
        tuning["items_per_thread"] = args.ipt
    torch.cuda.synchronize()
    t0 = time.time()
    plans = [loops.spmv_plan(A, panels=args.panels, stream=stream, layout=args.layout, **tuning)
             for A in copies]
    torch.cuda.synchronize()
    build_ms = (time.time() - t0) * 1e3 / len(plans)
    log(f"plans: {plans[0].layout}, {plans[0].panels} panels/segments, "
        f"{plans[0].device_bytes/1e6:.0f} MB each, "
        f"{build_ms:.1f} ms to build")

    n_panels = plans[0].panels
    layout = plans[0].layout
    n_copies = len(copies)
    nbytes = synth.spmv_bytes(R, C, N, 4)
    flops = 2 * N

    log(f"warmup {args.warmup}")
    for i in range(args.warmup):
        plans[i % len(plans)].execute(x, y, stream)
    torch.cuda.synchronize()

    # Timed region: K steps, events around each main kernel.
    evs = [(Event(), Event()) for _ in range(args.steps)]
    t_start, t_stop = Event(), Event()
    torch.cuda.synchronize()
    t_start.record(stream)
    for i in range(args.steps):
        plans[i % len(plans)].execute(x, y, stream, evs[i][0], evs[i][1])
    t_stop.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_ms(t_stop)
    ms_per_step = total_ms / args.steps
    kern_ms = [a.elapsed_ms(b) for a, b in evs]
    kern_avg = sum(kern_ms) / len(kern_ms)
    value = flops / (ms_per_step * 1e-3) / 1e9
    log(f"cold: {ms_per_step*1e3:.1f} us/step, kernel {kern_avg*1e3:.1f} us, "
        f"{value:.1f} GFLOP/s, {nbytes/kern_avg/1e6:.0f} GB/s")

    extras = {"plan": {"layout": layout, "panels": plans[0].panels, "blocks": plans[0].blocks,
                       "build_ms": round(build_ms, 2),
                       "device_bytes": plans[0].device_bytes}}
    if not args.quick:
        # warm: the same copy every step
        w0, w1 = Event(), Event()
        torch.cuda.synchronize()
        w0.record(stream)
        for i in range(args.steps):
            plans[0].execute(x, y, stream)
        w1.record(stream)
        torch.cuda.synchronize()
        warm_ms = w0.elapsed_ms(w1) / args.steps
        extras["warm_ms_per_step"] = round(warm_ms, 5)
        log(f"warm: {warm_ms*1e3:.1f} us/step")

        # measured HBM ceilings beside the 8 TB/s peak: this library's own
        # stream kernels (16-byte non-temporal loads / stores, the access of
        # the SpMV kernels' streamed arrays) over 2 GiB, best of a grid sweep;
        # torch's copy and reduction beside them
        extras["hbm_ceilings"] = hbm_ceilings(stream)
        props = torch.cuda.get_device_properties(0)
        extras["device"] = {"name": props.name,
                            "arch": getattr(props, "gcnArchName", ""),
                            "compute_units": props.multi_processor_count,
                            "memory_gib": round(props.total_memory / 2 ** 30, 1)}
        log(f"ceilings: read {extras['hbm_ceilings']['read_gbs']} GB/s, copy "
            f"{extras['hbm_ceilings']['copy_gbs']} GB/s (read + written)")

        # the drop-in itself: per-call merge_path_flat main kernel (events
        # around it) and the column-panel plan, rotating over the copies
        extras["merge_path"] = merge_path_legs(copies, x, y, stream, args, nbytes, flops)
        extras["plain_per_call_ms_per_step"] = extras["merge_path"]["per_call"]["ms_per_step"]
        extras["plain_per_call_gflops"] = extras["merge_path"]["per_call"]["gflops"]

        # all five schedules through loops_spmv, same matrix (copy 0)
        names = ["merge_path_flat", "work_oriented", "group_mapped", "thread_mapped", "original"]
        if args.schedules == "none":
            names = []
        elif args.schedules != "all":
            names = args.schedules.split(",")
        sched = {}
        for s in names:
            wss = loops.algorithms.workspace(s, copies[0])
            loops.run_spmv(s, copies[0], x, y, stream=stream, workspace=wss)
            torch.cuda.synchronize()
            reps = 10 if s in ("merge_path_flat", "work_oriented") else 3
            e0, e1 = Event(), Event()
            e0.record(stream)
            for _ in range(reps):
                loops.run_spmv(s, copies[0], x, y, stream=stream, workspace=wss)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_ms(e1) / reps
            sched[s] = {"ms": round(ms, 5), "gflops": round(flops / (ms * 1e-3) / 1e9, 2),
                        "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1)}
            log(f"schedule {s}: {ms*1e3:.1f} us")
        extras["schedules_warm"] = sched

    # CPU baseline leg: the oracle's sequential validator loop
    # (examples/spmv/helpers.hxx:106-112), timed on this host.
    import oracle
    log("cpu baseline")
    runs = []
    for _ in range(args.cpu_runs):
        c0 = time.perf_counter()
        oracle.spmv_f32(A_h.offsets, A_h.indices, A_h.values, x_h)
        runs.append(time.perf_counter() - c0)
    cpu_s = statistics.median(runs)
    cpu_gflops = flops / cpu_s / 1e9
    # every core this process may run on (its affinity set: the host's cores
    # as the box grants them), one thread per core
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    mt = max(1, min(256, ncores))
    runs_mt = []
    for _ in range(args.cpu_runs):
        c0 = time.perf_counter()
        oracle.spmv_f32_mt(A_h.offsets, A_h.indices, A_h.values, x_h, mt)
        runs_mt.append(time.perf_counter() - c0)
    cpu_mt_s = statistics.median(runs_mt)
    log(f"cpu: 1 core {cpu_s*1e3:.1f} ms, {mt} threads {cpu_mt_s*1e3:.1f} ms")

    if not args.quick and not args.no_forms:
        # bf16 value storage on the same matrix, the SpMM workload, then the
        # batched and served forms against plain per-call launches.
        import bench_forms
        del copies, plans
        torch.cuda.empty_cache()
        extras["bf16_storage_plan"] = bench_forms.spmv_bf16_plan(A_h, x_h, log=log)
        torch.cuda.empty_cache()
        extras["spmm_c3"] = bench_forms.spmm_c3(log=log)
        torch.cuda.empty_cache()
        extras["batched"] = bench_forms.batched(log=log)
        torch.cuda.empty_cache()
        extras["served"] = bench_forms.served(log=log)

    achieved = nbytes / (kern_avg * 1e-3) / 1e9
    out = {
        "metric": "SpMV GFLOP/s + % of HBM roofline; served req/s & p50/p99 vs plain per-call",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"c2_powerlaw_2m SpMV, loops_spmv_plan ({layout} layout; "
                        "per-call merge_path_flat and the five schedules beside it)",
            "rows": R, "cols": C, "nnz": N, "max_row": A_h.max_row(),
            "zipf": 1.8, "index": "int32", "values": "f32",
            "plan_layout": layout,
            "column_panels" if layout == "panels" else "column_segments": n_panels,
            "rotated_copies": n_copies, "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "frac_step": round(nbytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(layout),
            "kernel": ("spmv_blocks_kernel" if layout == "blocks" else
                       "spmv_panels_kernel" if n_panels > 1 else "spmv_merge_path_kernel"),
            "kernel_avg_ms": round(kern_avg, 5),
            "kernel_median_ms": round(statistics.median(kern_ms), 5),
            "kernel_launches": len(kern_ms),
            "algorithmic_bytes": nbytes,
        },
        "cpu_baseline": {
            "value": round(cpu_gflops, 3),
            "unit": "GFLOP/s",
            "cores": 1,
            "kind": "port",
            "sample": f"full c2 matrix ({N} nnz), median of {args.cpu_runs} runs of the "
                      "oracle's sequential fp32 validator loop",
            "host_cores": ncores,
            "host_cpus_online": os.cpu_count(),
            "all_cores": {"threads": mt, "value": round(flops / cpu_mt_s / 1e9, 3),
                          "what": "the same loop, rows split by equal nonzeros over one "
                                  "thread per host core"},
        },
    }
    if "hbm_ceilings" in extras:
        out["roofline"]["frac_of_measured_read"] = round(
            achieved / extras["hbm_ceilings"]["read_gbs"], 4)
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
