#!/usr/bin/env python3
"""Benchmark of the filter -> group-by -> join -> top-K device path (one GPU).

Default workload `batch` (the metric's configuration): a resident table of
500M rows (k1 i32 U[0,4), k2 i64 U[0,1e7) stored in k2 order, v1 i32 over the
full int32 range, v2 f64 U[0,1)) and 256 queries, each filtering v2 to
[lo, lo + s) with s log-uniform in [1e-4, 0.5], grouping by k1 or k2 and
computing SUM(v1), COUNT(*), MIN(v1), MAX(v2). A step is one batched
execution of all 256 queries (qm_batch_exec, inputs resident in HBM, results
device-resident); `value` is the batch's aggregate rows/s (256 x 500M rows per
step). The same 256 queries run one by one through qm_exec give the per-query
drop-in rate reported beside it.

`--workload zipf`: the single-query drop-in over 100M rows with Zipf keys
(filter a < 300000 AND b > 90 AND flag = 1 -> group by key -> SUM(b),
COUNT(*), MIN(a), MAX(a), AVG(b)); a step is one qm_exec.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "batched vs per-query rows/s + HBM % of peak; p50/p99 ms at fixed QPS, 1 GPU"


def copy_calibration(nbytes: int) -> None:
    """One plain device-to-device copy of `nbytes` (the runtime's copy kernel)
    through the HIP runtime the library is linked against."""
    hip = C.CDLL("libamdhip64.so.7")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def _traffic(kname: str, rows: int, fname: str):
    tfile = ROOT / "profiles" / fname
    if tfile.exists():
        rec = json.loads(tfile.read_text())
        if rec.get("kernel") == kname and rec.get("rows") == rows:
            return rec.get("hbm_bytes_per_launch")
    return None


def _exec_stats(lib, N, plan, table):
    rh = C.c_void_p()
    N.check(lib.qm_exec(plan.handle, table.handle, None, C.byref(rh)))
    st = N.Stats()
    N.check(lib.qm_result_stats(rh, C.byref(st)))
    lib.qm_result_close(rh)
    return st


def bench_zipf(args) -> dict:
    import numpy as np
    from querymode_amd import _native as N
    from querymode_amd import datagen
    from querymode_amd.client import Device, _Plan
    from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

    rows = args.rows or 100_000_000
    t0 = time.time()
    cols = datagen.filter_groupby_zipf(n=rows, seed=0)
    gen_s = time.time() - t0
    desc = QueryDescriptor("t", filters=[FilterOp("a", "lt", 300000), FilterOp("b", "gt", 90), FilterOp("flag", "eq", 1)],
                           groupBy=["key"],
                           aggregates=[AggregateOp("sum", "b"), AggregateOp("count", "*"), AggregateOp("min", "a"),
                                       AggregateOp("max", "a"), AggregateOp("avg", "b")])
    dev = Device(0)
    t0 = time.time()
    table = dev.register("t", cols)
    upload_s = time.time() - t0
    plan = _Plan(dev, desc)
    for _ in range(args.warmup):
        _exec_stats(dev.lib, N, plan, table)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(_exec_stats(dev.lib, N, plan, table))
    wall = time.perf_counter() - t0
    scan_ms = float(np.mean([s.scan_ms for s in stats]))
    alg_bytes = stats[-1].bytes_algorithmic
    achieved = alg_bytes / (scan_ms * 1e-3) / 1e9
    strategy = N.STRATEGIES.get(stats[-1].strategy)
    kname = "part_scan_fast_kernel<3, 1>" if strategy in ("part_hash", "part_direct") else "scan_agg_kernel"
    if args.copy_calibration:
        copy_calibration(1 << 30)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        n = min(args.cpu_sample_rows or rows, rows)
        t = O.Table({k: v[:n] for k, v in cols.items()})
        t1 = time.perf_counter()
        O.run(desc, {"t": t})
        cpu_s = time.perf_counter() - t1
        cpu = {"value": n / cpu_s, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"{'all' if n == rows else 'first'} {n} rows of the same table, same query: C "
                         f"restatement of the reference CPU path (per-page index-list filter + sorted "
                         f"intersection + hash aggregate), one thread, {cpu_s:.2f} s"}
    out = {
        "metric": METRIC, "value": rows / (wall / args.steps), "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f64",
        "data": "synthetic (deterministic splitmix64 generator, seed 0)",
        "config": {"workload": "single-query drop-in, 100M rows, Zipf(1.1) keys over 1M ids (filter_groupby_zipf)",
                   "rows": rows, "groups": int(stats[-1].groups), "selected": int(stats[-1].rows_selected),
                   "strategy": strategy, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": _traffic(kname, rows, "traffic_filter_groupby_zipf.json")},
        "kernel": {"name": kname, "avg_ms": scan_ms, "algorithmic_bytes": alg_bytes,
                   "query_device_ms": float(np.mean([s.device_ms for s in stats]))},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s, "upload_s": upload_s},
    }
    plan.close()
    dev.close()
    return out


def bench_batch(args) -> dict:
    import numpy as np
    from querymode_amd import _native as N
    from querymode_amd import datagen
    from querymode_amd.client import Device, _Plan

    rows = args.rows or 500_000_000
    nq = args.queries
    t0 = time.time()
    cols = datagen.batch_table(n=rows, seed=0, k2_order=True)
    gen_s = time.time() - t0
    descs = datagen.batch_queries(nq, seed=0)
    dev = Device(0)
    lib = dev.lib
    t0 = time.time()
    table = dev.register("batch", cols)
    upload_s = time.time() - t0
    plans = [_Plan(dev, d) for d in descs]
    arr = (C.c_void_p * nq)(*[p.handle for p in plans])

    def batch_step():
        outs = (C.c_void_p * nq)()
        N.check(lib.qm_batch_exec(dev.handle, arr, nq, table.handle, outs))
        st = N.Stats()
        N.check(lib.qm_result_stats(C.c_void_p(outs[0]), C.byref(st)))
        groups = 0
        for i in range(nq):
            s2 = N.Stats()
            lib.qm_result_stats(C.c_void_p(outs[i]), C.byref(s2))
            groups += s2.groups
            lib.qm_result_close(C.c_void_p(outs[i]))
        return st, groups

    for _ in range(args.warmup):
        batch_step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(batch_step())
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    batched_rows_s = nq * rows / (wall / args.steps)
    scan_ms = float(np.mean([s.scan_ms for s, _ in stats]))
    dev_ms = float(np.mean([s.device_ms for s, _ in stats]))
    alg_bytes = rows * sum(np.dtype(c.dtype).itemsize for c in cols.values())  # every column read once
    achieved = alg_bytes / (scan_ms * 1e-3) / 1e9
    # the per-query drop-in: the same queries one at a time (qm_exec)
    for p in plans[:4]:
        _exec_stats(lib, N, p, table)
    t0 = time.perf_counter()
    pq = [_exec_stats(lib, N, p, table) for p in plans]
    pq_wall = time.perf_counter() - t0
    per_query_rows_s = nq * rows / pq_wall
    strategies = {}
    for s in pq:
        k = N.STRATEGIES.get(s.strategy)
        strategies[k] = strategies.get(k, 0) + 1
    if args.copy_calibration:
        copy_calibration(1 << 30)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        n = min(args.cpu_sample_rows or 2_000_000, rows)
        t = O.Table({k: v[:n] for k, v in cols.items()})
        t1 = time.perf_counter()
        for d in descs:
            O.run(d, {"batch": t})
        cpu_s = time.perf_counter() - t1
        cpu = {"value": nq * n / cpu_s, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"all {nq} queries over the first {n} rows of the same table: C restatement of the "
                         f"reference CPU path (per-page index-list filter + hash aggregate per query), one "
                         f"thread, {cpu_s:.2f} s"}
    out = {
        "metric": METRIC, "value": batched_rows_s, "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f64",
        "data": "synthetic (deterministic splitmix64 generator, seed 0); table stored in k2 order",
        "config": {"workload": f"batched {nq}-query mix over a resident {rows // 1_000_000}M-row table "
                               "(range filter on v2, GROUP BY k1 or k2, SUM/COUNT/MIN/MAX)",
                   "rows": rows, "queries": nq, "groups_per_batch": int(stats[-1][1]),
                   "k2_order": True, "parallelism": "single GPU"},
        "batched_rows_per_s": batched_rows_s,
        "per_query_rows_per_s": per_query_rows_s,
        "batched_vs_per_query": batched_rows_s / per_query_rows_s,
        "per_query_ms_total": pq_wall * 1e3,
        "per_query_strategies": strategies,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": _traffic("batch_kernel<4>", rows, "traffic_batch_mix.json")},
        "kernel": {"name": "batch_kernel<4>", "avg_ms": scan_ms, "algorithmic_bytes": alg_bytes,
                   "batch_device_ms": dev_ms},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s, "upload_s": upload_s},
    }
    for p in plans:
        p.close()
    dev.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["batch", "zipf"], default="batch")
    ap.add_argument("--rows", type=int, default=0, help="table rows (default: the workload's)")
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--cpu-sample-rows", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, run one 1 GiB device-to-device copy (the known-size dispatch the "
                         "FETCH_SIZE / WRITE_SIZE counters are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: run it with --gpus 1 (this build has no multi-GPU path)",
              file=sys.stderr)
        return 2
    out = bench_batch(args) if args.workload == "batch" else bench_zipf(args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
