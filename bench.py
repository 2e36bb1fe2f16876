#!/usr/bin/env python3
"""Benchmark: the single-query drop-in over 100M resident rows.

Workload (filter_groupby_zipf, 100M rows): key i64 Zipf(1.1) over 1M ids,
a i32 U[0,1e6), b f64 N(100,25), flag i8 Bernoulli(0.5);
filter(a < 300000 AND b > 90 AND flag = 1) -> group by key ->
SUM(b), COUNT(*), MIN(a), MAX(a), AVG(b).

A step is one execution of that query over the resident table (inputs already
in HBM): fused filter+aggregate kernel, group-table compaction and
finalisation, device-resident result. Prints one JSON line.
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


def copy_calibration(nbytes: int) -> None:
    """One plain device-to-device copy of `nbytes` (the runtime's copy kernel)
    through the HIP runtime the library is linked against."""
    import ctypes as C
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--cpu-sample-rows", type=int, default=100_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, run one 1 GiB device-to-device copy (the known-size dispatch the "
                         "FETCH_SIZE / WRITE_SIZE counters are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md for how it would shard)",
              file=sys.stderr)
        return 2

    import numpy as np
    from querymode_amd import _native as N
    from querymode_amd import datagen
    from querymode_amd.client import Device, _Plan
    from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

    t0 = time.time()
    cols = datagen.filter_groupby_zipf(n=args.rows, seed=0)
    gen_s = time.time() - t0
    desc = QueryDescriptor("t", filters=[FilterOp("a", "lt", 300000), FilterOp("b", "gt", 90), FilterOp("flag", "eq", 1)],
                           groupBy=["key"],
                           aggregates=[AggregateOp("sum", "b"), AggregateOp("count", "*"), AggregateOp("min", "a"),
                                       AggregateOp("max", "a"), AggregateOp("avg", "b")])
    dev = Device(0)
    lib = dev.lib
    t0 = time.time()
    table = dev.register("t", cols)
    upload_s = time.time() - t0
    plan = _Plan(dev, desc)

    def step():
        rh = C.c_void_p()
        N.check(lib.qm_exec(plan.handle, table.handle, None, C.byref(rh)))
        st = N.Stats()
        N.check(lib.qm_result_stats(rh, C.byref(st)))
        lib.qm_result_close(rh)
        return st

    for _ in range(args.warmup):
        step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    rows_per_s = args.rows / (wall / args.steps)
    scan_ms = float(np.mean([s.scan_ms for s in stats]))
    dev_ms = float(np.mean([s.device_ms for s in stats]))
    alg_bytes = stats[-1].bytes_algorithmic
    achieved = alg_bytes / (scan_ms * 1e-3) / 1e9
    strategy = N.STRATEGIES.get(stats[-1].strategy)
    kname = "part_scan_fast_kernel<3, 1>" if strategy in ("part_hash", "part_direct") else "scan_agg_kernel"
    traffic = None
    tfile = ROOT / "profiles" / "traffic_filter_groupby_zipf.json"
    if tfile.exists():
        rec = json.loads(tfile.read_text())
        if rec.get("kernel") == kname and rec.get("rows") == args.rows:
            traffic = rec.get("hbm_bytes_per_launch")
    if args.copy_calibration:
        copy_calibration(1 << 30)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        n = min(args.cpu_sample_rows, args.rows)
        sample = {k: v[:n] for k, v in cols.items()}
        d2 = desc
        t = O.Table(sample)
        t1 = time.perf_counter()
        O.run(d2, {"t": t})
        cpu_s = time.perf_counter() - t1
        cpu = {"value": n / cpu_s, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"{'all' if n == args.rows else 'first'} {n} rows of the same table, same query: C "
                         f"restatement of the reference CPU path (per-page index-list filter + sorted "
                         f"intersection + hash aggregate), one thread, {cpu_s:.2f} s"}

    out = {
        "metric": "rows/s, single-query drop-in filter->group-by (100M rows, Zipf keys); HBM % of peak",
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64/f64",
        "data": "synthetic (deterministic splitmix64 generator, seed 0)",
        "config": {"workload": "filter_groupby_zipf", "rows": args.rows, "groups": int(stats[-1].groups),
                   "selected": int(stats[-1].rows_selected), "strategy": strategy,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "kernel": {"name": kname, "avg_ms": scan_ms, "algorithmic_bytes": alg_bytes,
                   "query_device_ms": dev_ms},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s, "upload_s": upload_s},
    }
    print(json.dumps(out))
    plan.close()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
