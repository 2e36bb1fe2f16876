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


JOIN_KERNEL = "scan_agg_kernel<true> (fused join probe)"


def _traffic_rec(kname: str, rows: int, fname: str):
    tfile = ROOT / "profiles" / fname
    if tfile.exists():
        rec = json.loads(tfile.read_text())
        if rec.get("kernel") == kname and rec.get("rows") == rows:
            return rec
    return None


def _traffic(kname: str, rows: int, fname: str):
    """HBM bytes per launch of the kernel from the committed FETCH_SIZE / WRITE_SIZE summary (None if the
    summary is of another kernel or table size)."""
    rec = _traffic_rec(kname, rows, fname)
    return rec.get("hbm_bytes_per_launch") if rec else None


def _traffic_build(kname: str, rows: int, fname: str):
    """The build the committed traffic summary was measured on."""
    rec = _traffic_rec(kname, rows, fname)
    return rec.get("build") if rec else None


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
    kname = "part_scan_fast_kernel<3, 1>" if strategy in ("part_hash", "part_direct") else "scan_agg_kernel<false>"
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
                     "traffic": _traffic(kname, rows, "traffic_filter_groupby_zipf.json"),
                     "traffic_build": _traffic_build(kname, rows, "traffic_filter_groupby_zipf.json")},
        "kernel": {"name": kname, "avg_ms": scan_ms, "algorithmic_bytes": alg_bytes,
                   "query_device_ms": float(np.mean([s.device_ms for s in stats]))},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s, "upload_s": upload_s},
    }
    plan.close()
    dev.close()
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_batch_worker(job):
    """One CPU worker: the batch table's rows [row0, row0 + n) regenerated
    (same bytes as the device table), every query of the mix over them through
    the C restatement of the reference CPU path. Returns its seconds."""
    n, row0, nq, k2_order = job
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as O
    from querymode_amd import datagen
    cols = datagen.batch_table(n=n, seed=0, row0=row0) if not k2_order else None
    descs = datagen.batch_queries(nq, seed=0)
    t = O.Table(cols)
    t1 = time.perf_counter()
    for d in descs:
        O.run(d, {"batch": t})
    return time.perf_counter() - t1


def cpu_batch_baseline(nq: int, sample_rows: int, workers: int) -> dict:
    """The reference CPU path (C restatement, `oracle/`) over `sample_rows`
    rows per worker, at one thread and at `workers` processes on disjoint row
    ranges (partial results are not merged: the merge of per-worker groups is
    left out of the multi-core figure)."""
    import multiprocessing as mp
    one = _cpu_batch_worker((sample_rows, 0, nq, False))
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        t0 = time.perf_counter()
        pool.map(_cpu_batch_worker, [(sample_rows, i * sample_rows, nq, False) for i in range(workers)])
        wall = time.perf_counter() - t0
    return {"value": nq * sample_rows / one, "unit": "rows/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"all {nq} queries over rows [0, {sample_rows}) of the same table: C restatement of the "
                      f"reference CPU path (per-page index-list filter + hash aggregate per query), one thread, "
                      f"{one:.2f} s",
            "multi_core": {"value": nq * sample_rows * workers / wall, "unit": "rows/s", "cores": workers,
                           "sample": f"{workers} processes, each all {nq} queries over its own {sample_rows} rows "
                                     f"of the table ({wall:.2f} s wall; per-worker partial groups not merged)"}}


def bench_batch(args) -> dict:
    import numpy as np
    from querymode_amd import _native as N
    from querymode_amd import datagen
    from querymode_amd.client import Device, _Plan

    rows = args.rows or 500_000_000
    nq = args.queries
    t0 = time.time()
    cols = datagen.batch_table(n=rows, seed=0, k2_order=args.k2_order)
    gen_s = time.time() - t0
    descs = datagen.batch_queries(nq, seed=0)
    dev = Device(0)
    lib = dev.lib
    t0 = time.time()
    table = dev.register("batch", cols)
    upload_s = time.time() - t0
    tuned = None
    if not args.no_autotune:
        from querymode_amd.autotune import Autotuner
        tuned = Autotuner(dev).tune_batch(table, descs).as_dict()
    plans = [_Plan(dev, d) for d in descs]
    arr = (C.c_void_p * nq)(*[p.handle for p in plans])
    k2 = [i for i, d in enumerate(descs) if d.groupBy == ["k2"]]

    def batch_step():
        outs = (C.c_void_p * nq)()
        N.check(lib.qm_batch_exec(dev.handle, arr, nq, table.handle, outs))
        per = []
        for i in range(nq):
            s2 = N.Stats()
            lib.qm_result_stats(C.c_void_p(outs[i]), C.byref(s2))
            per.append(s2)
            lib.qm_result_close(C.c_void_p(outs[i]))
        return per

    for _ in range(args.warmup):
        batch_step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(batch_step())
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    batched_rows_s = nq * rows / (wall / args.steps)
    last = stats[-1]
    groups = int(sum(s.groups for s in last))
    dev_ms = float(np.mean([st[0].device_ms for st in stats]))
    batch_bytes = rows * sum(np.dtype(c.dtype).itemsize for c in cols.values())  # every column once per pass
    strat = N.STRATEGIES.get(last[k2[0]].strategy) if k2 else None
    kernels = {}
    if strat == "part_runs":
        # the k2 queries' launch: row scatter (reads k2, v1, v2 once), record
        # scatter, group count, write pass (records in, result columns out)
        names = ["pr_scatter_kernel<0, 4>", "pr_scatter_kernel<1, 4>", "pr_count_kernel", "pr_runs_kernel<1, 2, 4>"]
        pm = [float(np.mean([st[k2[0]].pass_ms[j] for st in stats])) for j in range(4)]
        records = int(last[k2[0]].pass_records)
        k2_groups = int(sum(last[i].groups for i in k2))
        rec_bytes = 16  # key code / bucket word + v1 (4 B) + v2 (8 B)
        out_bytes = 8 + 8 + 8 + 4 + 8  # k2, SUM(v1), COUNT(*), MIN(v1), MAX(v2)
        alg = [rows * 20, records * rec_bytes, records * rec_bytes, records * rec_bytes + k2_groups * out_bytes]
        for j in range(4):
            kernels[names[j]] = {"avg_ms": pm[j], "algorithmic_bytes": alg[j],
                                 "achieved_gbs": alg[j] / (pm[j] * 1e-3) / 1e9}
        dom = max(range(4), key=lambda j: pm[j])
        kname, scan_ms, alg_bytes = names[dom], pm[dom], alg[dom]
    else:
        kname = "runs_kernel<1, 1, 2, 2>"
        scan_ms = float(np.mean([st[0].scan_ms for st in stats]))
        alg_bytes = rows * sum(np.dtype(cols[c].dtype).itemsize for c in ("k2", "v2", "v1"))
        records = None
    achieved = alg_bytes / (scan_ms * 1e-3) / 1e9
    # the same batch captured into a HIP graph (qm_batch_prepare) and replayed:
    # every pass runs again with one host launch and no host planning
    replay = None
    if not args.no_replay:
        from querymode_amd.client import PreparedBatch
        t0 = time.perf_counter()
        ph = C.c_void_p()
        N.check(lib.qm_batch_prepare(dev.handle, arr, nq, table.handle, C.byref(ph)))
        prep_s = time.perf_counter() - t0
        prep = PreparedBatch(dev, ph, table, nq)

        def replay_step():
            outs = (C.c_void_p * nq)()
            prep.run_handles(outs)
            s2 = N.Stats()
            lib.qm_result_stats(C.c_void_p(outs[0]), C.byref(s2))
            for i in range(nq):
                lib.qm_result_close(C.c_void_p(outs[i]))
            return s2.device_ms

        for _ in range(args.warmup):
            replay_step()
        t0 = time.perf_counter()
        rdev = [replay_step() for _ in range(args.steps)]
        rwall = (time.perf_counter() - t0) / args.steps
        info = prep.info()
        prep.close()
        replay = {"ms_per_step": rwall * 1e3, "device_ms": float(np.mean(rdev)), "rows_per_s": nq * rows / rwall,
                  "vs_batch_exec": ms_step / (rwall * 1e3), "prepare_s": prep_s,
                  "graph_nodes": info["graph_nodes"], "device_bytes": info["device_bytes"]}
        print(f"replay: {rwall * 1e3:.2f} ms/step (qm_batch_exec {ms_step:.2f}), {info['graph_nodes']} graph nodes",
              file=sys.stderr, flush=True)
    read_ceiling = None
    if not args.no_read_ceiling:
        g = C.c_double()
        N.check(lib.qm_measure_read_bw(dev.handle, 8 << 30, 10, C.byref(g)))
        read_ceiling = g.value
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
    served = None
    if args.serve_seconds > 0:
        served = serve_mix(args, dev, cols, rows, descs, [args.qps, args.qps_high])
    if args.copy_calibration:
        copy_calibration(1 << 30)
    for p in plans:
        p.close()
    dev.close()
    del cols
    cpu = None
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or min(16, os.cpu_count() or 1)
        cpu = cpu_batch_baseline(nq, min(args.cpu_sample_rows or 1_000_000, rows), workers)
    tfile = "traffic_batch_part_runs.json" if strat == "part_runs" else "traffic_batch_mix.json"
    out = {
        "metric": METRIC, "value": batched_rows_s, "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f64",
        "data": "synthetic (deterministic splitmix64 generator, seed 0)" + ("; table stored in k2 order" if args.k2_order else "; rows in generated order"),
        "config": {"workload": f"batched {nq}-query mix over a resident {rows // 1_000_000}M-row table "
                               "(range filter on v2, GROUP BY k1 or k2, SUM/COUNT/MIN/MAX)",
                   "rows": rows, "queries": nq, "groups_per_batch": groups,
                   "k2_order": bool(args.k2_order), "parallelism": "single GPU"},
        "batched_rows_per_s": batched_rows_s,
        "per_query_rows_per_s": per_query_rows_s,
        "batched_vs_per_query": batched_rows_s / per_query_rows_s,
        "per_query_ms_total": pq_wall * 1e3,
        "per_query_strategies": strategies,
        "batch_strategy_k2": strat,
        "served": served,
        "graph_replay": replay,
        "autotune": tuned,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "frac_of_measured_read": (achieved / read_ceiling) if read_ceiling else None,
                     "traffic": _traffic(kname, rows, tfile), "traffic_build": _traffic_build(kname, rows, tfile)},
        "kernel": {"name": kname, "avg_ms": scan_ms, "algorithmic_bytes": alg_bytes, "records": records},
        "kernels": kernels,
        "read_ceiling_gbs": read_ceiling,
        # the whole fused batch: every referenced column read once per pass
        "batch_hbm": {"device_ms": dev_ms, "algorithmic_bytes": batch_bytes,
                      "achieved": batch_bytes / (dev_ms * 1e-3) / 1e9,
                      "frac": batch_bytes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s, "upload_s": upload_s},
    }
    return out


def serve_mix(args, dev, cols, rows, descs, rates) -> dict:
    """Poisson arrivals at fixed rates through the continuous-batching server:
    40% group-by queries of the batched mix, 30% the 200M x 10M join +
    aggregate, 30% sort(v2 desc).limit(k), k in {10, 100, 1000}. One phase of
    `serve_seconds` per rate; the first rate's figures are the top-level ones."""
    import random
    import threading

    from querymode_amd import datagen
    from querymode_amd.client import QueryMode
    from querymode_amd.plan import AggregateOp, JoinDescriptor, QueryDescriptor
    from querymode_amd.serve import LANES, Server, percentile

    qm = QueryMode(dev)
    nbuild, nprobe = 10_000_000, 200_000_000
    qm.register_table("build", datagen.join_build(n=nbuild, seed=0))
    qm.register_table("probe", datagen.join_probe(n=nprobe, nids=nbuild, seed=0))
    join = QueryDescriptor("probe", join=JoinDescriptor(QueryDescriptor("build"), "fk", "id"), groupBy=["region"],
                           computedColumns=[("aw", "amount", "weight")],
                           aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*")])
    # warm the join index, top-K and group tables once
    for d in (join, descs[0], QueryDescriptor("batch", sortColumn="v2", sortDirection="desc", limit=10)):
        qm.executor.execute(d)

    def phase(qps: float, seed: int) -> dict:
        rng = random.Random(seed)
        mix, nrows = [], []
        t = 0.0
        while True:
            t += rng.expovariate(qps)
            if t > args.serve_seconds:
                break
            u = rng.random()
            if u < 0.4:
                mix.append((t, descs[rng.randrange(len(descs))]))
                nrows.append(rows)
            elif u < 0.7:
                mix.append((t, join))
                nrows.append(nprobe + nbuild)
            else:
                mix.append((t, QueryDescriptor("batch", sortColumn="v2", sortDirection="desc",
                                               limit=rng.choice([10, 100, 1000]))))
                nrows.append(rows)
        srv = Server(qm)
        reqs = [None] * len(mix)
        t0 = time.perf_counter()

        def producer():
            for i, (at, d) in enumerate(mix):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                reqs[i] = srv.submit(d)

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        for r in reqs:
            r.wait(600)
        elapsed = time.perf_counter() - t0
        srv.close()
        lat = [r.latency_ms for r in reqs]
        kinds = {"group-by": [], "join": [], "top-k": []}
        for r in reqs:
            k = "join" if r.desc.join is not None else ("group-by" if r.desc.aggregates else "top-k")
            kinds[k].append(r.latency_ms)
        gb = [r for r in reqs if r.fused]
        print(f"served {qps:.0f} QPS: {len(reqs)} requests, p50 {percentile(lat, 50):.1f} ms, "
              f"p99 {percentile(lat, 99):.1f} ms, {srv.launches} launches, overlap {srv.overlap_share:.2f}",
              file=sys.stderr, flush=True)
        return {"qps": qps, "seconds": args.serve_seconds, "requests": len(reqs),
                "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                "rows_per_s": sum(nrows) / elapsed, "achieved_qps": len(reqs) / elapsed,
                "launches": srv.launches, "mean_requests_per_launch": len(reqs) / max(1, srv.launches),
                "mean_group_by_batch": (sum(r.batch_size for r in gb) / len(gb)) if gb else None,
                "lanes": LANES, "max_concurrent_launches": srv.max_overlap,
                "overlap_share_of_busy_time": srv.overlap_share,
                "p50_ms_by_kind": {k: (percentile(v, 50) if v else None) for k, v in kinds.items()},
                "p99_ms_by_kind": {k: (percentile(v, 99) if v else None) for k, v in kinds.items()}}

    phases = [phase(q, 11 + i) for i, q in enumerate(rates)]
    out = dict(phases[0])
    out["mix"] = "40% group-by (batched mix), 30% join 200M x 10M + aggregate, 30% top-K v2 desc k in {10,100,1000}"
    out["phases"] = phases
    return out


def bench_join(args) -> dict:
    import numpy as np
    from querymode_amd import _native as N
    from querymode_amd import datagen
    from querymode_amd.client import Device, _Plan
    from querymode_amd.plan import AggregateOp, JoinDescriptor, QueryDescriptor

    nprobe = args.rows or 200_000_000
    nbuild = 10_000_000
    t0 = time.time()
    build = datagen.join_build(n=nbuild, seed=0)
    probe = datagen.join_probe(n=nprobe, nids=nbuild, seed=0)
    gen_s = time.time() - t0
    desc = QueryDescriptor("probe", join=JoinDescriptor(QueryDescriptor("build"), "fk", "id"), groupBy=["region"],
                           computedColumns=[("aw", "amount", "weight")],
                           aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*")])
    dev = Device(0)
    t0 = time.time()
    tb = dev.register("build", build)
    tp = dev.register("probe", probe)
    upload_s = time.time() - t0
    plan = _Plan(dev, desc)

    def step():
        rh = C.c_void_p()
        N.check(dev.lib.qm_exec(plan.handle, tp.handle, tb.handle, C.byref(rh)))
        st = N.Stats()
        N.check(dev.lib.qm_result_stats(rh, C.byref(st)))
        dev.lib.qm_result_close(rh)
        return st

    for _ in range(args.warmup):
        step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    wall = time.perf_counter() - t0
    rows = nprobe + nbuild
    scan_ms = float(np.mean([s.scan_ms for s in stats]))
    alg = nprobe * 16 + nbuild * 16  # fk + amount; id + region + weight (+ pad) read once
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        n = min(args.cpu_sample_rows or 10_000_000, nprobe)
        t1 = time.perf_counter()
        O.run(desc, {"build": O.Table(build), "probe": O.Table({k: v[:n] for k, v in probe.items()})})
        cpu_s = time.perf_counter() - t1
        cpu = {"value": (n + nbuild) / cpu_s, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the whole 10M-row build side and the first {n} probe rows, same query: C restatement "
                         f"of the reference hash join (build right, probe left) + aggregate, one thread, {cpu_s:.2f} s"}
    out = {
        "metric": METRIC, "value": rows / (wall / args.steps), "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f64",
        "data": "synthetic (deterministic splitmix64 generator, seed 0)",
        "config": {"workload": "hash join 200M probe x 10M build -> GROUP BY region SUM(amount*weight), COUNT(*)",
                   "probe_rows": nprobe, "build_rows": nbuild, "selected": int(stats[-1].rows_selected),
                   "strategy": N.STRATEGIES.get(stats[-1].strategy), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": alg / (scan_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (scan_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": _traffic(JOIN_KERNEL, rows, "traffic_join_probe.json"),
                     "traffic_build": _traffic_build(JOIN_KERNEL, rows, "traffic_join_probe.json")},
        "kernel": {"name": JOIN_KERNEL, "avg_ms": scan_ms, "algorithmic_bytes": alg,
                   "query_device_ms": float(np.mean([s.device_ms for s in stats]))},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s, "upload_s": upload_s},
    }
    plan.close()
    dev.close()
    return out


def bench_topk(args) -> dict:
    import numpy as np
    from querymode_amd import _native as N
    from querymode_amd import datagen
    from querymode_amd.client import Device, _Plan
    from querymode_amd.plan import QueryDescriptor

    rows = args.rows or 500_000_000
    t0 = time.time()
    cols = datagen.batch_table(n=rows, seed=0, k2_order=True)
    gen_s = time.time() - t0
    dev = Device(0)
    table = dev.register("batch", cols)
    per_k = {}
    for k in (10, 100, 1000):
        plan = _Plan(dev, QueryDescriptor("batch", sortColumn="v2", sortDirection="desc", limit=k))
        for _ in range(args.warmup):
            _exec_stats(dev.lib, N, plan, table)
        t0 = time.perf_counter()
        st = [_exec_stats(dev.lib, N, plan, table) for _ in range(args.steps)]
        wall = time.perf_counter() - t0
        per_k[k] = {"ms": wall / args.steps * 1e3, "scan_ms": float(np.mean([s.scan_ms for s in st])),
                    "device_ms": float(np.mean([s.device_ms for s in st]))}
        plan.close()
    alg = rows * 8
    scan_ms = per_k[100]["scan_ms"]
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        n = min(args.cpu_sample_rows or 100_000_000, rows)
        t = O.Table({"v2": cols["v2"][:n]})
        t1 = time.perf_counter()
        O.topk_heap_rows(t, "v2", True, 100)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": n / cpu_s, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {n} rows of v2, k = 100: C restatement of the reference binary-heap TopK, one "
                         f"thread, {cpu_s:.2f} s"}
    out = {
        "metric": METRIC, "value": rows / (per_k[100]["ms"] * 1e-3), "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": per_k[100]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (deterministic splitmix64 generator, seed 0)",
        "config": {"workload": "top-K: sort(v2 desc).limit(k) over 500M rows (value at k = 100)", "rows": rows,
                   "per_k": per_k, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": alg / (scan_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (scan_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "kernel": {"name": "topk_scan_kernel", "avg_ms": scan_ms, "algorithmic_bytes": alg},
        "cpu_baseline": cpu,
        "setup": {"generate_s": gen_s},
    }
    dev.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["batch", "zipf", "join", "topk"], default="batch")
    ap.add_argument("--rows", type=int, default=0, help="table rows (default: the workload's)")
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--k2-order", action="store_true",
                    help="store the batch table clustered by k2 (a named variant; the default keeps generated row order)")
    ap.add_argument("--cpu-sample-rows", type=int, default=0)
    ap.add_argument("--qps", type=float, default=40.0, help="served phase: Poisson arrival rate")
    ap.add_argument("--qps-high", type=float, default=120.0, help="served phase: second rate, near saturation")
    ap.add_argument("--serve-seconds", type=float, default=6.0, help="served phase length (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the planner's default knobs")
    ap.add_argument("--no-replay", action="store_true", help="skip the prepared-batch (HIP graph) replay phase")
    ap.add_argument("--no-read-ceiling", action="store_true", help="skip the streaming-read ceiling probe")
    ap.add_argument("--cpu-workers", type=int, default=0, help="CPU baseline processes (default: min(16, nproc))")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, run one 1 GiB device-to-device copy (the known-size dispatch the "
                         "FETCH_SIZE / WRITE_SIZE counters are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: run it with --gpus 1 (this build has no multi-GPU path)",
              file=sys.stderr)
        return 2
    out = {"batch": bench_batch, "zipf": bench_zipf, "join": bench_join, "topk": bench_topk}[args.workload](args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
