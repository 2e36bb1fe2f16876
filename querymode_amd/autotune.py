"""Empirical tuner for the device planner's knobs (qm_tune_set / qm_tune_get).

The library sizes hash tables, partitions and LDS replicas from the column
statistics; the knobs cover what a model cannot decide well: the partition
shape of the batched path over an unsorted key (keys per final partition, the
split between its two scatter passes, scatter workgroups per CU), the
shared-interval pass's wave rounds, how finely a batch over a clustered key
splits the table into chunks, the placement window, the run length from which
the run-sorted kernel beats the query-parallel one, the scatter shape and
pass-3 table fill (table size) of the partitioned group-by, and the top-K scan
width. Only the knobs of the kernels a workload runs are searched (from the
strategies its queries take). The tuner times a workload (one batch through
``qm_batch_exec``, or one query through ``qm_exec``) under candidate knob
values by coordinate descent and keeps the fastest setting whose results are
identical to the untuned run's (per query: result rows and selected rows).

Chosen settings can be kept in a JSON file keyed by the workload (device,
table rows, query count, a hash of the plans), so a serving process tunes a
recurring workload once.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import statistics
import time
from dataclasses import dataclass, field
from pathlib import Path

from . import _native as N
from .plan import QueryDescriptor, encode_plan

# candidate values per knob (the first pass starts from the library defaults),
# by the strategy a query of the workload takes: only the knobs of the
# kernels the workload runs are searched
BATCH_SPACES: dict[str, dict[str, list[float]]] = {
    # partitioned runs (a key in any row order): partition shape and scatter grid
    "part_runs": {
        "pr_scatter_bpc": [1, 2, 3, 4],
        "pr_b2": [0, 7, 8, 9, 10],
        "pr_max_s": [6, 5],
    },
    # shared-interval pass (small keys): workgroup rounds per resident slot
    "shared_interval": {"shared_waves": [2, 4, 8, 16]},
    # run-sorted kernels (a clustered key): chunking, staging blocks, placement window, kernel choice
    "sorted_runs": {
        "runs_chunks_per_cu": [8, 16, 32, 64, 128],
        "runs_qblock": [0, 16, 32, 64],
        "runs_win": [1024, 2048],
        "runs_min_len": [4, 8, 16, 32],
    },
}
# the union, for callers that name no workload
BATCH_SPACE: dict[str, list[float]] = {k: v for sp in BATCH_SPACES.values() for k, v in sp.items()}
QUERY_SPACES: dict[str, dict[str, list[float]]] = {
    # partitioned group-by: scatter regions, pass-3 table fill (table size), heavy hitters
    "part_hash": {"scatter_rpb": [0, 1, 2, 4], "part_load": [0.35, 0.5, 0.7], "hot": [0, 1]},
    "part_direct": {"scatter_rpb": [0, 1, 2, 4], "part_load": [0.35, 0.5, 0.7], "hot": [0, 1]},
    # top-K: scan workgroups per CU (each keeps k candidates)
    "topk": {"topk_bpc": [1, 2, 4]},
}
QUERY_SPACE: dict[str, list[float]] = QUERY_SPACES["part_hash"]


def space_for(strategies, spaces: dict[str, dict[str, list[float]]]) -> dict[str, list[float]]:
    """The knobs of the kernels behind the given strategies (in the order the
    strategies are first seen)."""
    out: dict[str, list[float]] = {}
    for s in strategies:
        for k, v in spaces.get(s, {}).items():
            out.setdefault(k, v)
    return out


@dataclass
class TuneResult:
    config: dict[str, float]
    ms: float
    default_ms: float
    trials: list[tuple[dict[str, float], float]] = field(default_factory=list)
    seconds: float = 0.0

    def as_dict(self) -> dict:
        start = self.trials[0][0] if self.trials else {}
        tried = []
        for cfg, ms in self.trials[1:]:
            diff = {k: v for k, v in cfg.items() if start.get(k) != v}
            tried.append({"set": diff, "ms": ms})
        return {"config": self.config, "ms": self.ms, "default_ms": self.default_ms, "trials": len(self.trials),
                "seconds": self.seconds, "tried": tried}


def workload_key(dev, table, descs: list[QueryDescriptor]) -> str:
    h = hashlib.sha256()
    for d in descs:
        h.update(encode_plan(d))
    info = dev.info()
    return f"{info['arch']}|rows={table.rows}|n={len(descs)}|{h.hexdigest()[:16]}"


def coordinate_descent(measure, start: dict[str, float], space: dict[str, list[float]], passes: int = 2):
    """Minimise measure(config) -> ms over one knob at a time; measure returns
    None for a configuration that must not be used (e.g. different results)."""
    best = dict(start)
    best_ms = measure(best)
    trials = [(dict(best), best_ms)]
    if best_ms is None:
        raise RuntimeError("the starting configuration failed")
    for _ in range(passes):
        changed = False
        for knob, values in space.items():
            for v in values:
                if best.get(knob) == v:
                    continue
                cand = dict(best)
                cand[knob] = v
                ms = measure(cand)
                trials.append((dict(cand), ms))
                if ms is not None and ms < best_ms * 0.99:  # ignore noise-level gains
                    best, best_ms, changed = cand, ms, True
        if not changed:
            break
    return best, best_ms, trials


class Autotuner:
    """Tunes the knobs of one Device (querymode_amd.client.Device)."""

    def __init__(self, dev, repeats: int = 3, warmup: int = 1):
        self.dev = dev
        self.repeats = repeats
        self.warmup = warmup

    # ---- knob access ----
    def current(self, names) -> dict[str, float]:
        return {k: self.dev.tune_get(k) for k in names}

    def apply(self, config: dict[str, float]) -> None:
        for k, v in config.items():
            self.dev.tune_set(k, v)

    # ---- workloads ----
    def _run_batch(self, table, plans, strategies=None):
        lib = self.dev.lib
        arr = (C.c_void_p * len(plans))(*[p.handle for p in plans])
        outs = (C.c_void_p * len(plans))()
        N.check(lib.qm_batch_exec(self.dev.handle, arr, len(plans), table.handle, outs))
        dev_ms, sig = 0.0, []
        for i in range(len(plans)):
            st = N.Stats()
            N.check(lib.qm_result_stats(C.c_void_p(outs[i]), C.byref(st)))
            dev_ms = max(dev_ms, st.device_ms)
            sig.append((st.groups, st.rows_selected))
            if strategies is not None:
                strategies.append(N.STRATEGIES.get(st.strategy, str(st.strategy)))
            lib.qm_result_close(C.c_void_p(outs[i]))
        return dev_ms, sig

    def _run_query(self, table, plans, right=None):
        lib = self.dev.lib
        rh = C.c_void_p()
        N.check(lib.qm_exec(plans[0].handle, table.handle, right.handle if right else None, C.byref(rh)))
        st = N.Stats()
        N.check(lib.qm_result_stats(rh, C.byref(st)))
        lib.qm_result_close(rh)
        return st.device_ms, [(st.groups, st.rows_selected)]

    def _tune(self, run, space, passes) -> TuneResult:
        start = self.current(space)
        ref_sig = None

        def measure(cfg):
            nonlocal ref_sig
            self.apply(cfg)
            for _ in range(self.warmup):
                run()
            times, sig = [], None
            for _ in range(self.repeats):
                t, sig = run()
                times.append(t)
            if ref_sig is None:
                ref_sig = sig
            elif sig != ref_sig:
                return None  # a knob must never change results
            return statistics.median(times)

        t0 = time.perf_counter()
        best, best_ms, trials = coordinate_descent(measure, start, space, passes)
        self.apply(best)
        return TuneResult(best, best_ms, trials[0][1], trials, time.perf_counter() - t0)

    def tune_batch(self, table, descs: list[QueryDescriptor], space=None, passes: int = 2) -> TuneResult:
        """Tune the knobs of the kernels the batch runs (from the strategies its
        queries take under the current knobs), or those of `space`."""
        from .client import _Plan
        plans = [_Plan(self.dev, d) for d in descs]
        try:
            if space is None:
                seen: list[str] = []
                self._run_batch(table, plans, seen)
                space = space_for(dict.fromkeys(seen), BATCH_SPACES)
            if not space:
                t, _ = self._run_batch(table, plans)
                return TuneResult({}, t, t)
            return self._tune(lambda: self._run_batch(table, plans), space, passes)
        finally:
            for p in plans:
                p.close()

    def tune_query(self, table, desc: QueryDescriptor, right=None, space=None, passes: int = 1) -> TuneResult:
        """Tune the knobs of the kernels the query runs (a top-K, or the
        strategy its group-by takes), or those of `space`."""
        from .client import _Plan
        plans = [_Plan(self.dev, desc)]
        try:
            if space is None:
                if desc.sortColumn and desc.limit is not None and not desc.aggregates:
                    space = QUERY_SPACES["topk"]
                else:
                    lib = self.dev.lib
                    rh = C.c_void_p()
                    N.check(lib.qm_exec(plans[0].handle, table.handle, right.handle if right else None, C.byref(rh)))
                    st = N.Stats()
                    N.check(lib.qm_result_stats(rh, C.byref(st)))
                    lib.qm_result_close(rh)
                    space = space_for([N.STRATEGIES.get(st.strategy, "")], QUERY_SPACES)
            if not space:
                t, _ = self._run_query(table, plans, right)
                return TuneResult({}, t, t)
            return self._tune(lambda: self._run_query(table, plans, right), space, passes)
        finally:
            plans[0].close()


# ---- cache of chosen settings ----
def load_cache(path: str | Path) -> dict:
    p = Path(path)
    return json.loads(p.read_text()) if p.exists() else {}


def save_choice(path: str | Path, key: str, config: dict[str, float]) -> None:
    p = Path(path)
    data = load_cache(p)
    data[key] = config
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


__all__ = ["Autotuner", "TuneResult", "coordinate_descent", "workload_key", "load_cache", "save_choice",
           "BATCH_SPACE", "BATCH_SPACES", "QUERY_SPACE", "QUERY_SPACES", "space_for"]
