"""Joins in key-range passes (the device counterpart of the reference's
partitioned join for build sides over the memory budget,
src/operators.ts:2466-2637): every pass builds the index of one range of the
build key and probes the whole probe side, aggregating into the same group
tables (fused join + aggregate) or appending its pairs (materialised join).

* Forced pass counts (the join_parts knob) on the three index kinds (DIRECT
  with unique keys, DIRECT with duplicate keys through row lists, open
  addressing for sparse keys): results equal the one-pass join and the oracle.
* A context whose HBM budget cannot hold the one-pass index: the planner picks
  several passes by itself (stats.join_parts > 1), the results are right, and
  the same query forced to one pass fails with MEMORY_EXCEEDED."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd import QueryMode, datagen
from querymode_amd import _native as N
from querymode_amd.client import _Plan
from querymode_amd.errors import QueryModeError
from querymode_amd.plan import AggregateOp, FilterOp, JoinDescriptor, QueryDescriptor

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu


def _agg_join(pt="jp_probe", bt="jp_build", rfilter=None):
    right = QueryDescriptor(bt, filters=[rfilter] if rfilter else [])
    return QueryDescriptor(pt, join=JoinDescriptor(right, "fk", "id"), groupBy=["region"],
                           filters=[FilterOp("amount", "gt", 5.0)],
                           computedColumns=[("aw", "amount", "weight")],
                           aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*"),
                                       AggregateOp("min", "weight"), AggregateOp("max", "amount")])


def _tables(kind: str, nb: int, npb: int, seed: int):
    rng = np.random.default_rng(seed)
    b = datagen.join_build(n=nb, seed=seed)
    p = datagen.join_probe(n=npb, nids=nb, seed=seed)
    if kind == "dup":      # every key three times (row lists per key)
        b = {k: np.concatenate([v, v, v]) for k, v in b.items()}
    elif kind == "sparse":  # ids spread over the int64 range: open addressing
        mult = np.int64(1_000_003_007)
        b["id"] = b["id"].astype(np.int64) * mult - (1 << 50)
        p["fk"] = p["fk"].astype(np.int64) * mult - (1 << 50)
    return b, p


def _rows_key(r):
    return tuple((0, v) if v is not None else (1, 0) for v in r.values())


@pytest.mark.parametrize("kind", ["unique", "dup", "sparse"])
def test_forced_passes_equal_one_pass(qm, kind):
    b, p = _tables(kind, 50_000, 400_000, seed=41)
    qm.register_table("jp_build", b)
    qm.register_table("jp_probe", p)
    tabs = {"jp_build": O.Table(b), "jp_probe": O.Table(p)}
    d_agg = _agg_join()
    d_rows = QueryDescriptor("jp_probe", join=JoinDescriptor(QueryDescriptor("jp_build"), "fk", "id"),
                             filters=[FilterOp("amount", "gt", 60.0)], projections=["fk", "amount", "region", "weight"])
    want_agg = O.run(d_agg, tabs)
    want_rows = sorted(O.run(d_rows, tabs), key=_rows_key)
    try:
        for parts in (0, 2, 3, 16, 257):
            print(f"join parts: {kind} index, {parts} passes", flush=True)
            qm.device.tune_set("join_parts", parts)
            r = qm.executor.execute(d_agg)
            assert r.stats["join_parts"] == max(1, parts), r.stats["join_parts"]
            assert_groups_equal(r.rows, want_agg, ["region"], float_cols=("sum_aw",), ctx=f"agg {kind} {parts}")
            got = qm.executor.execute(d_rows)
            assert got.stats["join_parts"] == max(1, parts)
            assert_rows_equal(sorted(got.rows, key=_rows_key), want_rows, ctx=f"rows {kind} {parts}")
            # a build-side filter combines with the range filter of each pass
            d2 = _agg_join(rfilter=FilterOp("region", "lt", 20))
            assert_groups_equal(qm.executor.execute(d2).rows, O.run(d2, tabs), ["region"], float_cols=("sum_aw",),
                                ctx=f"filtered {kind} {parts}")
    finally:
        qm.device.tune_set("join_parts", 0)


def test_limit_stops_after_enough_pairs(qm):
    b, p = _tables("unique", 20_000, 100_000, seed=42)
    qm.register_table("jp_build", b)
    qm.register_table("jp_probe", p)
    d = QueryDescriptor("jp_probe", join=JoinDescriptor(QueryDescriptor("jp_build"), "fk", "id"),
                        projections=["fk", "region"], limit=10)
    try:
        qm.device.tune_set("join_parts", 8)
        got = qm.executor.execute(d)
        assert got.rowCount == 10
        ids = set(b["id"].tolist())
        assert all(r["fk"] in ids for r in got.rows)
    finally:
        qm.device.tune_set("join_parts", 0)


def test_small_budget_picks_passes():
    b, p = _tables("unique", 2_000_000, 4_000_000, seed=43)
    d = _agg_join(pt="probe", bt="build")
    big = QueryMode.gpu(0)
    try:
        big.register_table("build", b)
        big.register_table("probe", p)
        used = big.device.info()["used"]
        big.device.tune_set("join_parts", 1)
        plan = _Plan(big.device, d)
        try:
            nb = C.c_uint64()
            N.check(big.device.lib.qm_plan_device_bytes(plan.handle, big.device.tables["probe"].handle,
                                                        big.device.tables["build"].handle, C.byref(nb)))
            one_pass = int(nb.value)
        finally:
            plan.close()
        want = big.executor.execute(d).rows
    finally:
        big.close()
    budget = used + int(one_pass * 0.6)  # one pass cannot fit, four can
    print(f"join parts: budget {budget >> 20} MiB, tables {used >> 20} MiB, one-pass index {one_pass >> 20} MiB",
          flush=True)
    small = QueryMode.gpu(0, hbm_budget_bytes=budget)
    try:
        small.register_table("build", b)
        small.register_table("probe", p)
        r = small.executor.execute(d)
        assert r.stats["join_parts"] >= 2, r.stats["join_parts"]
        assert_groups_equal(r.rows, want, ["region"], float_cols=("sum_aw",), ctx="small budget")
        assert_groups_equal(r.rows, O.run(d, {"build": O.Table(b), "probe": O.Table(p)}), ["region"],
                            float_cols=("sum_aw",), ctx="small budget vs oracle")
        small.device.tune_set("join_parts", 1)
        with pytest.raises(QueryModeError) as e:
            small.executor.execute(d)
        assert e.value.code == "MEMORY_EXCEEDED"
    finally:
        small.close()
