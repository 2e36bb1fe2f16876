"""Device path vs the CPU oracle on the same seeded inputs (sizes the oracle
finishes in seconds). Every query runs through the C-ABI (libqmgpu.so)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd import datagen
from querymode_amd.plan import AggregateOp, BigInt, FilterOp, JoinDescriptor, QueryDescriptor

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(qm, desc, tables, nulls=None):
    nulls = nulls or {}
    for name, cols in tables.items():
        qm.register_table(name, cols, nulls.get(name))
    got = qm.executor.execute(desc)
    want = O.run(desc, {n: O.Table(c, nulls.get(n)) for n, c in tables.items()})
    return got, want


@pytest.mark.parametrize("lean", [1, 0])
def test_filter_groupby_small(qm, lean):
    # lean 1: the shared-interval pass (the default for a small key);
    # lean 0: the fused scan with a replicated LDS table
    _say(f"case filter_groupby_small n=1M lean={lean}")
    cols = datagen.filter_groupby_small(n=1_000_000, seed=0)
    d = QueryDescriptor("t", filters=[FilterOp("status", "eq", 1)], groupBy=["category"],
                        aggregates=[AggregateOp("sum", "amount", "total"), AggregateOp("count", "*")],
                        sortColumn="total", sortDirection="desc")
    qm.device.tune_set("shared_lean", lean)
    try:
        got, want = run_both(qm, d, {"t": cols})
    finally:
        qm.device.tune_set("shared_lean", 1)
    assert got.stats["strategy"] == ("shared_interval" if lean else "direct_lds")
    assert_rows_equal(got.rows, want, float_cols={"total"}, ctx="c1")


@pytest.mark.parametrize("skewed", [True, False])
def test_filter_groupby_zipf(qm, skewed):
    _say(f"case filter_groupby_zipf skewed={skewed} n=2M")
    n = 2_000_000
    cols = datagen.filter_groupby_zipf(n=n, seed=1)
    if not skewed:
        cols["key"] = datagen.uniform_int(1, 99, 0, n, -(2 ** 62), 2 ** 63, np.int64)
    d = QueryDescriptor("t", filters=[FilterOp("a", "lt", 300000), FilterOp("b", "gt", 90), FilterOp("flag", "eq", 1)],
                        groupBy=["key"],
                        aggregates=[AggregateOp("sum", "b"), AggregateOp("count", "*"), AggregateOp("min", "a"),
                                    AggregateOp("max", "a"), AggregateOp("avg", "b")])
    got, want = run_both(qm, d, {"t": cols})
    assert got.stats["strategy"] == "part_hash"
    assert_groups_equal(got.rows, want, ["key"], float_cols={"sum_b", "avg_b"}, ctx="c2")


def test_global_aggregate_and_empty(qm):
    _say("case global aggregate")
    cols = datagen.filter_groupby_zipf(n=300_000, seed=2)
    cols["key"] = cols["a"].astype(np.int64) * 1_000_003 - 7  # int64 sums stay inside int64
    for flt in ([FilterOp("a", "lt", 500000)], [FilterOp("a", "lt", -5)]):
        d = QueryDescriptor("t", filters=flt, aggregates=[
            AggregateOp("sum", "b"), AggregateOp("count", "*"), AggregateOp("min", "a"), AggregateOp("max", "b"),
            AggregateOp("avg", "a"), AggregateOp("sum", "key"), AggregateOp("avg", "key")])
        got, want = run_both(qm, d, {"t": cols})
        assert_rows_equal(got.rows, want, float_cols={"sum_b", "max_b", "avg_a", "avg_key"}, ctx=str(flt))


def test_join_aggregate(qm):
    _say("case join aggregate build=100K probe=2M")
    build = datagen.join_build(n=100_000, seed=3)
    probe = datagen.join_probe(n=2_000_000, nids=100_000, seed=3)
    d = QueryDescriptor("p", join=JoinDescriptor(QueryDescriptor("b"), "fk", "id"), computedColumns=[("aw", "amount", "weight")],
                        groupBy=["region"], aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*")])
    got, want = run_both(qm, d, {"p": probe, "b": build})
    assert_groups_equal(got.rows, want, ["region"], float_cols={"sum_aw"}, ctx="join-agg")


def test_join_materialised_order(qm):
    _say("case join materialised")
    rng = np.random.default_rng(7)
    left = {"k": rng.integers(0, 300, 5000).astype(np.int64), "x": np.arange(5000, dtype=np.int32)}
    right = {"k": rng.integers(0, 400, 700).astype(np.int64), "x": rng.random(700), "y": np.arange(700, dtype=np.int32)}
    d = QueryDescriptor("l", join=JoinDescriptor(QueryDescriptor("r", filters=[FilterOp("y", "gte", 10)]), "k", "k"),
                        filters=[FilterOp("x", "lt", 4000)])
    got, want = run_both(qm, d, {"l": left, "r": right})
    assert_rows_equal(got.rows, want, float_cols={"right_x"}, ctx="join rows")


@pytest.mark.parametrize("payload", [1, 0])
def test_join_aggregate_build_columns_with_nulls(qm, payload):
    """Fused probe over unique build keys whose group key and values come from
    the build side (int16 / int8 / f32 with NULLs) and the probe side, with
    the build columns gathered from the packed per-key payload (payload=1) or
    per column (payload=0); both equal the oracle (src/operators.ts:2673-2713,
    src/partial-agg.ts:62-80)."""
    _say(f"case join aggregate build columns with nulls payload={payload}")
    rng = np.random.default_rng(23)
    nb, npr = 50_000, 400_000
    build = {"id": rng.permutation(nb).astype(np.int64) * 3 - 1000,
             "g": rng.integers(-3, 4, nb).astype(np.int8),
             "h": rng.integers(-200, 200, nb).astype(np.int16),
             "w": rng.random(nb).astype(np.float32)}
    bnull = {"g": rng.random(nb) < 0.05, "h": rng.random(nb) < 0.05, "w": rng.random(nb) < 0.05}
    probe = {"fk": (rng.integers(0, nb + nb // 9, npr) * 3 - 1000).astype(np.int64),
             "pg": rng.integers(0, 3, npr).astype(np.int32),
             "amount": rng.normal(10, 3, npr)}
    pnull = {"amount": rng.random(npr) < 0.05}
    d = QueryDescriptor("p", join=JoinDescriptor(QueryDescriptor("b"), "fk", "id"),
                        computedColumns=[("aw", "amount", "w")], groupBy=["g", "pg"],
                        aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*"), AggregateOp("min", "h"),
                                    AggregateOp("max", "w"), AggregateOp("sum", "h"), AggregateOp("count", "w"),
                                    AggregateOp("avg", "amount")])
    qm.device.tune_set("join_payload", payload)
    try:
        got, want = run_both(qm, d, {"p": probe, "b": build}, {"p": pnull, "b": bnull})
    finally:
        qm.device.tune_set("join_payload", 1)
    assert_groups_equal(got.rows, want, ["g", "pg"], float_cols={"sum_aw", "max_w", "avg_amount"}, ctx="join nulls")


@pytest.mark.parametrize("sparse", [False, True])
def test_join_heavily_duplicated_build_keys(qm, sparse):
    """Build keys repeated 100K and 40K times (row ids spread over more than
    one 262144-value window) and ~200 times each: every key's build rows come
    out ascending, i.e. in the reference's insertion order
    (src/operators.ts:2349-2436, 2673-2713); DIRECT (dense) and HASH (sparse) index."""
    _say(f"case join heavy duplicates sparse={sparse}")
    rng = np.random.default_rng(17)
    nb = 300_000
    bk = rng.integers(0, 1000, nb).astype(np.int64)
    pos = rng.permutation(nb)
    bk[pos[:100_000]] = 7
    bk[pos[100_000:140_000]] = 9
    pk = rng.integers(0, 1000, 2000).astype(np.int64)
    pk[[5, 700, 1999]] = 7
    if sparse:
        bk, pk = bk * 1_000_003 + 11, pk * 1_000_003 + 11
    left = {"k": pk, "x": np.arange(2000, dtype=np.int32)}
    right = {"k": bk, "y": np.arange(nb, dtype=np.int32)}
    d = QueryDescriptor("l", join=JoinDescriptor(QueryDescriptor("r"), "k", "k"), projections=["x", "y"])
    got, want = run_both(qm, d, {"l": left, "r": right})
    assert len(got.rows) > 300_000
    assert_rows_equal(got.rows, want, ctx="join heavy duplicates")
    # the fused aggregate over the same duplicated keys (the unique-key index is refused)
    right["g"] = (np.arange(nb) % 5).astype(np.int32)
    d = QueryDescriptor("l", join=JoinDescriptor(QueryDescriptor("r"), "k", "k"), groupBy=["g"],
                        aggregates=[AggregateOp("count", "*"), AggregateOp("sum", "y"), AggregateOp("min", "x")])
    got, want = run_both(qm, d, {"l": left, "r": right})
    assert_groups_equal(got.rows, want, ["g"], ctx="join heavy duplicates aggregate")


@pytest.mark.parametrize("k", [10, 100, 1000])
def test_topk(qm, k):
    _say(f"case topk k={k}")
    cols = datagen.batch_table(n=1_000_000, seed=5)
    d = QueryDescriptor("t", sortColumn="v2", sortDirection="desc", limit=k)
    got, want = run_both(qm, d, {"t": cols})
    assert_rows_equal(got.rows, want, ctx=f"topk {k}")


def test_topk_filter_offset_asc(qm):
    _say("case topk offset asc")
    cols = datagen.batch_table(n=500_000, seed=6)
    cols["k1"] = cols["k1"]  # ties: sort on a 4-valued column, tie-break by row id
    d = QueryDescriptor("t", filters=[FilterOp("v2", "lt", 0.3)], sortColumn="k1", sortDirection="asc", limit=50,
                        offset=20, projections=["k2", "k1"])
    got, want = run_both(qm, d, {"t": cols})
    assert_rows_equal(got.rows, want, ctx="topk asc offset ties")


def test_select_limit_offset(qm):
    _say("case select limit")
    cols = datagen.filter_groupby_zipf(n=200_000, seed=8)
    d = QueryDescriptor("t", filters=[FilterOp("b", "between", [95.5, 96.5])], limit=37, offset=5, projections=["a", "b"])
    got, want = run_both(qm, d, {"t": cols})
    assert_rows_equal(got.rows, want, ctx="select")
    d2 = QueryDescriptor("t", filters=[FilterOp("a", "gte", 999000)], sortColumn="b")
    got, want = run_both(qm, d2, {"t": cols})
    assert_rows_equal(got.rows, want, ctx="select sort no limit")


@pytest.mark.parametrize("nullable", [False, True])
def test_partitioned_direct_and_nulls(qm, nullable):
    """Dense 1M-slot key domain (too large for LDS): partitioned by key slice."""
    _say(f"case partitioned direct nullable={nullable}")
    n = 2_000_000
    cols = datagen.batch_table(n=n, seed=9)
    cols["k2"] = cols["k2"] % 1_000_000
    nulls = None
    if nullable:
        rng = np.random.default_rng(3)
        nulls = {"t": {"k2": rng.random(n) < 0.01, "v2": rng.random(n) < 0.05, "v1": rng.random(n) < 0.02}}
    d = QueryDescriptor("t", filters=[FilterOp("v2", "lt", 0.5)], groupBy=["k2"],
                        aggregates=[AggregateOp("sum", "v2"), AggregateOp("count", "*"), AggregateOp("count", "v1"),
                                    AggregateOp("min", "v1"), AggregateOp("max", "v1"), AggregateOp("sum", "v1"),
                                    AggregateOp("avg", "v2")])
    got, want = run_both(qm, d, {"t": cols}, nulls)
    assert got.stats["strategy"] == "part_direct"
    assert_groups_equal(got.rows, want, ["k2"], float_cols={"sum_v2", "avg_v2"}, ctx="part direct")


def test_partitioned_hash_nullable_two_keys(qm):
    _say("case partitioned hash nullable + two-column key")
    n = 1_500_000
    cols = datagen.filter_groupby_zipf(n=n, seed=10)
    rng = np.random.default_rng(11)
    nulls = {"t": {"key": rng.random(n) < 0.02, "b": rng.random(n) < 0.03}}
    d = QueryDescriptor("t", filters=[FilterOp("flag", "eq", 1)], groupBy=["key"],
                        aggregates=[AggregateOp("sum", "b"), AggregateOp("count", "b"), AggregateOp("max", "a")])
    got, want = run_both(qm, d, {"t": cols}, nulls)
    assert got.stats["strategy"] == "part_hash"
    assert_groups_equal(got.rows, want, ["key"], float_cols={"sum_b"}, ctx="part hash nulls")
    cols2 = {"x": (cols["a"] % 3000).astype(np.int32), "y": (cols["key"] % 700).astype(np.int64), "b": cols["b"]}
    d2 = QueryDescriptor("u", groupBy=["x", "y"], aggregates=[AggregateOp("sum", "b"), AggregateOp("count", "*")])
    got, want = run_both(qm, d2, {"u": cols2})
    assert_groups_equal(got.rows, want, ["x", "y"], float_cols={"sum_b"}, ctx="two keys")
