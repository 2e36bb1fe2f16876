"""HAVING, multi-column ORDER BY and DISTINCT on the device against the oracle
(oracle/sql_post.py, pinned in tests/test_sql_post_oracle.py).

Data holds what these stages decide on: NULLs, NaN, +-Infinity, -0 next to
+0, ties in every key, int64 values beyond 2^53 and bigint literals.
Group-by results are compared as multisets unless an ORDER BY fixes their
order; row results are compared in order (a full tie keeps row order on both
sides)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd.plan import AggregateOp, BigInt, FilterOp, QueryDescriptor
from querymode_amd.sql import sql_to_descriptor

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu

SAMPLE = {
    "id": np.array([1, 2, 3, 4, 5], dtype=np.int32),
    "age": np.array([30, 25, 35, 28, 32], dtype=np.int32),
    "dept": np.array([0, 0, 1, 1, 0], dtype=np.int32),
    "salary": np.array([120000, 95000, 110000, 105000, 130000], dtype=np.int32),
}


def _table(n=60_000, seed=11):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 300, n).astype(np.int64)
    v = rng.normal(0, 100, n)
    special = rng.random(n)
    v[special < 0.01] = np.nan
    v[(special >= 0.01) & (special < 0.015)] = np.inf
    v[(special >= 0.015) & (special < 0.02)] = -np.inf
    v[(special >= 0.02) & (special < 0.05)] = 0.0
    v[(special >= 0.05) & (special < 0.08)] = -0.0
    w = rng.integers(-50, 50, n).astype(np.int32)
    big = (rng.integers(0, 5, n).astype(np.int64) + (1 << 60))
    a = rng.integers(0, 6, n).astype(np.int32)
    cols = {"k": k, "v": v, "w": w, "big": big, "a": a}
    nulls = {"v": rng.random(n) < 0.03, "w": rng.random(n) < 0.05, "a": rng.random(n) < 0.04}
    return cols, nulls


def _both(qm, d, cols, nulls, name="post_t"):
    got = qm.executor.execute(d).rows
    want = O.run(d, {name: O.Table(cols, nulls)})
    return got, want


def test_reference_known_answers(qm):
    print("sql post: the reference's SQL integration expectations on the device", flush=True)
    qm.register_table("data", SAMPLE)
    ex = qm.executor
    ids = lambda sql: [r["id"] for r in ex.execute(sql_to_descriptor(sql)).rows]  # noqa: E731
    assert ids("SELECT id FROM data ORDER BY dept ASC, salary DESC") == [5, 1, 2, 3, 4]
    assert ids("SELECT id FROM data ORDER BY dept ASC, salary DESC LIMIT 3") == [5, 1, 2]
    r = ex.execute(sql_to_descriptor("SELECT id FROM data ORDER BY age ASC, salary DESC"))
    assert [x["id"] for x in r.rows] == [2, 4, 1, 5, 3] and r.columns == ["id"]
    r = ex.execute(sql_to_descriptor("SELECT dept FROM data GROUP BY dept HAVING SUM(salary) > 300000"))
    assert r.rows == [{"dept": 0}] and r.columns == ["dept"]
    r = ex.execute(sql_to_descriptor("SELECT dept FROM data GROUP BY dept HAVING COUNT(*) >= 3"))
    assert r.rows == [{"dept": 0}]
    assert [x["dept"] for x in ex.execute(sql_to_descriptor("SELECT DISTINCT dept FROM data")).rows] == [0, 1]


HAVINGS = [
    ("cmp", "gt", "n", 200),
    ("cmp", "le", "sum_w", -10),
    ("cmp", "gt", "max_v", 250.5),                       # +Infinity and NaN groups pass, NaN is greatest
    ("cmp", "lt", "max_v", math.inf),
    ("cmp", "eq", "min_w", -50),
    ("cmp", "ne", "min_w", -50),
    ("cmp", "ge", "sum_v", 0),
    ("cmp", "gt", "max_big", BigInt((1 << 60) + 3)),    # exact int64 beyond 2^53
    ("cmp", "eq", "max_big", float(1 << 60)),           # a number literal above 2^53
    ("cmp_col", "gt", "sum_w", "n"),
    ("between", "avg_w", -1.5, 1.5, False),
    ("between", "avg_w", -1.5, 1.5, True),
    ("in", "min_w", [-50, -49, None], False),            # not found -> NULL -> dropped
    ("in", "min_w", [-50, -49, None], True),
    ("isnull", "sum_v", True),
    ("and", ("cmp", "gt", "n", 195), ("not", ("cmp", "lt", "max_v", 300))),
    ("or", ("cmp", "gt", "n", 215), ("in", "min_w", [-48, None], False)),
    ("not", ("or", ("cmp", "lt", "sum_w", 0), ("isnull", "max_v", False))),
]


@pytest.mark.parametrize("hv", range(len(HAVINGS)))
def test_having(qm, hv):
    cols, nulls = _table()
    qm.register_table("post_t", cols, nulls)
    d = QueryDescriptor("post_t", groupBy=["k"], filters=[FilterOp("a", "lt", 5)],
                        aggregates=[AggregateOp("count", "*", "n"), AggregateOp("sum", "w"), AggregateOp("sum", "v"),
                                    AggregateOp("min", "w"), AggregateOp("max", "v"), AggregateOp("avg", "w"),
                                    AggregateOp("max", "big")],
                        having=HAVINGS[hv])
    print(f"sql post: HAVING {HAVINGS[hv]}", flush=True)
    got, want = _both(qm, d, cols, nulls)
    assert_groups_equal(got, want, ["k"], float_cols=("sum_v", "avg_w"), ctx=f"having {hv}")


ORDERS = [
    [("a", "asc"), ("v", "desc")],
    [("v", "asc"), ("a", "desc"), ("w", "asc")],
    [("w", "desc"), ("big", "asc")],
    [("a", "desc"), ("w", "desc"), ("v", "asc"), ("k", "asc")],
]


@pytest.mark.parametrize("oi", range(len(ORDERS)))
@pytest.mark.parametrize("lim", [(None, None), (100, None), (50, 1000)])
def test_multi_order_by_rows(qm, oi, lim):
    cols, nulls = _table(n=40_000, seed=12)
    qm.register_table("post_t", cols, nulls)
    d = QueryDescriptor("post_t", filters=[FilterOp("k", "lt", 150)], projections=["k", "v", "a", "w"],
                        orderBy=ORDERS[oi], limit=lim[0], offset=lim[1])
    print(f"sql post: ORDER BY {ORDERS[oi]} limit/offset {lim}", flush=True)
    got, want = _both(qm, d, cols, nulls)
    assert_rows_equal(got, want, ctx=f"order {oi} {lim}")


def test_multi_order_by_groups(qm):
    cols, nulls = _table(n=40_000, seed=13)
    qm.register_table("post_t", cols, nulls)
    d = QueryDescriptor("post_t", groupBy=["a"], aggregates=[AggregateOp("count", "*", "n"), AggregateOp("max", "v")],
                        orderBy=[("max_v", "desc"), ("a", "asc")], limit=4, having=("cmp", "gt", "n", 10))
    got, want = _both(qm, d, cols, nulls)
    assert_rows_equal(got, want, ctx="grouped order")


@pytest.mark.parametrize("case", ["a_v", "a", "all", "sorted", "sorted_limit"])
def test_distinct(qm, case):
    cols, nulls = _table(n=30_000, seed=14)
    qm.register_table("post_t", cols, nulls)
    base = dict(filters=[FilterOp("k", "lt", 200)])
    if case == "a_v":
        d = QueryDescriptor("post_t", projections=["a", "v"], distinct=["a", "v"], **base)
    elif case == "a":
        d = QueryDescriptor("post_t", projections=["a", "k", "w"], distinct=["a"], **base)
    elif case == "all":
        d = QueryDescriptor("post_t", projections=["a", "w"], distinct=[], **base)
    elif case == "sorted":
        d = QueryDescriptor("post_t", projections=["a", "w"], distinct=["a", "w"], orderBy=[("w", "desc"), ("a", "asc")],
                            **base)
    else:
        d = QueryDescriptor("post_t", projections=["a", "w"], distinct=["a", "w"], sortColumn="w",
                            sortDirection="desc", limit=20, offset=5, **base)
    print(f"sql post: DISTINCT {case}", flush=True)
    got, want = _both(qm, d, cols, nulls)
    assert_rows_equal(got, want, ctx=f"distinct {case}")


def test_having_in_a_batch(qm):
    cols, nulls = _table(n=50_000, seed=15)
    qm.register_table("post_t", cols, nulls)
    descs = [QueryDescriptor("post_t", groupBy=["k"], filters=[FilterOp("w", "lt", t)],
                             aggregates=[AggregateOp("count", "*", "n"), AggregateOp("sum", "w")],
                             having=("cmp", "gt", "n", 80 + t), orderBy=[("n", "desc"), ("k", "asc")], limit=10)
             for t in (-20, 0, 20, 40)]
    batch = qm.executor.execute_batch(descs)
    for i, (d, b) in enumerate(zip(descs, batch)):
        want = O.run(d, {"post_t": O.Table(cols, nulls)})
        assert_rows_equal(b.rows, want, ctx=f"batch {i}")
        assert_rows_equal(qm.executor.execute(d).rows, want, ctx=f"solo {i}")
