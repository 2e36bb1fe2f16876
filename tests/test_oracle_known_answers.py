"""The CPU oracle against the known answers in the reference's own tests.

Each case cites the reference test it restates. String columns of those
tests (region "us"/"eu"/…) are carried as small integer codes, since the
device path is numeric; the expected values are the reference's, or — for the
5M-row generator tables — closed forms derived from the generator's
periodicity, not recomputed with the oracle's own formulas.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd.plan import AggregateOp, FilterOp, JoinDescriptor, QueryDescriptor


def _q(table="t", **kw) -> QueryDescriptor:
    return QueryDescriptor(table=table, **kw)


def _by(rows, key):
    return {r[key] if not (isinstance(r[key], float) and math.isnan(r[key])) else "nan": r for r in rows}


# ---------------------------------------------------------------- partial-agg.test.ts

def test_update_partial_agg_sum_count_min_max():
    # src/partial-agg.test.ts:22-32: values 10, 20, 5 -> sum 35, count 3, min 5, max 20
    t = O.Table({"price": np.array([10, 20, 5], dtype=np.float64)})
    r = O.run(_q(aggregates=[AggregateOp("sum", "price"), AggregateOp("count", "price"),
                             AggregateOp("min", "price"), AggregateOp("max", "price"),
                             AggregateOp("avg", "price")]), {"t": t})
    assert r == [{"sum_price": 35.0, "count_price": 3, "min_price": 5.0, "max_price": 20.0,
                  "avg_price": 35.0 / 3}]


def test_compute_partial_agg_sum():
    # src/partial-agg.test.ts:34-47: amounts 10, 20, 30 -> sum 60, count 3
    t = O.Table({"amount": np.array([10, 20, 30], dtype=np.int32)})
    r = O.run(_q(aggregates=[AggregateOp("sum", "amount"), AggregateOp("count", "*")]), {"t": t})
    assert r == [{"sum_amount": 60, "count_*": 3}]


def test_compute_partial_agg_group_by():
    # src/partial-agg.test.ts:49-68: us: 10 + 30 = 40, eu: 20 + 5 = 25 (us=0, eu=1)
    t = O.Table({"region": np.array([0, 1, 0, 1], dtype=np.int32),
                 "amount": np.array([10, 20, 30, 5], dtype=np.int32)})
    r = O.run(_q(groupBy=["region"], aggregates=[AggregateOp("sum", "amount")]), {"t": t})
    # first-seen group order (Map insertion order)
    assert r == [{"region": 0, "sum_amount": 40}, {"region": 1, "sum_amount": 25}]


def test_merge_partials_equals_whole():
    # src/partial-agg.test.ts:70-91: partials [10, 20] and [5, 100] merge to
    # sum 135, count 4, min 5, max 100 — the merged result is the result over all rows
    t = O.Table({"amount": np.array([10, 20, 5, 100], dtype=np.int64)})
    r = O.run(_q(aggregates=[AggregateOp("sum", "amount"), AggregateOp("count", "*"),
                             AggregateOp("min", "amount"), AggregateOp("max", "amount")]), {"t": t})
    assert r == [{"sum_amount": 135, "count_*": 4, "min_amount": 5, "max_amount": 100}]


def test_finalize_avg_sum_min_max_count():
    # src/partial-agg.test.ts:126-148: v = 10, 20, 30 -> avg 20, sum 60, min 10, max 30, count 3
    t = O.Table({"v": np.array([10, 20, 30], dtype=np.int32)})
    r = O.run(_q(aggregates=[AggregateOp("avg", "v"), AggregateOp("sum", "v"), AggregateOp("min", "v"),
                             AggregateOp("max", "v"), AggregateOp("count", "v")]), {"t": t})
    assert r == [{"avg_v": 20.0, "sum_v": 60, "min_v": 10, "max_v": 30, "count_v": 3}]


def test_nan_and_infinite_group_keys():
    # src/partial-agg.test.ts:150-177: keys NaN, NaN, +Inf, -Inf -> 3 groups,
    # sum_v NaN: 3, +Inf: 3, -Inf: 4
    t = O.Table({"g": np.array([np.nan, np.nan, np.inf, -np.inf]),
                 "v": np.array([1, 2, 3, 4], dtype=np.float64)})
    r = O.run(_q(groupBy=["g"], aggregates=[AggregateOp("sum", "v")]), {"t": t})
    assert len(r) == 3
    g = _by(r, "g")
    assert g["nan"]["sum_v"] == 3.0
    assert g[math.inf]["sum_v"] == 3.0
    assert g[-math.inf]["sum_v"] == 4.0


def test_empty_input_aggregates():
    # src/partial-agg.ts:306-352: an empty input gives null aggregates, count 0
    t = O.Table({"v": np.array([], dtype=np.float64)})
    r = O.run(_q(aggregates=[AggregateOp("sum", "v"), AggregateOp("count", "*"), AggregateOp("min", "v"),
                             AggregateOp("avg", "v")]), {"t": t})
    assert r == [{"sum_v": None, "count_*": 0, "min_v": None, "avg_v": None}]


# ---------------------------------------------------------------- conformance.test.ts

TRUTH = [
    # src/conformance.test.ts:702-739 (matchesFilter truth table, numeric cases)
    (42, "eq", 42, True), (41, "eq", 42, False),
    (42, "neq", 42, False), (43, "neq", 42, True),
    (10, "gt", 10, False), (11, "gt", 10, True), (10, "gte", 10, True), (9, "gte", 10, False),
    (10, "lt", 10, False), (9, "lt", 10, True), (10, "lte", 10, True), (11, "lte", 10, False),
    (2, "in", [1, 2, 3], True), (4, "in", [1, 2, 3], False),
]


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float64, np.int8, np.float32])
@pytest.mark.parametrize("v,op,lit,want", TRUTH)
def test_matches_filter_truth_table(dtype, v, op, lit, want):
    t = O.Table({"x": np.array([v], dtype=dtype)})
    got = O.filter_rows(t, [FilterOp("x", op, lit)])
    assert (len(got) == 1) == want


@pytest.mark.parametrize("op,lit", [("eq", 0), ("gt", -999), ("in", [0]), ("lt", 1), ("neq", 5),
                                    ("between", [-1, 1]), ("not_between", [5, 6]), ("not_in", [7])])
def test_null_never_matches(op, lit):
    # src/conformance.test.ts:735-739: a null value never matches (decode.ts:485-491)
    t = O.Table({"x": np.array([0], dtype=np.int64)}, nulls={"x": np.array([True])})
    assert len(O.filter_rows(t, [FilterOp("x", op, lit)])) == 0
    assert len(O.filter_rows(t, [FilterOp("x", "is_null")])) == 1


# ------------------------------------------- operators-conformance.test.ts generator

def _generate_rows(n: int) -> dict:
    """generateRows (src/operators-conformance.test.ts:101-113) with region /
    category as codes i % 4 / i % 3."""
    i = np.arange(n, dtype=np.int64)
    return {"id": i.astype(np.int32), "region": (i % 4).astype(np.int32), "category": (i % 3).astype(np.int32),
            "amount": ((i * 7 + 13) % 1000).astype(np.int32), "score": ((i * 31) % 997) / 10.0}


@pytest.fixture(scope="module")
def rows5m():
    return O.Table(_generate_rows(5_000_000))


def test_groupby_region_5m_closed_form(rows5m):
    # src/operators-conformance.test.ts:408-447 (sum/avg/min/max/count by region, 5M).
    # amount = (7i + 13) % 1000 repeats every 1000 rows; within a period the rows
    # of region r (i = 4m + r) take every value v < 1000 with v = c (mod 4),
    # c = (3r + 1) % 4, exactly once (28m mod 1000 runs over all multiples of 4).
    # So per period: 250 values, sum 250c + 4 * (0 + ... + 249); 5000 periods.
    r = O.run(_q(groupBy=["region"], aggregates=[
        AggregateOp("sum", "amount", "sum_amount"), AggregateOp("avg", "amount", "avg_amount"),
        AggregateOp("min", "amount", "min_amount"), AggregateOp("max", "amount", "max_amount"),
        AggregateOp("count", "amount", "cnt")]), {"t": rows5m})
    assert [x["region"] for x in r] == [0, 1, 2, 3]  # first-seen order
    for x in r:
        c = (3 * x["region"] + 1) % 4
        per = 250 * c + 4 * (249 * 250 // 2)
        assert x["sum_amount"] == 5000 * per
        assert x["avg_amount"] == per / 250 == c + 498
        assert x["min_amount"] == c
        assert x["max_amount"] == 996 + c
        assert x["cnt"] == 1_250_000


def test_topk_100_5m_all_999(rows5m):
    # src/operators-conformance.test.ts:576-594: the top 100 amounts of 5M rows
    # are all 999. 7i + 13 = 999 (mod 1000) <=> i = 998 (mod 1000), since
    # 7 * 143 = 1001; under ties by row id ascending the rows are 998 + 1000k.
    r = O.run(_q(sortColumn="amount", sortDirection="desc", limit=100, projections=["id", "amount"]),
              {"t": rows5m})
    assert [x["amount"] for x in r] == [999] * 100
    assert [x["id"] for x in r] == [998 + 1000 * k for k in range(100)]


def test_sort_ascending_extremes():
    # src/operators-conformance.test.ts:558-574: ascending sort starts at MIN(amount) and ends at MAX
    t = O.Table(_generate_rows(20_000))
    r = O.run(_q(sortColumn="amount", sortDirection="asc", projections=["id", "amount"]), {"t": t})
    assert r[0]["amount"] == 0 and r[-1]["amount"] == 999
    assert len(r) == 20_000
    assert all(a["amount"] <= b["amount"] for a, b in zip(r, r[1:]))


def test_inner_join_1m_x_500k():
    # src/operators-conformance.test.ts:194-218: 1M x 500K inner join on id gives
    # exactly 500,000 rows and value = id * 10
    left = O.Table(_generate_rows(1_000_000))
    rid = np.arange(500_000, dtype=np.int32)
    right = O.Table({"id": rid, "value": rid * 10})
    tables = {"l": left, "r": right}
    j = JoinDescriptor(_q("r"), "id", "id")
    agg = O.run(_q("l", join=j, aggregates=[AggregateOp("count", "*"), AggregateOp("sum", "value"),
                                            AggregateOp("min", "id"), AggregateOp("max", "id"),
                                            AggregateOp("sum", "id")]), tables)
    assert agg == [{"count_*": 500_000, "sum_value": 10 * (499_999 * 500_000 // 2), "min_id": 0,
                    "max_id": 499_999, "sum_id": 499_999 * 500_000 // 2}]
    rows = O.run(_q("l", join=j, filters=[FilterOp("id", "lt", 3000)],
                    projections=["id", "region", "amount", "value"]), tables)
    assert len(rows) == 3000
    assert all(x["value"] == x["id"] * 10 for x in rows)
    assert [x["id"] for x in rows] == list(range(3000))  # probe order


def test_inner_join_1m_x_1m_count():
    # src/operators-conformance.test.ts:314-359: 1M x 1M on id gives 1,000,000 rows
    left = O.Table(_generate_rows(1_000_000))
    right = O.Table({"id": np.arange(1_000_000, dtype=np.int32)[::-1].copy(),
                     "value": np.arange(1_000_000, dtype=np.int32)})
    r = O.run(_q("l", join=JoinDescriptor(_q("r"), "id", "id"), aggregates=[AggregateOp("count", "*")]),
              {"l": left, "r": right})
    assert r == [{"count_*": 1_000_000}]
