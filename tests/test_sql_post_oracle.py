"""Pins of the oracle's SQL post-stages (oracle/sql_post.py, oracle.run):
HAVING, multi-column ORDER BY and DISTINCT.

* Known answers: the reference's SQL integration fixture and expectations
  (src/sql/integration.test.ts:6-12, 147-247), with the string columns coded
  as integers (dept: eng = 0, sales = 1; name: the row id).
* Closed forms: SQL three-valued AND / OR / NOT tables; compare() and
  looseEqual() against exact rational arithmetic (fractions.Fraction) on
  random int / float pairs, including fractions, +-0, +-Infinity and NaN.
* Brute force: the multi-column sort against Python's stable tuple sort, and
  DISTINCT against first occurrence by canonical key, on NULL-free data.
"""
from __future__ import annotations

import math
import random
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from oracle import sql_post as SP
from querymode_amd.plan import AggregateOp, QueryDescriptor
from querymode_amd.sql import sql_to_descriptor

# id, name -> id, age, dept (eng 0, sales 1), salary
SAMPLE = {
    "id": np.array([1, 2, 3, 4, 5], dtype=np.int32),
    "age": np.array([30, 25, 35, 28, 32], dtype=np.int32),
    "dept": np.array([0, 0, 1, 1, 0], dtype=np.int32),
    "salary": np.array([120000, 95000, 110000, 105000, 130000], dtype=np.int32),
}
NAME = {1: "Alice", 2: "Bob", 3: "Charlie", 4: "Diana", 5: "Eve"}


def _run(sql):
    return O.run(sql_to_descriptor(sql), {"data": O.Table(SAMPLE)})


def test_known_multi_order_by():
    rows = _run("SELECT id FROM data ORDER BY dept ASC, salary DESC")
    assert [NAME[r["id"]] for r in rows] == ["Eve", "Alice", "Bob", "Charlie", "Diana"]
    rows = _run("SELECT id FROM data ORDER BY dept ASC, salary DESC LIMIT 3")
    assert [NAME[r["id"]] for r in rows] == ["Eve", "Alice", "Bob"]
    rows = _run("SELECT id FROM data ORDER BY age ASC, salary DESC")
    assert [NAME[r["id"]] for r in rows] == ["Bob", "Diana", "Alice", "Eve", "Charlie"]
    assert list(rows[0].keys()) == ["id"]


def test_known_having():
    rows = _run("SELECT dept FROM data GROUP BY dept HAVING SUM(salary) > 300000")
    assert rows == [{"dept": 0}]  # eng 345K > 300K, sales 215K; sum_salary not returned
    rows = _run("SELECT dept FROM data GROUP BY dept HAVING COUNT(*) >= 3")
    assert rows == [{"dept": 0}]


def test_distinct_first_rows():
    rows = _run("SELECT DISTINCT dept FROM data")
    assert rows == [{"dept": 0}, {"dept": 1}]
    rows = O.run(QueryDescriptor("data", projections=["dept", "id"], distinct=["dept"]), {"data": O.Table(SAMPLE)})
    assert rows == [{"dept": 0, "id": 1}, {"dept": 1, "id": 3}]  # the first row of each key


T, F, N = True, False, None


@pytest.mark.parametrize("a", [T, F, N])
@pytest.mark.parametrize("b", [T, F, N])
def test_three_valued_tables(a, b):
    # Kleene logic: AND is the minimum, OR the maximum, of F < N < T
    rank = {F: 0, N: 1, T: 2}
    inv = {0: F, 1: N, 2: T}
    leaf = {T: ("isnull", "x", False), F: ("isnull", "y", False), N: ("cmp", "eq", "z", 1)}
    row = {"x": None, "y": 5, "z": None}
    assert SP.evaluate(leaf[a], row) is a
    assert SP.evaluate(("and", leaf[a], leaf[b]), row) is inv[min(rank[a], rank[b])]
    assert SP.evaluate(("or", leaf[a], leaf[b]), row) is inv[max(rank[a], rank[b])]
    assert SP.evaluate(("not", leaf[a]), row) is inv[2 - rank[a]]


def _sign(x):
    return (x > 0) - (x < 0)


def _values(rng):
    specials = [0.0, -0.0, 1.5, -1.5, 2.0 ** 53, 2.0 ** 63, -2.0 ** 63, 1e300, -1e-300, 7.25]
    ints = [0, 1, -1, 7, -8, 2 ** 53 + 1, 2 ** 62, -(2 ** 63), 2 ** 63 - 1]
    out = []
    for _ in range(400):
        if rng.random() < 0.5:
            out.append(rng.choice(ints) if rng.random() < 0.5 else rng.randint(-10 ** 6, 10 ** 6))
        else:
            out.append(rng.choice(specials) if rng.random() < 0.5 else rng.uniform(-1e6, 1e6))
    return out


def test_compare_is_exact_for_finite_numbers():
    rng = random.Random(5)
    vals = _values(rng)
    for _ in range(4000):
        a, b = rng.choice(vals), rng.choice(vals)
        want = _sign(Fraction(a) - Fraction(b))
        got = SP.compare(a, b)
        assert _sign(got) == want, (a, b, got)
        if isinstance(a, int) or isinstance(b, int):
            # looseEqual: exact where one side is a safe integer, else the double compare
            fa = isinstance(a, float) and not SP._safe_int(a)
            fb = isinstance(b, float) and not SP._safe_int(b)
            if not (fa or fb):
                assert SP.loose_equal(a, b) == (Fraction(a) == Fraction(b)), (a, b)


def test_compare_nan_and_infinity():
    inf, nan = math.inf, math.nan
    assert SP.compare(nan, nan) == 0 and SP.compare(nan, 1.0) > 0 and SP.compare(1.0, nan) < 0
    assert SP.compare(5, nan) < 0 and SP.compare(nan, 5) > 0          # bigint vs NaN: NaN last
    assert SP.compare(5, inf) < 0 and SP.compare(5, -inf) > 0
    assert math.isnan(SP.compare(inf, inf))                             # Infinity - Infinity
    assert SP.evaluate(("cmp", "ge", "v", 1), {"v": inf}) is True
    assert SP.evaluate(("cmp_col", "le", "a", "b"), {"a": inf, "b": inf}) is False  # NaN difference
    assert SP.evaluate(("cmp_col", "eq", "a", "b"), {"a": inf, "b": inf}) is True
    assert SP.evaluate(("cmp", "gt", "v", 5), {"v": nan}) is True       # NaN sorts greatest
    assert SP.evaluate(("cmp", "eq", "v", 0), {"v": -0.0}) is True
    assert SP.evaluate(("in", "v", [1, None], False), {"v": 2}) is None
    assert SP.evaluate(("in", "v", [1, None], True), {"v": 1}) is False
    assert SP.evaluate(("between", "v", 1, 2.5, False), {"v": 2}) is True


def test_order_by_matches_tuple_sort():
    rng = np.random.default_rng(3)
    rows = [{"a": int(x), "b": float(y), "i": i} for i, (x, y) in
            enumerate(zip(rng.integers(0, 5, 300), rng.integers(0, 7, 300) / 2))]
    got = SP.order_by(rows, [("a", "desc"), ("b", "asc")])
    want = sorted(rows, key=lambda r: (-r["a"], r["b"]))  # stable
    assert [r["i"] for r in got] == [r["i"] for r in want]


def test_order_by_nulls_last_both_directions():
    rows = [{"a": None, "i": 0}, {"a": 2.0, "i": 1}, {"a": math.nan, "i": 2}, {"a": -1.0, "i": 3}, {"a": None, "i": 4}]
    asc = [r["i"] for r in SP.order_by(rows, [("a", "asc")])]
    dsc = [r["i"] for r in SP.order_by(rows, [("a", "desc")])]
    assert asc == [3, 1, 2, 0, 4]   # NaN above every number, NULLs last, ties stable
    assert dsc == [2, 1, 3, 0, 4]   # descending flips NaN to the front; NULLs stay last


def test_distinct_matches_first_occurrence():
    rng = np.random.default_rng(4)
    pool = [0.0, -0.0, math.nan, 1.5, None]
    rows = [{"a": int(rng.integers(0, 4)), "b": pool[int(rng.integers(0, len(pool)))], "i": i} for i in range(500)]
    got = SP.distinct(rows, ["a", "b"])

    def canon(v):
        if v is None:
            return ("null",)
        if isinstance(v, float) and math.isnan(v):
            return ("nan",)
        return (v + 0.0 if isinstance(v, float) else v,)  # -0.0 + 0.0 == 0.0

    first = {}
    for r in rows:
        first.setdefault((canon(r["a"]), canon(r["b"])), r["i"])
    assert [r["i"] for r in got] == sorted(first.values())
    assert len(got) == 4 * 4  # 0.0 and -0.0 are one value


def test_having_drops_having_only_aggregates():
    cols = {"k": np.array([1, 1, 2, 3, 3, 3], dtype=np.int64), "v": np.array([1, 2, 3, 4, 5, 6], dtype=np.int32)}
    d = QueryDescriptor("t", groupBy=["k"], aggregates=[AggregateOp("count", "*", "n"), AggregateOp("sum", "v")],
                        projections=["k", "n"], having=("cmp", "gt", "sum_v", 3))
    rows = O.run(d, {"t": O.Table(cols)})
    assert sorted(r["k"] for r in rows) == [3] and all(set(r) == {"k", "n"} for r in rows)
