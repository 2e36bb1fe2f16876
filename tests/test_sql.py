"""SQL front end: lowering to QueryDescriptor (src/sql/compiler.ts:39-140,
287-310, 365-615)."""
from __future__ import annotations

import pytest

from querymode_amd.errors import QueryModeError
from querymode_amd.plan import AggregateOp, FilterOp
from querymode_amd.sql import sql_to_descriptor as S


def test_c2_query():
    d = S("SELECT key, SUM(b), COUNT(*), MIN(a), MAX(a), AVG(b) FROM t "
          "WHERE a < 300000 AND b > 90 AND flag = 1 GROUP BY key")
    assert d.table == "t"
    assert d.filters == [FilterOp("a", "lt", 300000), FilterOp("b", "gt", 90), FilterOp("flag", "eq", 1)]
    assert [a.out_name() for a in d.aggregates] == ["sum_b", "count_*", "min_a", "max_a", "avg_b"]
    assert d.groupBy == ["key"] and d.projections == ["key", "sum_b", "count_*", "min_a", "max_a", "avg_b"]


def test_join_product_aggregate():
    d = S("SELECT region, SUM(amount * weight) AS aw, COUNT(*) FROM probe p JOIN build b ON p.fk = b.id "
          "GROUP BY region")
    assert d.join.leftKey == "fk" and d.join.rightKey == "id" and d.join.right.table == "build"
    assert d.aggregates[0] == AggregateOp("sum", "multiply(amount,weight)", "aw")
    # ON written right-side first swaps the keys (compiler.ts:574-586)
    d2 = S("SELECT COUNT(*) FROM probe JOIN build ON build.id = probe.fk")
    assert (d2.join.leftKey, d2.join.rightKey) == ("fk", "id")
    d3 = S("SELECT COUNT(*) FROM probe JOIN build USING (id)")
    assert (d3.join.leftKey, d3.join.rightKey) == ("id", "id")


def test_order_limit_offset():
    d = S("SELECT * FROM t ORDER BY v2 DESC LIMIT 10 OFFSET 3")
    assert (d.sortColumn, d.sortDirection, d.limit, d.offset) == ("v2", "desc", 10, 3)
    assert S("SELECT a FROM t ORDER BY a").sortDirection == "asc"


def test_or_groups_and_precedence():
    d = S("SELECT * FROM t WHERE a > 2 AND b < 50 OR a = 1")
    assert d.filters == [] and d.filterGroups == [[FilterOp("a", "gt", 2), FilterOp("b", "lt", 50)],
                                                  [FilterOp("a", "eq", 1)]]
    d = S("SELECT * FROM t WHERE a = 1 OR (b < 2 OR a BETWEEN 1 AND 3)")
    assert d.filters == [] and d.filterGroups == [[FilterOp("a", "eq", 1)], [FilterOp("b", "lt", 2)],
                                                  [FilterOp("a", "between", [1, 3])]]


def test_not_forms_on_the_filter_path():
    # the NOT forms the reference's compiler pushes down (compiler.ts:383-451)
    assert S("SELECT * FROM t WHERE a NOT BETWEEN 2 AND 5").filters == [FilterOp("a", "not_between", [2, 5])]
    assert S("SELECT * FROM t WHERE NOT a BETWEEN 2 AND 5").filters == [FilterOp("a", "not_between", [2, 5])]
    assert S("SELECT * FROM t WHERE a NOT IN (1, 2)").filters == [FilterOp("a", "not_in", [1, 2])]
    assert S("SELECT * FROM t WHERE a IS NOT NULL").filters == [FilterOp("a", "is_not_null", 0)]
    d = S("SELECT * FROM t WHERE a NOT BETWEEN 1 AND 2 OR b IS NULL")
    assert d.filterGroups == [[FilterOp("a", "not_between", [1, 2])], [FilterOp("b", "is_null", 0)]]


@pytest.mark.parametrize("where", [
    "NOT (NOT (a > 5))", "NOT a > 5", "NOT (a < 5)", "NOT (i >= 2.5)", "NOT (a > 5 AND b < 50)",
    "NOT (a > 5 OR b < 50)", "NOT a IN (1, 2)", "NOT a IS NULL", "NOT a NOT BETWEEN 2 AND 5",
    "a > 2 AND (b < 50 OR a = 1)", "(a = 1 OR b > 50) AND a < 10", "a > 1 OR NOT (b < 2)",
    "(a > 1 OR b < 2) AND (a < 0 OR b > 9)",
])
def test_row_by_row_shapes_are_unsupported(where):
    """Clauses the reference's compiler leaves to its row-by-row evaluator
    (compiler.ts:90-109, executor.ts:89-91), whose comparisons put NaN last and
    compare integer columns exactly against fractions (evaluator.ts:169-200):
    lowering them to filter-path predicates would change their rows (NOT (x < 5)
    keeps NaN there; NOT (i >= 2.5) keeps i = 2), so they are refused."""
    with pytest.raises(QueryModeError) as e:
        S("SELECT * FROM t WHERE " + where)
    assert e.value.code == "UNSUPPORTED"


def test_literals():
    assert S("SELECT * FROM t WHERE 5 < a").filters == [FilterOp("a", "gt", 5)]
    assert S("SELECT * FROM t WHERE a >= -2.5").filters == [FilterOp("a", "gte", -2.5)]
    assert S("SELECT * FROM t WHERE a <> 1e3").filters == [FilterOp("a", "neq", 1000.0)]


@pytest.mark.parametrize("sql", [
    "SELECT * FROM t WHERE c LIKE 'f%'",
    "SELECT * FROM t WHERE a = NULL",
    "SELECT * FROM t WHERE a > b",
    "SELECT * FROM t WHERE c = 'x'",
    "SELECT * FROM l LEFT JOIN r ON l.k = r.k",
    "SELECT k FROM t GROUP BY k HAVING SUM(a) * 2 > 3",
    "SELECT k FROM t GROUP BY k HAVING COUNT(*) > 'x'",
    "SELECT k, COUNT(*) FROM t HAVING COUNT(*) > 1",
])
def test_off_path_is_unsupported(sql):
    with pytest.raises(QueryModeError) as e:
        S(sql)
    assert e.value.code == "UNSUPPORTED"


@pytest.mark.parametrize("sql", ["SELEC * FROM t", "SELECT * FROM t WHERE", "SELECT * FROM t LIMIT x",
                                 "SELECT * FROM t WHERE a >"])
def test_malformed_is_invalid_format(sql):
    with pytest.raises(QueryModeError) as e:
        S(sql)
    assert e.value.code == "INVALID_FORMAT"


def test_post_stages_lowering():
    d = S("SELECT * FROM t ORDER BY a, b DESC LIMIT 5")
    assert d.orderBy == [("a", "asc"), ("b", "desc")] and d.sortColumn is None and d.limit == 5
    d = S("SELECT DISTINCT a, b FROM t")
    assert d.distinct == ["a", "b"] and d.projections == ["a", "b"]
    d = S("SELECT DISTINCT * FROM t")
    assert d.distinct == []
    # HAVING-only aggregates are computed but not projected (compiler.ts:80-88, 159-164)
    d = S("SELECT dept FROM data GROUP BY dept HAVING SUM(salary) > 300000")
    assert d.projections == ["dept"] and [a.out_name() for a in d.aggregates] == ["sum_salary"]
    assert d.having == ("cmp", "gt", "sum_salary", 300000)
    # a SELECT alias names the HAVING column; a literal on the left flips the comparison
    d = S("SELECT k, COUNT(*) AS n FROM t GROUP BY k HAVING 5 < COUNT(*) AND MAX(v) IS NOT NULL")
    assert d.having == ("and", ("cmp", "gt", "n", 5), ("isnull", "max_v", True))
    assert d.projections == ["k", "n"] and [a.out_name() for a in d.aggregates] == ["n", "max_v"]
    d = S("SELECT k, SUM(a) FROM t GROUP BY k HAVING SUM(a) > SUM(b) OR k IN (1, NULL)")
    assert d.having == ("or", ("cmp_col", "gt", "sum_a", "sum_b"), ("in", "k", [1, None], False))
