"""SQL WHERE golden vectors from the reference's PyLance conformance fixtures
(tests/golden/pylance_filters.json, written by make_pylance_vectors.py from
wasm/tests/conformance/fixtures/*.expected.json), run through the SQL front
end and the CPU oracle's filter (src/operators.ts:439-547, src/decode.ts:477-578).

Every vector is accounted for: it either runs on the numeric path and must
return exactly the fixture's rows, or it is listed below with the reason it is
not on that path. The GPU run of the same vectors is in test_gpu_golden.py.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd.errors import QueryModeError
from querymode_amd.sql import sql_to_descriptor

GOLDEN = json.loads((Path(__file__).parent / "golden" / "pylance_filters.json").read_text())

# Vectors off the numeric device path (string column c, LIKE, NULL literals,
# column-to-column comparisons): the SQL front end rejects them as UNSUPPORTED.
OFF_PATH = {
    "combo_between_and_like", "equals_empty_string", "in_string", "in_string_with_null", "in_with_null_in_list",
    "like_contains", "like_empty", "like_exact", "like_literal_percent",
    "like_literal_underscore", "like_match_empty", "like_mixed", "like_multiple_single", "like_null_pattern",
    "like_prefix", "like_single_char", "like_start_end", "like_suffix", "not_equals_empty_string", "not_like",
    "not_in_with_null_in_list", "null_column_like", "compare_with_null_column", "null_in_both_columns",
    "null_equals_null", "null_greater_than", "null_less_than", "null_not_equals_null", "between_both_null",
    "between_null_high", "between_null_low",
}
# Clauses the reference's SQL compiler does not push to the filter path (NOT
# over a comparison or over AND, an AND holding an OR): its row-by-row
# evaluator runs them (src/sql/compiler.ts:90-109, src/sql/executor.ts:89-91),
# so the SQL front end rejects them as UNSUPPORTED.
ROW_BY_ROW = {"and_or_precedence_2", "and_or_precedence_4", "multiple_not", "not_precedence_1", "not_precedence_2",
              "nested_and_or"}
# Vectors that lower to device predicates but test the string column c, which
# the numeric table does not carry.
STRING_COLUMN = {"is_null_string"}
# The SQL engine that wrote the fixtures rejects an int64 column compared with a
# float literal; the reference coerces the literal (BigInt(Math.trunc(x)),
# src/decode.ts:521-556), which on this table gives the rows listed here
# (row ids of the golden table: a = [1, 2, 3, null, 5, null, 7, 0, -1, 10]).
ENGINE_ERRORS = {
    "int_compare_float": [2, 4, 6, 9],   # a > 2.5  -> a > 2
    "int_equals_float": [2],              # a = 3.0  -> a = 3
    "in_float_for_int": [0, 1, 2],        # a IN (1.0, 2.0, 3.0)
}


def _table() -> O.Table:
    t = GOLDEN["table"]
    cols, nulls = {}, {}
    for name, dt in (("a", np.int64), ("b", np.int64), ("d", np.float64)):
        vals = t[name]
        nulls[name] = np.array([v is None for v in vals])
        cols[name] = np.array([0 if v is None else v for v in vals], dtype=dt)
    return O.Table(cols, nulls)


def _vectors():
    return [pytest.param(v, id=v["name"]) for v in GOLDEN["vectors"]]


def test_every_vector_is_classified():
    names = {v["name"] for v in GOLDEN["vectors"]}
    assert len(names) == 72
    assert OFF_PATH <= names and set(ENGINE_ERRORS) <= names and STRING_COLUMN <= names and ROW_BY_ROW <= names
    assert not (OFF_PATH & ROW_BY_ROW)
    errors = {v["name"] for v in GOLDEN["vectors"] if "engine_error" in v}
    assert errors == set(ENGINE_ERRORS)


@pytest.mark.parametrize("vec", _vectors())
def test_pylance_vector(vec):
    sql = "SELECT * FROM t WHERE " + vec["where"]
    if vec["name"] in OFF_PATH or vec["name"] in ROW_BY_ROW:
        with pytest.raises(QueryModeError) as e:
            sql_to_descriptor(sql)
        assert e.value.code == "UNSUPPORTED"
        return
    d = sql_to_descriptor(sql)
    if vec["name"] in STRING_COLUMN:
        cols = {f.column for f in d.filters} | {f.column for g in d.filterGroups for f in g}
        assert "c" in cols
        return
    got = O.filter_rows(_table(), d.filters, d.filterGroups).tolist()
    want = ENGINE_ERRORS.get(vec["name"], vec.get("rows"))
    assert got == want, f"{vec['where']}: rows {got} vs fixture {want}"
