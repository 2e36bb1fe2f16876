"""The filter matrix on the device: 6 column types x 15 operators with NULLs,
NaN and +-0, exact bigint literals on int64, ToInt32 coercion of fractional
and out-of-range literals on int32, and AND / OR groups. Device row ids must
equal the numpy restatement (and the oracle) exactly (src/operators.ts:550-666,
src/decode.ts:477-578)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd import qmcb
from querymode_amd.plan import BigInt, FilterOp, QueryDescriptor

from .test_oracle_bruteforce import ALL_DTYPES, CASES, INT32_LITERALS, N, _column, _expect, to_int32

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rows(qm, table, filters, groups=()):
    res = qm.executor.execute(QueryDescriptor(table, filters=list(filters), filterGroups=[list(g) for g in groups],
                                              projections=["rid"]))
    b = qmcb.decode(res.columnarData)
    return np.asarray(b.columns[0].data[:b.row_count], dtype=np.int64) if b.row_count else np.zeros(0, np.int64)


def test_filter_matrix_on_device(qm):
    rng = np.random.default_rng(20241204)
    cols, nulls = {}, {}
    for i, dt in enumerate(ALL_DTYPES):
        cols[f"x{i}"] = _column(rng, dt, N)
        nulls[f"x{i}"] = rng.random(N) < 0.05
    cols["rid"] = np.arange(N, dtype=np.int64)
    qm.register_table("fm", cols, nulls)
    ot = O.Table(cols, nulls)
    for i, dt in enumerate(ALL_DTYPES):
        c = f"x{i}"
        _say(f"filter matrix: {np.dtype(dt).name}")
        for op, lit in CASES:
            got = _rows(qm, "fm", [FilterOp(c, op, lit)])
            want = np.nonzero(_expect(op, cols[c], lit, nulls[c]))[0]
            assert np.array_equal(got, want), (np.dtype(dt).name, op, lit, len(got), len(want))
            assert np.array_equal(got, O.filter_rows(ot, [FilterOp(c, op, lit)]))


def test_int64_bigint_literals_on_device(qm):
    rng = np.random.default_rng(3)
    x = rng.integers(-(2 ** 62), 2 ** 62, N).astype(np.int64)
    lits = [BigInt(2 ** 62), BigInt(-(2 ** 61)), BigInt(5), BigInt(2 ** 53 + 1)]
    for j, lit in enumerate(lits):
        x[j * 100:(j + 1) * 100] = int(lit)
        x[(j + 4) * 100:(j + 5) * 100] = int(lit) + 1
    qm.register_table("bi", {"x": x, "rid": np.arange(N, dtype=np.int64)})
    for lit in lits:
        for op in ("lt", "gte", "eq", "neq", "lte", "gt"):
            got = _rows(qm, "bi", [FilterOp("x", op, lit)])
            want = np.nonzero(_expect(op, x, int(lit), np.zeros(N, bool)))[0]
            assert np.array_equal(got, want), (op, int(lit))


def test_int32_literal_coercion_on_device(qm):
    rng = np.random.default_rng(7)
    specials = np.array(sorted({to_int32(l) + d for l in INT32_LITERALS for d in (-1, 0, 1)}), dtype=np.int64)
    x = np.concatenate([rng.integers(-(2 ** 31), 2 ** 31, N - len(specials) * 10), np.repeat(specials, 10)])
    x = np.clip(x, -(2 ** 31), 2 ** 31 - 1).astype(np.int32)
    qm.register_table("i32", {"x": x, "rid": np.arange(len(x), dtype=np.int64)})
    z = np.zeros(len(x), bool)
    for lit in INT32_LITERALS:
        for op in ("eq", "neq", "lt", "lte", "gt", "gte", "between"):
            if op == "between":
                f, want = FilterOp("x", op, [-lit, lit]), (x >= to_int32(-lit)) & (x <= to_int32(lit))
            else:
                f, want = FilterOp("x", op, lit), _expect(op, x, to_int32(lit), z)
            got = _rows(qm, "i32", [f])
            assert np.array_equal(got, np.nonzero(want)[0]), (op, lit)


def test_and_or_groups_on_device(qm):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 100, N).astype(np.int32)
    b = rng.normal(0, 1, N)
    c = rng.integers(-5, 5, N).astype(np.int64)
    d = rng.integers(-3, 3, N).astype(np.int16)
    e = rng.normal(0, 1, N).astype(np.float32)
    qm.register_table("ao", {"a": a, "b": b, "c": c, "d": d, "e": e, "rid": np.arange(N, dtype=np.int64)})
    got = _rows(qm, "ao", [FilterOp("a", "gte", 10)],
                [[FilterOp("b", "gt", 0.5), FilterOp("c", "lt", 0)], [FilterOp("a", "in", [11, 12, 13])],
                 [FilterOp("c", "eq", 4)], [FilterOp("d", "eq", -2), FilterOp("e", "lt", -0.25)]])
    m = (a >= 10) & (((b > 0.5) & (c < 0)) | np.isin(a, [11, 12, 13]) | (c == 4) | ((d == -2) & (e < np.float32(-0.25))))
    assert np.array_equal(got, np.nonzero(m)[0])


def test_sql_pushed_clauses_with_nan_and_fractions_on_device(qm):
    """SQL clauses the reference pushes to the filter path (a top-level OR of
    AND trees, NOT BETWEEN, NOT IN), over a float column holding NaN and an
    int32 column compared with fractional literals: device rows equal the
    oracle's filter (src/sql/compiler.ts:90-109, src/decode.ts:477-578). The
    row-by-row shapes (NOT over a comparison, AND over OR) are refused before
    anything runs."""
    from querymode_amd.errors import QueryModeError
    from querymode_amd.sql import sql_to_descriptor

    rng = np.random.default_rng(13)
    x = rng.normal(0, 1, N)
    x[rng.random(N) < 0.1] = np.nan
    i = rng.integers(-4, 5, N).astype(np.int32)
    nulls = {"x": rng.random(N) < 0.03, "i": rng.random(N) < 0.03}
    cols = {"x": x, "i": i, "rid": np.arange(N, dtype=np.int64)}
    qm.register_table("sq", cols, nulls)
    ot = O.Table(cols, nulls)
    for where in ["x > 0.5 OR i >= 2.5", "x NOT BETWEEN -0.5 AND 0.5", "i NOT IN (1.5, 2, -3)",
                  "x < -1 AND i <= -2.5 OR x >= 1 AND i > 1.5 OR x IS NULL", "NOT x BETWEEN -1 AND 1 OR i = 2.5",
                  "i BETWEEN -2.5 AND 2.5 AND x <> 0.25"]:
        _say(f"sql on device: {where}")
        d = sql_to_descriptor("SELECT rid FROM sq WHERE " + where)
        got = _rows(qm, "sq", d.filters, d.filterGroups)
        assert np.array_equal(got, O.filter_rows(ot, d.filters, d.filterGroups)), where
    for where in ["NOT (x < 0.5)", "NOT (i >= 2.5)", "x > 0 AND (i = 1 OR x < -1)"]:
        with pytest.raises(QueryModeError) as e:
            sql_to_descriptor("SELECT rid FROM sq WHERE " + where)
        assert e.value.code == "UNSUPPORTED"
