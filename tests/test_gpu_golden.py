"""The device path against the reference's own known answers (the same
vectors the oracle is pinned on in test_golden_pylance.py and
test_oracle_known_answers.py), through the C-ABI."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from querymode_amd.plan import AggregateOp, JoinDescriptor, QueryDescriptor
from querymode_amd.sql import sql_to_descriptor

from .test_golden_pylance import ENGINE_ERRORS, OFF_PATH, ROW_BY_ROW, STRING_COLUMN

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "pylance_filters.json").read_text())


def _golden_table(qm):
    t = GOLDEN["table"]
    cols, nulls = {}, {}
    for name, dt in (("a", np.int64), ("b", np.int64), ("d", np.float64)):
        nulls[name] = np.array([v is None for v in t[name]])
        cols[name] = np.array([0 if v is None else v for v in t[name]], dtype=dt)
    cols["rid"] = np.arange(len(t["a"]), dtype=np.int64)
    qm.register_table("t", cols, nulls)


def test_pylance_vectors_on_device(qm):
    _golden_table(qm)
    ran = 0
    for vec in GOLDEN["vectors"]:
        if vec["name"] in OFF_PATH or vec["name"] in STRING_COLUMN or vec["name"] in ROW_BY_ROW:
            continue
        print(f"vector {vec['name']}: {vec['where']}", flush=True)
        d = sql_to_descriptor("SELECT rid FROM t WHERE " + vec["where"])
        got = [r["rid"] for r in qm.executor.execute(d).rows]
        want = ENGINE_ERRORS.get(vec["name"], vec.get("rows"))
        assert got == want, f"{vec['name']} ({vec['where']}): rows {got} vs fixture {want}"
        ran += 1
    assert ran == 72 - len(OFF_PATH) - len(STRING_COLUMN) - len(ROW_BY_ROW)


def _generate_rows(n: int) -> dict:
    i = np.arange(n, dtype=np.int64)
    return {"id": i.astype(np.int32), "region": (i % 4).astype(np.int32), "category": (i % 3).astype(np.int32),
            "amount": ((i * 7 + 13) % 1000).astype(np.int32), "score": ((i * 31) % 997) / 10.0}


def test_groupby_region_5m_closed_form(qm):
    # src/operators-conformance.test.ts:408-447; closed form as in the oracle test
    qm.register_table("g5", _generate_rows(5_000_000))
    r = qm.executor.execute(QueryDescriptor(table="g5", groupBy=["region"], aggregates=[
        AggregateOp("sum", "amount", "sum_amount"), AggregateOp("avg", "amount", "avg_amount"),
        AggregateOp("min", "amount", "min_amount"), AggregateOp("max", "amount", "max_amount"),
        AggregateOp("count", "amount", "cnt")])).rows
    assert sorted(x["region"] for x in r) == [0, 1, 2, 3]
    for x in r:
        c = (3 * x["region"] + 1) % 4
        per = 250 * c + 4 * (249 * 250 // 2)
        assert x["sum_amount"] == 5000 * per
        assert x["avg_amount"] == c + 498
        assert x["min_amount"] == c and x["max_amount"] == 996 + c
        assert x["cnt"] == 1_250_000


def test_topk_100_5m_all_999(qm):
    # src/operators-conformance.test.ts:576-594
    qm.register_table("g5", _generate_rows(5_000_000))
    r = qm.executor.execute(QueryDescriptor(table="g5", sortColumn="amount", sortDirection="desc", limit=100,
                                            projections=["id", "amount"])).rows
    assert [x["amount"] for x in r] == [999] * 100
    assert [x["id"] for x in r] == [998 + 1000 * k for k in range(100)]


def test_inner_join_1m_x_500k(qm):
    # src/operators-conformance.test.ts:194-218: 500,000 rows, value = id * 10
    qm.register_table("jl", _generate_rows(1_000_000))
    rid = np.arange(500_000, dtype=np.int32)
    qm.register_table("jr", {"id": rid, "value": rid * 10})
    j = JoinDescriptor(QueryDescriptor(table="jr"), "id", "id")
    agg = qm.executor.execute(QueryDescriptor(table="jl", join=j, aggregates=[
        AggregateOp("count", "*"), AggregateOp("sum", "value"), AggregateOp("min", "id"),
        AggregateOp("max", "id")])).rows
    assert agg == [{"count_*": 500_000, "sum_value": 10 * (499_999 * 500_000 // 2), "min_id": 0,
                    "max_id": 499_999}]
    rows = qm.executor.execute(QueryDescriptor(table="jl", join=j, projections=["id", "value"])).rows
    assert len(rows) == 500_000
    assert [x["id"] for x in rows] == list(range(500_000))
    assert all(x["value"] == 10 * x["id"] for x in rows[::997])


def test_partial_agg_vectors(qm):
    # src/partial-agg.test.ts:49-68, 126-177
    qm.register_table("pa", {"region": np.array([0, 1, 0, 1], np.int32), "amount": np.array([10, 20, 30, 5], np.int32)})
    r = qm.executor.execute(QueryDescriptor(table="pa", groupBy=["region"],
                                            aggregates=[AggregateOp("sum", "amount")])).rows
    assert sorted((x["region"], x["sum_amount"]) for x in r) == [(0, 40), (1, 25)]
    qm.register_table("nk", {"g": np.array([np.nan, np.nan, np.inf, -np.inf]), "v": np.array([1.0, 2, 3, 4])})
    r = qm.executor.execute(QueryDescriptor(table="nk", groupBy=["g"], aggregates=[AggregateOp("sum", "v")])).rows
    assert len(r) == 3
    by = {("nan" if math.isnan(x["g"]) else x["g"]): x["sum_v"] for x in r}
    assert by == {"nan": 3.0, math.inf: 3.0, -math.inf: 4.0}
