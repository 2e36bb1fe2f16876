"""The Node-API addon (qm_napi.node) and its JavaScript host
(querymode_amd/napi/querymode_gpu.js): exports, error codes without a device,
and the JS plan encoder / QMCB decoder against the Python ones byte for byte."""
from __future__ import annotations

import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from querymode_amd import qmcb as Q
from querymode_amd.plan import AggregateOp, BigInt, FilterOp, JoinDescriptor, QueryDescriptor, encode_plan

ROOT = Path(__file__).resolve().parent.parent
JS = ROOT / "querymode_amd" / "napi" / "querymode_gpu.js"
ADDON = ROOT / "querymode_amd" / "lib" / "qm_napi.node"
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not ADDON.exists(), reason="node or the addon is not available")

EXPORTS = ["init", "shutdown", "lastError", "deviceInfo", "tableCreate", "tableDrop", "tableRows", "planCompile",
           "planClose", "exec", "batchExec", "batchSubmit", "planDeviceBytes", "resultRows", "resultQmcb", "resultQmcbRange",
           "resultStats",
           "resultClose"]


def _node(code: str) -> str:
    res = subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    return res.stdout


def test_addon_exports():
    out = _node(f"const q=require({json.dumps(str(JS))}); console.log(JSON.stringify(Object.keys(q.addon)))")
    assert json.loads(out) == EXPORTS


def test_init_without_device_throws_code():
    out = _node(f"""const q=require({json.dumps(str(JS))});
      try {{ const c=q.addon.init(0,0); q.addon.shutdown(c); console.log('device'); }}
      catch (e) {{ console.log(JSON.stringify({{code: e.code, msg: e.message}})); }}""")
    if out.strip() == "device":
        pytest.skip("device present")
    e = json.loads(out)
    assert e["code"] == "QUERY_FAILED" and e["msg"]


def _revive_py(v):
    if isinstance(v, dict) and "$bigint" in v:
        return BigInt(int(v["$bigint"]))
    if isinstance(v, list):
        return [_revive_py(x) for x in v]
    return v


CASES = [
    {"table": "t", "filters": [["a", "lt", 300000], ["b", "gt", 90], ["flag", "eq", 1]], "groupBy": ["key"],
     "aggregates": [["sum", "b", None], ["count", "*", None], ["min", "a", None], ["max", "a", None], ["avg", "b", None]]},
    {"table": "t", "filters": [["k", "eq", {"$bigint": "9007199254740993"}], ["x", "between", [-2.5, 7]],
                               ["y", "in", [1, 2, {"$bigint": "-5"}]], ["z", "is_null", None]],
     "filterGroups": [[["a", "gte", 0.5]], [["b", "not_in", [3]], ["c", "not_between", [0, 1]]]],
     "sortColumn": "x", "sortDirection": "desc", "limit": 10, "offset": 3, "projections": ["x", "y"]},
    {"table": "probe", "join": {"right": {"table": "build", "filters": [["region", "lt", 32]]}, "leftKey": "fk",
                                "rightKey": "id"},
     "computedColumns": [["aw", "amount", "weight"]], "groupBy": ["region"],
     "aggregates": [["sum", "aw", "total"], ["count", "*", None]]},
    {"table": "t", "groupBy": ["k"], "aggregates": [["count", "*", "n"], ["max", "v", None]],
     "having": ["or", ["and", ["cmp", "gt", "n", 5], ["not", ["isnull", "max_v", True]]],
                ["in", "max_v", [1.5, None, {"$bigint": "7"}], False]],
     "orderBy": [["n", "desc"], ["k", "asc"]], "limit": 4, "projections": ["k", "n"]},
    {"table": "t", "groupBy": ["k"], "aggregates": [["sum", "a", None], ["sum", "b", None]],
     "having": ["and", ["cmp_col", "ge", "sum_a", "sum_b"], ["between", "sum_a", -2.5, {"$bigint": "9"}, True]]},
    {"table": "t", "projections": ["a", "b"], "distinct": ["a"], "orderBy": [["b", "asc"], ["a", "desc"]]},
    {"table": "t", "projections": ["a"], "distinct": [], "sortColumn": "a", "limit": 3},
]


def _py_desc(c) -> QueryDescriptor:
    f = lambda l: [FilterOp(a, op, _revive_py(v)) for a, op, v in l]  # noqa: E731
    d = QueryDescriptor(table=c["table"], filters=f(c.get("filters", [])),
                        filterGroups=[f(g) for g in c.get("filterGroups", [])],
                        groupBy=c.get("groupBy", []), aggregates=[AggregateOp(*a) for a in c.get("aggregates", [])],
                        sortColumn=c.get("sortColumn"), sortDirection=c.get("sortDirection"), limit=c.get("limit"),
                        offset=c.get("offset"), projections=c.get("projections", []),
                        computedColumns=[tuple(x) for x in c.get("computedColumns", [])],
                        distinct=c.get("distinct"), orderBy=[tuple(o) for o in c.get("orderBy", [])],
                        having=_revive_py(c["having"]) if "having" in c else None)
    if "join" in c:
        j = c["join"]
        d.join = JoinDescriptor(QueryDescriptor(table=j["right"]["table"], filters=f(j["right"].get("filters", []))),
                                j["leftKey"], j["rightKey"])
    return d


@pytest.mark.parametrize("case", CASES)
def test_js_plan_encoding_matches_python(case):
    js = f"""const q=require({json.dumps(str(JS))});
      const rev=(v)=>(v&&typeof v==='object'&&'$bigint' in v)?BigInt(v['$bigint']):(Array.isArray(v)?v.map(rev):v);
      const c={json.dumps(case)};
      const F=(l)=>(l||[]).map(([column,op,value])=>({{column,op,value:rev(value)}}));
      const d={{table:c.table, filters:F(c.filters), filterGroups:(c.filterGroups||[]).map(F), groupBy:c.groupBy,
        aggregates:(c.aggregates||[]).map(([fn,column,alias])=>({{fn,column,alias}})), sortColumn:c.sortColumn,
        sortDirection:c.sortDirection, limit:c.limit, offset:c.offset, projections:c.projections,
        computedColumns:c.computedColumns, distinct:c.distinct, orderBy:c.orderBy,
        having:c.having===undefined?undefined:rev(c.having)}};
      if (c.join) d.join={{right:{{table:c.join.right.table, filters:F(c.join.right.filters)}}, leftKey:c.join.leftKey,
        rightKey:c.join.rightKey, type:'inner'}};
      console.log(q.encodePlan(d).toString('hex'));"""
    assert _node(js).strip() == encode_plan(_py_desc(case)).hex()


def test_js_qmcb_decode_matches_python():
    n = 37
    rng = np.random.default_rng(1)
    cols = [Q.ColumnarColumn("k", Q.DTYPE_I64, rng.integers(-2 ** 62, 2 ** 62, n), n),
            Q.ColumnarColumn("s", Q.DTYPE_F64, rng.normal(0, 1, n), n, Q.pack_bits(rng.random(n) < 0.3)),
            Q.ColumnarColumn("c", Q.DTYPE_I32, rng.integers(-9, 9, n).astype(np.int32), n)]
    buf = Q.encode(Q.ColumnarBatch(cols, n))
    js = f"""const q=require({json.dumps(str(JS))});
      const b=q.decodeQMCB(Buffer.from('{buf.hex()}','hex'));
      console.log(JSON.stringify(q.qmcbToRows(b), (k,v)=>typeof v==='bigint'?v.toString():v));"""
    got = json.loads(_node(js))
    want = Q.to_rows(Q.decode(buf))
    assert len(got) == n
    for g, w in zip(got, want):
        assert int(g["k"]) == w["k"] and g["c"] == w["c"]
        assert (g["s"] is None and w["s"] is None) or g["s"] == w["s"]
