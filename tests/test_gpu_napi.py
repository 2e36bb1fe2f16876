"""End to end from JavaScript: the Node-API addon and querymode_gpu.js drive
the device path (register -> DataFrame builder -> exec -> QMCB -> rows) on the
reference's generateRows tables (src/operators-conformance.test.ts:101-113),
checked against the same closed forms as the oracle tests."""
from __future__ import annotations

import json
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
JS = ROOT / "querymode_amd" / "napi" / "querymode_gpu.js"
NODE = shutil.which("node")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(NODE is None or not (ROOT / "querymode_amd" / "lib" / "qm_napi.node").exists(),
                                 reason="node or the addon is not available")]

SCRIPT = r"""
const q = require(%s);
const N = 1000000;
const id = new Int32Array(N), region = new Int32Array(N), amount = new Int32Array(N);
for (let i = 0; i < N; i++) { id[i] = i; region[i] = i %% 4; amount[i] = (i * 7 + 13) %% 1000; }
const R = 500000;
const rid = new Int32Array(R), value = new Int32Array(R);
for (let i = 0; i < R; i++) { rid[i] = i; value[i] = i * 10; }
const qm = q.QueryMode.gpu(0);
qm.registerTable('rows', { id, region, amount });
qm.registerTable('right', { id: rid, value });
const out = {};
out.group = qm.table('rows').groupBy('region').aggregate('sum', 'amount', 'sum_amount')
  .aggregate('avg', 'amount', 'avg_amount').aggregate('min', 'amount', 'min_amount')
  .aggregate('max', 'amount', 'max_amount').aggregate('count', 'amount', 'cnt').exec().rows;
out.topk = qm.table('rows').sort('amount', 'desc').limit(100).select('id', 'amount').exec().rows;
out.join = qm.table('rows').join(qm.table('right'), { left: 'id', right: 'id' })
  .aggregate('count', '*').aggregate('sum', 'value').exec().rows;
(async () => {
  const op = qm.table('rows').filter('amount', 'eq', 999).select('id').toOperator(1000);
  out.pulled = (await q.drainPipeline(op)).length;
  // a batch through the Promise-returning batchSubmit equals each query run alone
  const frames = [0, 250, 500, 750].map((lo) => qm.table('rows').filter('amount', 'gte', lo)
    .filter('amount', 'lt', lo + 300).groupBy('region').aggregate('sum', 'amount').aggregate('count', '*'));
  const pending = qm.executor.executeBatch(frames.map((f) => f.desc));
  out.batch = (await pending).map((r) => r.rows);
  out.solo = frames.map((f) => f.exec().rows);
  // cursor: the top-K result stays on the device, 7 rows per pull
  const topDesc = qm.table('rows').sort('amount', 'desc').limit(100).select('id', 'amount').desc;
  out.cursor = [];
  out.cursorBatches = 0;
  for await (const b of qm.executor.cursor(topDesc, 7)) { out.cursor.push(...b); out.cursorBatches++; }
  qm.close();
  console.log(JSON.stringify(out, (k, v) => typeof v === 'bigint' ? v.toString() : v));
})().catch((e) => { console.error(e); process.exit(1); });
"""


def test_js_host_end_to_end():
    print("node: generateRows(1M) group-by / top-K / join / batch through qm_napi.node", flush=True)
    res = subprocess.run([NODE, "-e", SCRIPT % json.dumps(str(JS))], capture_output=True, text=True, timeout=120,
                         cwd=ROOT)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert len(out["group"]) == 4
    for x in out["group"]:
        c = (3 * x["region"] + 1) % 4
        per = 250 * c + 4 * (249 * 250 // 2)          # one 1000-row period, 1000 periods
        assert int(x["sum_amount"]) == 1000 * per
        assert x["avg_amount"] == c + 498
        assert int(x["min_amount"]) == c and int(x["max_amount"]) == 996 + c
        assert int(x["cnt"]) == 250_000
    assert [int(r["amount"]) for r in out["topk"]] == [999] * 100
    assert [int(r["id"]) for r in out["topk"]] == [998 + 1000 * k for k in range(100)]
    assert [(int(r["count_*"]), int(r["sum_value"])) for r in out["join"]] == [(500_000, 10 * 499_999 * 500_000 // 2)]
    assert out["pulled"] == 1000
    key = lambda r: int(r["region"])  # noqa: E731
    assert [sorted(b, key=key) for b in out["batch"]] == [sorted(x, key=key) for x in out["solo"]]
    assert all(len(b) == 4 for b in out["batch"])
    assert out["cursor"] == out["topk"] and out["cursorBatches"] == 15


SCRIPT_CLOSE = r"""
const q = require(%s);
const N = 2000000;
const region = new Int32Array(N), amount = new Int32Array(N);
for (let i = 0; i < N; i++) { region[i] = i %% 4; amount[i] = (i * 7 + 13) %% 1000; }
const qm = q.QueryMode.gpu(0);
qm.registerTable('rows', { region, amount });
const frames = [0, 250, 500, 750].map((lo) => qm.table('rows').filter('amount', 'gte', lo)
  .filter('amount', 'lt', lo + 300).groupBy('region').aggregate('sum', 'amount').aggregate('count', '*'));
(async () => {
  const pending = qm.executor.executeBatch(frames.map((f) => f.desc));
  // while the batch runs: the table is registered again under its name (the old one is
  // dropped) and the executor is closed; both take effect once the batch settles
  qm.registerTable('rows', { region: region.slice(0, 1000), amount: amount.slice(0, 1000) });
  qm.close();
  const got = (await pending).map((r) => r.rows);
  console.log(JSON.stringify(got, (k, v) => typeof v === 'bigint' ? v.toString() : v));
})().catch((e) => { console.error(e); process.exit(1); });
"""


def test_js_close_while_batch_runs():
    """Dropping the table, re-registering its name and closing the executor
    while executeBatch is pending must not free what the running batch uses:
    the batch's results are whole, and the context closes after them."""
    print("node: close while a batch runs", flush=True)
    res = subprocess.run([NODE, "-e", SCRIPT_CLOSE % json.dumps(str(JS))], capture_output=True, text=True,
                         timeout=120, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    got = json.loads(res.stdout.strip().splitlines()[-1])
    assert len(got) == 4
    for lo, rows in zip([0, 250, 500, 750], got):
        assert len(rows) == 4
        total = sum(int(r["count_*"]) for r in rows)
        # amount = (7 i + 13) mod 1000 takes each value 2000 times over 2M rows
        assert total == 2000 * len(range(lo, min(lo + 300, 1000)))
