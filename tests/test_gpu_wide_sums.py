"""Float SUM / AVG over columns whose values span a wide exponent range: every
group's sum must be within 1e-12 relative of math.fsum (exactly rounded),
whatever the other groups of the column hold (src/partial-agg.ts:62-80; the
north star's FP64 tolerance). Groups of ~1e20 and ~1e-20 values share one
column; {1e300, -1e300, 1} cancels to 1; normal x 10^U[-3, 6] and
normal x 10^U[-300, 300] values; a group that cancels to exactly 0."""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import REL_TOL

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _wide_table(seed=0):
    rng = np.random.default_rng(seed)
    parts = [
        (0, rng.normal(0, 1, 20_000) * 1e20),
        (1, rng.normal(0, 1, 20_000) * 1e-20),
        (2, np.array([1e300, -1e300, 1.0])),
        (3, rng.normal(0, 1, 50_000) * 10.0 ** rng.uniform(-3, 6, 50_000)),
        (4, rng.normal(0, 1, 5_000) * 10.0 ** rng.uniform(-300, 300, 5_000)),
        (5, np.array([1e300, 1e-300, -1e300, -1e-300])),           # exactly 0
        (6, np.concatenate([np.full(1000, 0.1), np.full(999, -0.1)])),
        (7, np.array([2.0 ** 53, 1.0, -(2.0 ** 53), 2.0 ** -1074])),
    ]
    g = np.concatenate([np.full(len(v), k, dtype=np.int32) for k, v in parts])
    v = np.concatenate([p[1] for p in parts])
    perm = rng.permutation(len(v))  # groups interleaved in row order
    return {"g": g[perm], "v": v[perm].astype(np.float64), "f": rng.normal(0, 1, len(v)).astype(np.float64)}, parts


def _check(rows, parts, key="g"):
    by = {r[key]: r for r in rows}
    assert len(by) == len(parts)
    for k, vals in parts:
        r = by[k]
        s = math.fsum(vals.tolist())
        assert r["n"] == len(vals)
        assert math.isclose(r["s"], s, rel_tol=REL_TOL, abs_tol=0.0) or r["s"] == s, (k, r["s"], s)
        a = s / len(vals)
        assert math.isclose(r["a"], a, rel_tol=REL_TOL, abs_tol=0.0) or r["a"] == a, (k, r["a"], a)


def test_wide_range_sums_per_group(qm):
    _say("wide-range float sums: 8 groups in one column")
    cols, parts = _wide_table()
    qm.register_table("w", cols)
    d = QueryDescriptor("w", groupBy=["g"], aggregates=[AggregateOp("sum", "v", "s"), AggregateOp("avg", "v", "a"),
                                                         AggregateOp("count", "*", "n")])
    _check(qm.executor.execute(d).rows, parts)
    # through a batch (the fused passes leave wide sums to the per-query path) and with a filter
    got = qm.executor.execute_batch([d, QueryDescriptor("w", filters=[FilterOp("f", "gt", -100.0)], groupBy=["g"],
                                                        aggregates=d.aggregates)])
    _check(got[0].rows, parts)
    _check(got[1].rows, parts)


def test_wide_range_sums_many_groups(qm):
    """A hashed key of 3000 groups, each with its own magnitude and some exact cancellation."""
    _say("wide-range float sums: 3000 hashed groups")
    rng = np.random.default_rng(5)
    n = 300_000
    key = rng.integers(0, 3000, n).astype(np.int64) * 1_000_003  # sparse ids: hash table
    mag = 10.0 ** rng.uniform(-250, 250, 3000)
    v = rng.normal(0, 1, n) * mag[key // 1_000_003]
    v[::7] = -v[::7]  # pairs of opposite values in some groups
    qm.register_table("wh", {"k": key, "v": v})
    rows = qm.executor.execute(QueryDescriptor("wh", groupBy=["k"], aggregates=[
        AggregateOp("sum", "v", "s"), AggregateOp("avg", "v", "a"), AggregateOp("count", "*", "n")])).rows
    vals = {}
    for k_, x in zip(key.tolist(), v.tolist()):
        vals.setdefault(k_, []).append(x)
    _check([dict(r, g=r["k"]) for r in rows], [(k_, np.array(xs)) for k_, xs in vals.items()])


def test_wide_range_f32_and_product(qm):
    _say("wide-range float sums: f32 column and SUM(a * b)")
    rng = np.random.default_rng(9)
    n = 40_000
    g = rng.integers(0, 8, n).astype(np.int32)
    a = (rng.normal(0, 1, n) * 10.0 ** (rng.uniform(-30, 30, n))).astype(np.float32)
    b = rng.normal(0, 1, n) * 10.0 ** rng.uniform(-100, 100, n)
    qm.register_table("wf", {"g": g, "a": a, "b": b})
    rows = qm.executor.execute(QueryDescriptor("wf", groupBy=["g"], computedColumns=[("ab", "a", "b")], aggregates=[
        AggregateOp("sum", "a", "s"), AggregateOp("avg", "a", "a"), AggregateOp("count", "*", "n"),
        AggregateOp("sum", "ab", "p")])).rows
    parts = [(k, a[g == k].astype(np.float64)) for k in range(8)]
    _check(rows, parts)
    for r in rows:
        m = g == r["g"]
        s = math.fsum((a[m].astype(np.float64) * b[m]).tolist())
        assert math.isclose(r["p"], s, rel_tol=REL_TOL, abs_tol=0.0) or r["p"] == s, (r["g"], r["p"], s)
