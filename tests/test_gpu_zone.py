"""Zone maps: a single aggregate query skips the 65,536-row pages its range
filters cannot match at either end of the table (the reference skips pages on
their min/max, src/decode.ts:9-100). Every result must equal the oracle's and
the same query with zone maps off; pagesSkipped must be the count a numpy
restatement of the page test gives."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu

Z = 65536
N = 3_000_001  # 46 pages, the last one partial


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _table():
    rng = np.random.default_rng(17)
    k = np.sort(rng.integers(0, 1_000_000_000, N)).astype(np.int64)   # sorted: pages are disjoint ranges
    t = np.sort(rng.random(N)) * 100.0                                # clustered float column
    g = rng.integers(0, 7, N).astype(np.int32)
    v = rng.integers(-(2 ** 31), 2 ** 31, N, dtype=np.int64).astype(np.int32)
    x = rng.random(N)                                                 # unclustered: nothing to skip
    cols = {"k": k, "t": t, "g": g, "v": v, "x": x}
    nm = np.zeros(N, bool)
    nm[5 * Z: 9 * Z] = True                                           # pages 5..8 of `t` are all NULL
    nm[20 * Z + 17] = True
    return cols, {"t": nm}


def _expected_skipped(cols, nulls, ranges):
    """Pages outside [first, last] of the pages every range can match."""
    pages = (N + Z - 1) // Z
    can = np.ones(pages, bool)
    for col, lo, hi in ranges:
        a = cols[col]
        nm = nulls.get(col, np.zeros(N, bool))
        for p in range(pages):
            seg, sn = a[p * Z:(p + 1) * Z], nm[p * Z:(p + 1) * Z]
            vals = seg[~sn]
            if vals.size == 0 or vals.max() < lo or vals.min() > hi:
                can[p] = False
    idx = np.nonzero(can)[0]
    if idx.size == 0:
        return pages - 1
    return int(pages - (idx[-1] - idx[0] + 1))


AGGS = [AggregateOp("count", "*"), AggregateOp("sum", "v"), AggregateOp("min", "v"), AggregateOp("max", "t"),
        AggregateOp("avg", "v")]


def _cases(cols):
    k = cols["k"]
    lo_mid, hi_mid = int(k[20 * Z + 5]), int(k[23 * Z + 100])
    return [
        # (descriptor, the inclusive (column, lo, hi) ranges the page test sees)
        (QueryDescriptor("z", filters=[FilterOp("k", "between", [lo_mid, hi_mid])], groupBy=["g"], aggregates=AGGS),
         [("k", lo_mid, hi_mid)]),
        (QueryDescriptor("z", filters=[FilterOp("k", "lt", int(k[Z + 3]))], aggregates=AGGS),
         [("k", -(2 ** 63), int(k[Z + 3]) - 1)]),
        (QueryDescriptor("z", filters=[FilterOp("k", "gte", int(k[N - 10]))], groupBy=["g"], aggregates=AGGS),
         [("k", int(k[N - 10]), 2 ** 63 - 1)]),
        (QueryDescriptor("z", filters=[FilterOp("k", "gt", 2_000_000_000)], groupBy=["g"], aggregates=AGGS),
         [("k", 2_000_000_001, 2 ** 63 - 1)]),                                        # nothing can match
        (QueryDescriptor("z", filters=[FilterOp("t", "gte", 40.0), FilterOp("t", "lt", 41.5)], groupBy=["g"],
                         aggregates=AGGS), [("t", 40.0, np.nextafter(41.5, 0.0))]),
        (QueryDescriptor("z", filters=[FilterOp("t", "lt", 12.0), FilterOp("x", "lt", 0.5)], groupBy=["g"],
                         aggregates=AGGS), [("t", -np.inf, np.nextafter(12.0, -1.0)), ("x", -np.inf, 0.5)]),
        (QueryDescriptor("z", filters=[FilterOp("k", "between", [lo_mid, hi_mid])], groupBy=["k"],
                         aggregates=AGGS[:4]), [("k", lo_mid, hi_mid)]),                   # sorted group key
        (QueryDescriptor("z", filters=[FilterOp("x", "lt", 0.001)], groupBy=["g"], aggregates=AGGS),
         [("x", -np.inf, 0.001)]),                                                     # unclustered: no skip
        (QueryDescriptor("z", filters=[FilterOp("t", "is_not_null"), FilterOp("k", "lt", int(k[8 * Z]))],
                         aggregates=AGGS), [("t", -np.inf, np.inf), ("k", -(2 ** 63), int(k[8 * Z]) - 1)]),
    ]


def test_zone_maps_skip_pages_and_keep_results(qm):
    _say(f"zone maps: {N} rows, {(N + Z - 1) // Z} pages")
    cols, nulls = _table()
    qm.register_table("z", cols, nulls)
    ot = O.Table(cols, nulls)
    dev = qm.device
    ex = qm.executor
    cases = _cases(cols)
    try:
        for i, (d, ranges) in enumerate(cases):
            _say(f"  case {i}")
            dev.tune_set("zone_maps", 1)
            got = ex.execute(d)
            dev.tune_set("zone_maps", 0)
            off = ex.execute(d)
            # the same groups (a hash table's slot order may differ between runs)
            assert_groups_equal(got.rows, off.rows, d.groupBy, ctx=f"zone on/off case {i}")
            assert off.pagesSkipped == 0
            want = O.run(d, {"z": ot})
            assert_groups_equal(got.rows, want, d.groupBy, ctx=f"zone case {i}")
            assert got.pagesSkipped == _expected_skipped(cols, nulls, ranges), (i, got.pagesSkipped)
            assert got.stats["rows_scanned"] <= N
    finally:
        dev.tune_set("zone_maps", 1)
    # the first cases do skip pages
    assert _expected_skipped(cols, nulls, cases[0][1]) == 42


def test_zone_maps_row_outputs(qm):
    """Top-K, ORDER BY and plain selections over the skipped-page view: the
    same rows in the same order (ties by row id) as the oracle and as the
    query with zone maps off."""
    _say("zone maps: row outputs")
    cols, nulls = _table()
    qm.register_table("z", cols, nulls)
    ot = O.Table(cols, nulls)
    k = cols["k"]
    lo, hi = int(k[20 * Z + 5]), int(k[23 * Z + 100])
    rng_k = [FilterOp("k", "between", [lo, hi])]
    descs = [
        (QueryDescriptor("z", filters=rng_k, projections=["k", "v"], sortColumn="v", sortDirection="desc", limit=100),
         42),
        (QueryDescriptor("z", filters=rng_k + [FilterOp("x", "lt", 0.01)], projections=["k", "x", "v"],
                         limit=50, offset=7), 42),
        (QueryDescriptor("z", filters=[FilterOp("t", "between", [40.0, 40.05])], projections=["t", "g"],
                         sortColumn="t", sortDirection="asc"), None),
        (QueryDescriptor("z", filters=[FilterOp("k", "lt", int(k[Z + 3]))], projections=["k", "v", "t"],
                         sortColumn="k", sortDirection="desc", limit=10, offset=2), 44),
    ]
    dev = qm.device
    ex = qm.executor
    try:
        for i, (d, skipped) in enumerate(descs):
            _say(f"  rows case {i}")
            dev.tune_set("zone_maps", 1)
            got = ex.execute(d)
            dev.tune_set("zone_maps", 0)
            off = ex.execute(d)
            assert bytes(got.columnarData) == bytes(off.columnarData), i
            want = O.run(d, {"z": ot})
            assert_rows_equal(got.rows, want, ctx=f"zone rows case {i}")
            if skipped is not None:
                assert got.pagesSkipped == skipped, (i, got.pagesSkipped)
            else:
                assert got.pagesSkipped > 0, i
    finally:
        dev.tune_set("zone_maps", 1)
