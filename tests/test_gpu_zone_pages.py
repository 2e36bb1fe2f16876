"""Zone maps inside the page span and with OR groups (src/decode.ts:9-100
canSkipPage, per page): pages no selected row can lie on are skipped by the
scan, selection and top-K kernels tile by tile, wherever they are. Every
result equals the oracle's and the same query with zone maps off (row outputs
byte for byte); pagesSkipped equals a numpy restatement of the page test.
A Parquet file decoded on the host and uploaded gives the oracle's results."""
from __future__ import annotations

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from oracle import oracle as O
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu

Z = 65536
N = 3_000_001  # 46 pages, the last one partial


def _table():
    rng = np.random.default_rng(71)
    k = np.sort(rng.integers(0, 1_000_000_000, N)).astype(np.int64)   # pages are disjoint key ranges
    g = rng.integers(0, 7, N).astype(np.int32)
    v = rng.integers(-1000, 1000, N).astype(np.int32)
    t = rng.random(N)
    nt = np.zeros(N, bool)
    nt[10 * Z: 14 * Z] = True          # pages 10..13 of t: all NULL
    return {"k": k, "g": g, "v": v, "t": t}, {"t": nt}


def _can(cols, nulls, groups):
    """Per page: can a row pass AND-group 0 and (if any) one OR group? groups:
    [[(col, lo, hi), ...] (AND), [...], ...] with group 0 the top-level AND."""
    pages = (N + Z - 1) // Z
    can = np.ones(pages, bool)

    def meets(col, lo, hi, p):
        a, m = cols[col][p * Z:(p + 1) * Z], nulls.get(col, np.zeros(N, bool))[p * Z:(p + 1) * Z]
        vals = a[~m]
        return vals.size > 0 and vals.max() >= lo and vals.min() <= hi

    for p in range(pages):
        ok = all(meets(c, lo, hi, p) for c, lo, hi in groups[0])
        if ok and len(groups) > 1:
            ok = any(all(meets(c, lo, hi, p) for c, lo, hi in g) for g in groups[1:])
        can[p] = ok
    return can


def _run(qm, d, cols, nulls):
    dev, ex = qm.device, qm.executor
    try:
        dev.tune_set("zone_maps", 1)
        on = ex.execute(d)
        dev.tune_set("zone_maps", 0)
        off = ex.execute(d)
    finally:
        dev.tune_set("zone_maps", 1)
    return on, off, O.run(d, {"zp": O.Table(cols, nulls)})


def test_or_groups_skip_interior_pages(qm):
    cols, nulls = _table()
    qm.register_table("zp", cols, nulls)
    k = cols["k"]
    a1, b1 = int(k[3 * Z + 10]), int(k[5 * Z + 70])
    a2, b2 = int(k[30 * Z + 5]), int(k[31 * Z + 9])
    groups = [[FilterOp("k", "between", [a1, b1])], [FilterOp("k", "between", [a2, b2])]]
    can = _can(cols, nulls, [[], [("k", a1, b1)], [("k", a2, b2)]])
    want_skip = int((~can).sum())
    assert want_skip >= 40
    rows = QueryDescriptor("zp", filterGroups=groups, projections=["k", "v"])
    on, off, want = _run(qm, rows, cols, nulls)
    assert bytes(on.columnarData) == bytes(off.columnarData)
    assert_rows_equal(on.rows, want, ctx="or rows")
    assert on.pagesSkipped == want_skip and off.pagesSkipped == 0
    top = QueryDescriptor("zp", filterGroups=groups, projections=["k", "v"], sortColumn="v", sortDirection="desc",
                          limit=50)
    on, off, want = _run(qm, top, cols, nulls)
    assert bytes(on.columnarData) == bytes(off.columnarData)
    assert_rows_equal(on.rows, want, ctx="or top-k")
    assert on.pagesSkipped == want_skip
    agg = QueryDescriptor("zp", filterGroups=groups, groupBy=["g"],
                          aggregates=[AggregateOp("count", "*"), AggregateOp("sum", "v"), AggregateOp("max", "k")])
    on, off, want = _run(qm, agg, cols, nulls)
    assert_groups_equal(on.rows, off.rows, ["g"], ctx="or agg on/off")
    assert_groups_equal(on.rows, want, ["g"], ctx="or agg")
    assert on.pagesSkipped == want_skip


def test_all_null_pages_inside_the_span(qm):
    cols, nulls = _table()
    qm.register_table("zp", cols, nulls)
    k = cols["k"]
    hi = int(k[20 * Z])
    d = QueryDescriptor("zp", filters=[FilterOp("t", "is_not_null"), FilterOp("k", "lt", hi)],
                        projections=["k", "t"])
    can = _can(cols, nulls, [[("t", -np.inf, np.inf), ("k", -(2 ** 63), hi - 1)]])
    assert not can[10:14].any() and can[9] and can[14]      # interior: pages 10..13
    on, off, want = _run(qm, d, cols, nulls)
    assert bytes(on.columnarData) == bytes(off.columnarData)
    assert_rows_equal(on.rows, want, ctx="null pages")
    assert on.pagesSkipped == int((~can).sum())
    # an ORDER BY over the same selection
    d2 = QueryDescriptor("zp", filters=d.filters, projections=["k", "t"], sortColumn="t", sortDirection="asc",
                         limit=25)
    on, off, want = _run(qm, d2, cols, nulls)
    assert bytes(on.columnarData) == bytes(off.columnarData)
    assert_rows_equal(on.rows, want, ctx="null pages top-k")


def test_parquet_upload(qm, tmp_path):
    rng = np.random.default_rng(72)
    n = 300_000
    data = {"id": np.arange(n, dtype=np.int64), "cat": rng.integers(0, 16, n).astype(np.int32),
            "amount": rng.random(n) * 1000, "flag": rng.random(n) < 0.5}
    mask = rng.random(n) < 0.07
    path = tmp_path / "orders.parquet"
    pq.write_table(pa.table({"id": data["id"], "cat": data["cat"], "amount": pa.array(data["amount"], mask=mask),
                             "flag": data["flag"]}), path, row_group_size=50_000)
    qm.register_parquet("pq", str(path))
    ot = O.Table(data, {"amount": mask})
    d = QueryDescriptor("pq", filters=[FilterOp("amount", "gt", 250.0)], groupBy=["cat"],
                        aggregates=[AggregateOp("sum", "amount"), AggregateOp("count", "amount"),
                                    AggregateOp("max", "id")])
    assert_groups_equal(qm.executor.execute(d).rows, O.run(d, {"pq": ot}), ["cat"], float_cols=("sum_amount",),
                        ctx="parquet agg")
    d2 = QueryDescriptor("pq", filters=[FilterOp("cat", "eq", 3)], projections=["id", "amount"],
                         sortColumn="amount", sortDirection="desc", limit=20)
    assert_rows_equal(qm.executor.execute(d2).rows, O.run(d2, {"pq": ot}), ctx="parquet top-k")


def test_batch_skips_pages_no_query_can_match(qm):
    """A batch whose range filters all fall on a clustered column: the pages
    no query of the batch can match are not read by the shared-interval pass
    (GROUP BY a small key) nor by the partitioned-runs row scatter (GROUP BY a
    key in any row order). Results equal the same batch with zone maps off,
    byte for byte, and the oracle; pagesSkipped counts those pages."""
    from querymode_amd import datagen

    n = 3_000_001
    cols = datagen.batch_table(n=n, seed=81)
    cols["v2"] = np.sort(cols["v2"])                       # clustered: pages are disjoint value ranges
    cols["k2"] = datagen.uniform_int(81, 72, 0, n, 0, 60_000, np.int64)
    qm.register_table("zb", cols)
    ranges = [(0.20, 0.26), (0.21, 0.22), (0.60, 0.63), (0.605, 0.61), (0.90, 0.9005)]
    aggs = [AggregateOp("sum", "v1"), AggregateOp("count", "*"), AggregateOp("min", "v1"), AggregateOp("max", "v2")]
    descs = [QueryDescriptor("zb", filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", hi)], groupBy=[key],
                             aggregates=aggs) for lo, hi in ranges for key in ("k1", "k2")]
    v = cols["v2"]
    pages = (n + Z - 1) // Z
    can = np.zeros(pages, bool)
    for p in range(pages):
        a = v[p * Z:(p + 1) * Z]
        can[p] = any(a.max() >= lo and a.min() < hi for lo, hi in ranges)
    want_skip = int((~can).sum())
    assert want_skip >= 30
    ex, dev = qm.executor, qm.device
    try:
        dev.tune_set("zone_maps", 1)
        on = ex.execute_batch(descs)
        dev.tune_set("zone_maps", 0)
        off = ex.execute_batch(descs)
    finally:
        dev.tune_set("zone_maps", 1)
    assert {r.stats["strategy"] for r in on} == {"shared_interval", "part_runs"}, [r.stats["strategy"] for r in on]
    ot = O.Table(cols)
    for i, (d, a, b) in enumerate(zip(descs, on, off)):
        assert bytes(a.columnarData) == bytes(b.columnarData), i
        assert a.pagesSkipped == want_skip and b.pagesSkipped == 0, (i, a.pagesSkipped)
        if i < 4:
            assert_groups_equal(a.rows, O.run(d, {"zb": ot}), d.groupBy, ctx=f"zone batch q{i}")
    # one query with no filter: every page may hold its rows, nothing is skipped
    more = ex.execute_batch(descs[:2] + [QueryDescriptor("zb", groupBy=["k1"], aggregates=aggs)])
    assert all(r.pagesSkipped == 0 for r in more)
    assert bytes(more[0].columnarData) == bytes(on[0].columnarData)
