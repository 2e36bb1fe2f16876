"""Batched group-by over a key stored in any row order (the partitioned-runs
launch, k_pruns.hip): many queries, each a range filter on one column and a
GROUP BY on an integer key whose rows are spread over the whole table. Every
batched result must equal the oracle's and the same query run alone through
qm_exec, bit for bit (integer SUM / COUNT / MIN / MAX, float MIN / MAX)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd import datagen
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import assert_groups_equal

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _table(n, kdom, seed=0, klo=0, kdt=np.int64):
    cols = datagen.batch_table(n=n, seed=seed)  # generated row order
    cols["k2"] = datagen.uniform_int(seed, 72, 0, n, klo, kdom, np.int64).astype(kdt)
    return cols


def _mix(name, nq, seed, min_sel=1e-3):
    return [QueryDescriptor(name, filters=d.filters, groupBy=d.groupBy, aggregates=d.aggregates)
            for d in datagen.batch_queries(nq, seed=seed, min_sel=min_sel)]


def _check(qm, name, cols, descs, nulls=None):
    qm.register_table(name, cols, nulls)
    batch = qm.executor.execute_batch(descs)
    ot = O.Table(cols, nulls)
    for i, (d, got) in enumerate(zip(descs, batch)):
        want = O.run(d, {name: ot})
        assert_groups_equal(got.rows, want, d.groupBy, ctx=f"{name} batch q{i}")
    return batch


@pytest.mark.parametrize("kdom", [60_000, 1_000_000])
def test_unsorted_batch_matches_oracle_and_solo(qm, kdom):
    """96 queries (half GROUP BY k2 with k2 uniform in generated row order) over
    3M rows: ~50 rows per key (the 500M-row mix's ratio) and ~3 rows per key."""
    n = 3_000_000
    _say(f"unsorted batch: {n} rows, k2 domain {kdom}, 96 queries")
    cols = _table(n, kdom, seed=kdom % 97)
    descs = _mix("u", 96, seed=kdom)
    batch = _check(qm, "u", cols, descs)
    k2 = [i for i, d in enumerate(descs) if d.groupBy == ["k2"]]
    assert all(batch[i].stats["strategy"] == "part_runs" for i in k2), [batch[i].stats["strategy"] for i in k2]
    for i in k2[:12] + [i for i in range(len(descs)) if i not in k2][:4]:
        solo = qm.executor.execute(descs[i])
        assert_groups_equal(batch[i].rows, solo.rows, descs[i].groupBy, ctx=f"solo q{i}")
    # bytes are the same on a second run (results in ascending key order per query)
    again = qm.executor.execute_batch([descs[i] for i in k2[:8]])
    for j, i in enumerate(k2[:8]):
        assert bytes(again[j].columnarData) == bytes(batch[i].columnarData), i
        keys = [r["k2"] for r in batch[i].rows]
        assert keys == sorted(keys)


def test_unsorted_batch_edges(qm):
    """NULL state values, empty / never-true / full / unfiltered ranges,
    NaN filter values, negative keys, a ragged row count."""
    n = 1_000_003
    _say("unsorted batch edges")
    cols = _table(n, 20_000, seed=3, klo=-7_000)
    rng = np.random.default_rng(5)
    cols["v2"][rng.random(n) < 0.01] = np.nan
    nulls = {"v1": rng.random(n) < 0.1, "v2": rng.random(n) < 0.05}
    a4 = [AggregateOp("sum", "v1"), AggregateOp("count", "*"), AggregateOp("min", "v1"), AggregateOp("max", "v2")]
    a5 = [AggregateOp("count", "v1"), AggregateOp("count", "*"), AggregateOp("max", "v1"), AggregateOp("min", "v2")]
    descs = [
        QueryDescriptor("e", filters=[FilterOp("v2", "gte", 0.7), FilterOp("v2", "lt", 0.2)], groupBy=["k2"],
                        aggregates=a4),                                                        # never true
        QueryDescriptor("e", filters=[FilterOp("v2", "gt", 2.0)], groupBy=["k2"], aggregates=a4),   # empty
        QueryDescriptor("e", filters=[FilterOp("v2", "gte", 0.0)], groupBy=["k2"], aggregates=a4),  # full
        QueryDescriptor("e", filters=[FilterOp("v2", "between", [0.25, 0.5])], groupBy=["k2"], aggregates=a4),
        QueryDescriptor("e", groupBy=["k2"], aggregates=a4),                                    # no filter
        QueryDescriptor("e", filters=[FilterOp("v2", "lt", 0.001)], groupBy=["k2"], aggregates=a4),
        QueryDescriptor("e", filters=[FilterOp("v2", "is_not_null")], groupBy=["k2"], aggregates=a4),
    ] + [QueryDescriptor("e", filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", lo + w)], groupBy=["k2"],
                         aggregates=a4) for lo, w in ((0.1, 0.3), (0.35, 0.01), (0.5, 0.5), (0.0, 0.999))]
    batch = _check(qm, "e", cols, descs, nulls)
    assert sum(r.stats["strategy"] == "part_runs" for r in batch) >= 8
    # a second signature (COUNT(col), MAX of an int, MIN of a float) and NULL-guarded outputs
    descs2 = [QueryDescriptor("e", filters=[FilterOp("v2", "lt", x)], groupBy=["k2"], aggregates=a5)
              for x in (0.05, 0.5, 0.95)]
    _check(qm, "e", cols, descs2, nulls)


@pytest.mark.parametrize("kdt", [np.int8, np.int16, np.int32])
def test_unsorted_narrow_keys(qm, kdt):
    n = 400_000
    lo = {np.int8: -100, np.int16: -3000, np.int32: -50_000}[kdt]
    dom = {np.int8: 200, np.int16: 6000, np.int32: 100_000}[kdt]
    _say(f"unsorted narrow key {np.dtype(kdt).name}")
    cols = _table(n, dom, seed=11, klo=lo, kdt=kdt)
    _check(qm, "k", cols, _mix("k", 16, seed=17))


def test_unsorted_skewed_key_falls_back_or_matches(qm):
    """One key holding a third of the rows: its partition cannot fit the LDS
    of one workgroup, so the launch takes a smaller partition or the
    per-query path; either way every result equals the oracle."""
    n = 600_000
    _say("unsorted skewed key")
    cols = _table(n, 50_000, seed=19)
    cols["k2"][::3] = 777
    _check(qm, "z", cols, _mix("z", 24, seed=23))


def test_unsorted_tiny_tables(qm):
    for n in (1, 7, 100, 8193):
        _say(f"unsorted tiny table {n}")
        cols = _table(n, 50, seed=n)
        _check(qm, f"t{n}", cols, _mix(f"t{n}", 8, seed=n, min_sel=0.05))


def test_unsorted_output_kinds(qm):
    """AVG, MIN of a float32 column (stored as f32), MAX and SUM of an int16
    column (MAX stored as int16), with NULLs in both: every output kind of the
    write pass against the oracle."""
    n = 700_001
    _say("unsorted output kinds")
    cols = _table(n, 30_000, seed=29)
    rng = np.random.default_rng(31)
    cols["f3"] = rng.normal(0, 50, n).astype(np.float32)
    cols["s16"] = rng.integers(-30_000, 30_000, n).astype(np.int16)
    nulls = {"f3": rng.random(n) < 0.07, "s16": rng.random(n) < 0.04}
    aggs = [AggregateOp("avg", "v1"), AggregateOp("min", "f3"), AggregateOp("max", "s16"), AggregateOp("count", "*"),
            AggregateOp("sum", "s16")]
    descs = [QueryDescriptor("o", filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", lo + w)], groupBy=["k2"],
                             aggregates=aggs) for lo, w in ((0.0, 0.001), (0.2, 0.3), (0.0, 1.0), (0.6, 0.05))]
    batch = _check(qm, "o", cols, descs, nulls)
    print([r.stats["strategy"] for r in batch], flush=True)


def test_unsorted_narrow_minmax_extremes(qm):
    """MIN / MAX of int32 and int8 columns keep 32-bit order codes in the write
    pass. Groups whose values are all INT32_MAX (MIN's code is 0), all
    INT32_MIN (MAX's code is 0), all NULL, or mixed with NULLs; int8 at -128 /
    127; record widths 2, 3 and 4 words."""
    n = 240_000
    _say("unsorted narrow MIN / MAX extremes")
    cols = _table(n, 24_000, seed=37)
    k = cols["k2"]
    rng = np.random.default_rng(41)
    v1 = rng.integers(-(2 ** 31), 2 ** 31, n, dtype=np.int64)
    v1[k % 5 == 0] = 2 ** 31 - 1
    v1[k % 5 == 1] = -(2 ** 31)
    v1[(k % 5 == 2) & (rng.random(n) < 0.5)] = 2 ** 31 - 1
    v1[(k % 5 == 3) & (rng.random(n) < 0.5)] = -(2 ** 31)
    cols["v1"] = v1.astype(np.int32)
    i8 = rng.integers(-128, 128, n).astype(np.int8)
    i8[k % 7 == 0] = 127
    i8[k % 7 == 1] = -128
    cols["i8"] = i8
    ranges = ((0.0, 1.0), (0.1, 0.4), (0.5, 0.02), (0.0, 0.3))

    def qs(name, aggs):
        return [QueryDescriptor(name, filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", lo + w)],
                                groupBy=["k2"], aggregates=aggs) for lo, w in ranges]

    # no NULLs: 4-word records (key / bucket word, v1, v2)
    a = qs("x", [AggregateOp("max", "v1"), AggregateOp("count", "*"), AggregateOp("min", "v1"),
                 AggregateOp("max", "v2")])
    b = _check(qm, "x", cols, a)
    assert all(r.stats["strategy"] == "part_runs" for r in b), [r.stats["strategy"] for r in b]
    # int8 only: 2-word records
    c = qs("x", [AggregateOp("min", "i8"), AggregateOp("max", "i8"), AggregateOp("count", "*")])
    b = _check(qm, "x", cols, c)
    assert all(r.stats["strategy"] == "part_runs" for r in b), [r.stats["strategy"] for r in b]
    # NULLs in v1, whole groups NULL (keys % 5 == 4): a NULL-bit word, NULL-guarded outputs
    nulls = {"v1": (k % 5 == 4) | (rng.random(n) < 0.1)}
    d = qs("y", [AggregateOp("min", "v1"), AggregateOp("max", "v1"), AggregateOp("count", "*")])
    b = _check(qm, "y", cols, d, nulls)
    assert all(r.stats["strategy"] == "part_runs" for r in b), [r.stats["strategy"] for r in b]
