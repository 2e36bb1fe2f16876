"""A full ORDER BY (no LIMIT) in key-range passes (the device counterpart of
the reference's external sort for inputs over the memory budget,
src/operators.ts:2811-3010): a histogram of the top key bits cuts the key
space into ranges; each pass sorts the rows of one range stably; NULL / NaN
rows come last in row order.

* Forced pass counts (the sort_parts knob) give byte-identical results to
  one pass, with NaN, NULLs, -0 / +0, ties, both directions and OFFSET; and
  equal the oracle.
* A context whose HBM budget cannot hold the one-pass temporaries sorts in
  passes by itself, while the same sort forced to one pass fails with
  MEMORY_EXCEEDED."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd import QueryMode
from querymode_amd.errors import QueryModeError
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import assert_rows_equal

pytestmark = pytest.mark.gpu


def _table(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 10, n)
    u = rng.random(n)
    f[u < 0.01] = np.nan
    f[(u >= 0.01) & (u < 0.05)] = 0.0
    f[(u >= 0.05) & (u < 0.08)] = -0.0
    f[(u >= 0.08) & (u < 0.2)] = np.round(f[(u >= 0.08) & (u < 0.2)])  # ties
    cols = {"f": f, "i": rng.integers(-1000, 1000, n).astype(np.int32),
            "w": rng.integers(-(2 ** 62), 2 ** 62, n).astype(np.int64), "k": rng.integers(0, 5000, n).astype(np.int32)}
    nulls = {"f": rng.random(n) < 0.05, "i": rng.random(n) < 0.03}
    return cols, nulls


@pytest.mark.parametrize("col,direction,offset", [("f", "asc", None), ("f", "desc", 1234), ("i", "asc", None),
                                                  ("i", "desc", None), ("w", "asc", 50_000)])
def test_forced_passes_equal_one_pass(qm, col, direction, offset):
    cols, nulls = _table(120_003, seed=51)
    qm.register_table("sp_t", cols, nulls)
    d = QueryDescriptor("sp_t", filters=[FilterOp("k", "lt", 4000)], sortColumn=col, sortDirection=direction,
                        offset=offset, projections=["k", "f", "i", "w"])
    try:
        qm.device.tune_set("sort_parts", 1)
        one = qm.executor.execute(d)
        for parts in (2, 7, 64):
            print(f"sort parts: {col} {direction} offset {offset}, {parts} passes", flush=True)
            qm.device.tune_set("sort_parts", parts)
            got = qm.executor.execute(d)
            assert bytes(got.columnarData) == bytes(one.columnarData), parts
    finally:
        qm.device.tune_set("sort_parts", 0)
    want = O.run(d, {"sp_t": O.Table(cols, nulls)})
    assert_rows_equal(one.rows, want, ctx=f"oracle {col} {direction}")


def test_grouped_sort_in_passes(qm):
    cols, nulls = _table(80_000, seed=52)
    qm.register_table("sp_t", cols, nulls)
    d = QueryDescriptor("sp_t", groupBy=["k"], aggregates=[AggregateOp("count", "*", "n"), AggregateOp("max", "f")],
                        sortColumn="max_f", sortDirection="desc")
    try:
        qm.device.tune_set("sort_parts", 1)
        one = qm.executor.execute(d)
        qm.device.tune_set("sort_parts", 9)
        got = qm.executor.execute(d)
    finally:
        qm.device.tune_set("sort_parts", 0)
    assert bytes(got.columnarData) == bytes(one.columnarData)


def test_small_budget_sorts_in_passes():
    cols, nulls = _table(3_000_000, seed=53)
    d = QueryDescriptor("t", sortColumn="f", sortDirection="asc", projections=["f", "k"])
    big = QueryMode.gpu(0)
    try:
        big.register_table("t", cols, nulls)
        used = big.device.info()["used"]
        big.device.tune_set("sort_parts", 1)
        want = bytes(big.executor.execute(d).columnarData)
    finally:
        big.close()
    n = 3_000_000
    # what the query holds at once besides the sort scratch: row ids (8 B/row), the gathered columns
    # (12 B/row), the order (8 B/row) and the permuted columns (12 B/row); room for passes on top,
    # not for one pass's ~50 B/row of radix sort scratch
    budget = used + n * 60
    print(f"sort parts: budget {budget >> 20} MiB, tables {used >> 20} MiB", flush=True)
    small = QueryMode.gpu(0, hbm_budget_bytes=budget)
    try:
        small.register_table("t", cols, nulls)
        got = small.executor.execute(d)
        assert bytes(got.columnarData) == want
        small.device.tune_set("sort_parts", 1)
        with pytest.raises(QueryModeError) as e:
            small.executor.execute(d)
        assert e.value.code == "MEMORY_EXCEEDED"
    finally:
        small.close()
