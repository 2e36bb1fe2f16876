"""Edge cases of the run-sorted batch kernels (one sorted group key, one range
column): tables smaller than one chunk and ragged ones, one run spanning every
chunk, keys at the int64 extremes, narrow key dtypes, NULLs in the filter and
aggregate columns, never-true and unfiltered queries, and the largest bound
set the bucket search takes (510 distinct bounds) next to one it does not.
Each batched result must equal the oracle's and the same query run alone."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import assert_groups_equal

pytestmark = pytest.mark.gpu

AGGS_INT = [AggregateOp("count", "*"), AggregateOp("sum", "v1"), AggregateOp("min", "v1"), AggregateOp("max", "v2")]
# four states (COUNT(*), SUM(v1), COUNT(v1), MIN(v1)): what the run-sorted kernels hold
AGGS_NULL = [AggregateOp("count", "*"), AggregateOp("sum", "v1"), AggregateOp("count", "v1"), AggregateOp("min", "v1"),
             AggregateOp("avg", "v1")]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def runs_any_len(qm):
    """Let the run-sorted kernels take keys with short runs (tiny tables)."""
    dev = qm.device
    old = dev.tune_get("runs_min_len")
    dev.tune_set("runs_min_len", 0.0)
    yield qm
    dev.tune_set("runs_min_len", old)


def _sorted_key(rng, n, nruns, lo=0, stride=1000, dtype=np.int64):
    """Ascending key of about nruns runs, stride apart (a key domain over 64
    values keeps filtered queries off the shared-interval pass)."""
    cuts = np.sort(rng.integers(0, n, max(0, nruns - 1)))
    key = np.zeros(n, np.int64)
    np.add.at(key, cuts, 1)
    return (np.cumsum(key) * stride + lo).astype(dtype)


def _cols(n, seed, key):
    rng = np.random.default_rng(seed)
    return {
        "k": key,
        "v1": rng.integers(-(2 ** 31), 2 ** 31, n, dtype=np.int64).astype(np.int32),
        "v2": rng.random(n),
    }


def _mix(name, aggs, rng, nq):
    descs = [QueryDescriptor(name, groupBy=["k"], aggregates=aggs),                      # unfiltered
             QueryDescriptor(name, filters=[FilterOp("v2", "gte", 0.7), FilterOp("v2", "lt", 0.2)],
                             groupBy=["k"], aggregates=aggs),                            # never true
             QueryDescriptor(name, filters=[FilterOp("v2", "gte", 0.0)], groupBy=["k"], aggregates=aggs),
             QueryDescriptor(name, filters=[FilterOp("v2", "gt", 2.0)], groupBy=["k"], aggregates=aggs)]
    while len(descs) < nq:
        s = float(np.exp(rng.uniform(np.log(1e-3), np.log(0.6))))
        lo = float(rng.uniform(0.0, 1.0 - s))
        descs.append(QueryDescriptor(name, filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", lo + s)],
                                     groupBy=["k"], aggregates=aggs))
    return descs


def _check(qm, name, cols, descs, nulls=None, strategy="sorted_runs", solo=4):
    qm.register_table(name, cols, nulls)
    batch = qm.executor.execute_batch(descs)
    ot = O.Table(cols, nulls)
    for i, (d, got) in enumerate(zip(descs, batch)):
        want = O.run(d, {name: ot})
        assert_groups_equal(got.rows, want, d.groupBy, ctx=f"{name} q{i}")
        if strategy is not None:
            assert got.stats["strategy"] == strategy, (name, i, got.stats["strategy"])
    for i, d in enumerate(descs[:solo]):
        one = qm.executor.execute(d)
        assert_groups_equal(one.rows, batch[i].rows, d.groupBy, ctx=f"{name} solo q{i}")
    return batch


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1023, 1024, 1025, 4097, 100_003])
def test_runs_small_and_ragged(runs_any_len, n):
    qm = runs_any_len
    _say(f"runs edges: {n} rows")
    rng = np.random.default_rng(n)
    cols = _cols(n, n, _sorted_key(rng, n, max(2, n // 3), lo=-4000))
    one_key = len(np.unique(cols["k"])) == 1  # a one-value key takes the shared-interval pass
    _check(qm, f"rs{n}", cols, _mix(f"rs{n}", AGGS_INT, rng, 12), strategy=None if one_key else "sorted_runs")


def test_runs_one_run_spans_every_chunk(qm):
    n = 2_000_000
    _say(f"runs edges: one key over {n} rows, then two runs")
    rng = np.random.default_rng(11)
    cols = _cols(n, 11, np.full(n, 42, np.int64))
    b = _check(qm, "r1", cols, _mix("r1", AGGS_INT, rng, 16), strategy=None)
    assert [(r["k"], r["count_*"]) for r in b[0].rows] == [(42, n)]
    key = np.where(np.arange(n) < 1_234_567, -5, 10 ** 9).astype(np.int64)
    _check(qm, "r2", _cols(n, 12, key), _mix("r2", AGGS_INT, rng, 16))


@pytest.mark.parametrize("kdt", [np.int8, np.int16, np.int32])
def test_runs_narrow_keys(qm, kdt):
    n = 600_001
    _say(f"runs edges: sorted {np.dtype(kdt).name} key")
    rng = np.random.default_rng(21)
    info = np.iinfo(kdt)
    lo, nruns, stride = {np.int8: (int(info.min), 80, 3), np.int16: (-30000, 2000, 30),
                         np.int32: (-(2 ** 31), 2000, 2 ** 21)}[kdt]
    cols = _cols(n, 21, _sorted_key(rng, n, nruns, lo=lo, stride=stride, dtype=kdt))
    assert np.all(np.diff(cols["k"].astype(np.int64)) >= 0)  # no wrap-around in the narrow type
    name = f"rn{np.dtype(kdt).name}"
    _check(qm, name, cols, _mix(name, AGGS_INT, rng, 24))


def test_runs_int64_extreme_keys(qm):
    n = 500_000
    _say("runs edges: keys at the int64 extremes")
    rng = np.random.default_rng(31)
    edges = np.array([-(2 ** 63), -(2 ** 63) + 1, -1, 0, 1, 2 ** 62, 2 ** 63 - 2, 2 ** 63 - 1], np.int64)
    key = np.sort(edges[rng.integers(0, len(edges), n)])
    _check(qm, "rx", _cols(n, 31, key), _mix("rx", AGGS_INT, rng, 16))


def test_runs_nulls(qm):
    n = 1_500_007
    _say("runs edges: NULLs in the filter and aggregate columns")
    rng = np.random.default_rng(41)
    cols = _cols(n, 41, _sorted_key(rng, n, 20_000))
    nulls = {"v1": rng.random(n) < 0.2, "v2": rng.random(n) < 0.1}
    descs = _mix("rnull", AGGS_NULL, rng, 20)
    descs.append(QueryDescriptor("rnull", filters=[FilterOp("v2", "is_not_null")], groupBy=["k"],
                                 aggregates=AGGS_NULL))
    _check(qm, "rnull", cols, descs, nulls, strategy=None)
    # without the IS NOT NULL query the batch runs on the run-sorted kernels
    got = qm.executor.execute_batch(descs[:-1])
    assert {r.stats["strategy"] for r in got} == {"sorted_runs"}


def test_runs_bound_limit(qm):
    """255 queries with distinct two-sided ranges give 510 bounds, the most the
    bucket search holds; one more query (512 bounds) goes to another kernel."""
    n = 800_000
    _say("runs edges: 510 and 512 distinct bounds")
    rng = np.random.default_rng(51)
    cols = _cols(n, 51, _sorted_key(rng, n, 5000))
    descs = []
    while len(descs) < 256:
        lo = float(rng.uniform(0.0, 0.9))
        s = float(rng.uniform(1e-3, 0.1))
        descs.append(QueryDescriptor("rb", filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", lo + s)],
                                     groupBy=["k"], aggregates=AGGS_INT))
    _check(qm, "rb", cols, descs[:255], solo=2)
    b = _check(qm, "rb", cols, descs, strategy=None, solo=0)
    assert b[0].stats["strategy"] != "sorted_runs"
