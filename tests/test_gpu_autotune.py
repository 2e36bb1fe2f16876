"""The empirical tuner on the device: it picks valid knob values and the
results under the tuned knobs equal the untuned results byte for byte."""
from __future__ import annotations

import sys

import pytest

from querymode_amd import datagen
from querymode_amd.autotune import Autotuner, workload_key
from querymode_amd.errors import QueryModeError
from querymode_amd.plan import QueryDescriptor

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_autotune_batch_keeps_results(qm):
    n = 2_000_000
    _say(f"autotune: {n} rows, 32 queries")
    cols = datagen.batch_table(n=n, seed=2, k2_order=True)
    cols["k2"] = datagen.sorted_uniform_i64(2, 71, n, 0, 20_000)
    qm.register_table("t", cols)
    descs = [QueryDescriptor("t", filters=d.filters, groupBy=d.groupBy, aggregates=d.aggregates)
             for d in datagen.batch_queries(32, seed=2, min_sel=1e-3)]
    dev = qm.device
    before = [bytes(r.columnarData) for r in qm.executor.execute_batch(descs)]
    tuner = Autotuner(dev, repeats=2)
    space = {"runs_chunks_per_cu": [4, 16, 64], "runs_win": [1024, 2048]}
    defaults = tuner.current(space)
    res = tuner.tune_batch(qm.executor._table("t"), descs, space=space)
    assert res.ms > 0 and res.default_ms > 0 and len(res.trials) >= 2
    for k, v in res.config.items():
        assert dev.tune_get(k) == v
    after = [bytes(r.columnarData) for r in qm.executor.execute_batch(descs)]
    assert after == before
    assert workload_key(dev, qm.executor._table("t"), descs).startswith(dev.info()["arch"])
    tuner.apply(defaults)


def test_tune_knob_errors(qm):
    dev = qm.device
    with pytest.raises(QueryModeError):
        dev.tune_set("no_such_knob", 1)
    old = dev.tune_get("runs_win")
    dev.tune_set("runs_win", 5000)        # clamped to the knob's range
    assert dev.tune_get("runs_win") == 2048
    dev.tune_set("runs_win", 100)
    assert dev.tune_get("runs_win") == 1024
    dev.tune_set("runs_win", old)


def test_autotune_unsorted_batch_searches_its_kernels(qm):
    """A batch over a key in any row order: the tuner searches the partition
    shape and scatter grid of the partitioned-runs launch and the wave rounds
    of the shared-interval pass (the kernels the batch runs), keeps results
    byte for byte, and leaves the chosen knobs set."""
    import numpy as np

    from querymode_amd.autotune import BATCH_SPACES, space_for
    n = 3_000_000
    _say(f"autotune unsorted: {n} rows, 32 queries")
    cols = datagen.batch_table(n=n, seed=4)
    cols["k2"] = datagen.uniform_int(4, 72, 0, n, 0, 60_000, np.int64)
    qm.register_table("u", cols)
    descs = [QueryDescriptor("u", filters=d.filters, groupBy=d.groupBy, aggregates=d.aggregates)
             for d in datagen.batch_queries(32, seed=4, min_sel=1e-3)]
    dev = qm.device
    before = qm.executor.execute_batch(descs)
    strategies = {r.stats["strategy"] for r in before}
    assert "part_runs" in strategies, strategies
    tuner = Autotuner(dev, repeats=1)
    names = list(space_for(strategies, BATCH_SPACES))
    assert set(BATCH_SPACES["part_runs"]) <= set(names)
    defaults = tuner.current(names)
    try:
        res = tuner.tune_batch(qm.executor._table("u"), descs, passes=1)
        tried = {k for cfg, _ in res.trials for k, v in cfg.items() if v != defaults.get(k)}
        assert tried == set(names), tried
        for k, v in res.config.items():
            assert dev.tune_get(k) == v
        after = qm.executor.execute_batch(descs)
        assert [bytes(r.columnarData) for r in after] == [bytes(r.columnarData) for r in before]
    finally:
        tuner.apply(defaults)
