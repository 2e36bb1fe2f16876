"""Every planner knob setting gives the same results, byte for byte (the
empirical tuner relies on it): the same batch and single queries through the
run-sorted kernels with each staging block width, chunk count and placement
window, the query-parallel kernel, the segmented kernel, the fused scan, and
both shared-interval passes."""
from __future__ import annotations

import sys

import pytest

from querymode_amd import datagen
from querymode_amd.plan import QueryDescriptor

pytestmark = pytest.mark.gpu

SETTINGS = [
    {"runs_qblock": 0}, {"runs_qblock": 16}, {"runs_qblock": 32},
    {"runs_win": 1024}, {"runs_chunks_per_cu": 2}, {"runs_chunks_per_cu": 128},
    {"runs_min_q": 100000},                       # sorted keys: query-parallel kernel
    {"runs_min_q": 100000, "qp_min": 100000},     # sorted keys: segmented kernel
    {"shared_lean": 0},                           # shared groups: fused multi-group pass
    {"sorted1": 0},                               # single sorted-key query: segmented kernel
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bytes(results):
    return [bytes(r.columnarData) for r in results]


def test_knobs_keep_results(qm):
    n = 2_000_000
    _say(f"knobs: {n} rows, 48 queries, {len(SETTINGS)} settings")
    cols = datagen.batch_table(n=n, seed=7, k2_order=True)
    cols["k2"] = datagen.sorted_uniform_i64(7, 71, n, 0, 5000)
    qm.register_table("kn", cols)
    descs = [QueryDescriptor("kn", filters=d.filters, groupBy=d.groupBy, aggregates=d.aggregates)
             for d in datagen.batch_queries(48, seed=7, min_sel=1e-3)]
    dev = qm.device
    names = sorted({k for s in SETTINGS for k in s})
    defaults = {k: dev.tune_get(k) for k in names}
    ex = qm.executor
    want_batch = _bytes(ex.execute_batch(descs))
    want_single = _bytes([ex.execute(d) for d in descs[:6]])
    try:
        for setting in SETTINGS:
            _say(f"  setting {setting}")
            for k, v in setting.items():
                dev.tune_set(k, v)
            assert _bytes(ex.execute_batch(descs)) == want_batch, setting
            assert _bytes([ex.execute(d) for d in descs[:6]]) == want_single, setting
            for k in setting:
                dev.tune_set(k, defaults[k])
    finally:
        for k, v in defaults.items():
            dev.tune_set(k, v)
