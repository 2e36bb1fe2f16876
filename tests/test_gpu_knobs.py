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


def test_topk_blocks_keep_results(qm):
    """Top-K with 1, 2 and 8 scan workgroups per CU (each keeps its own
    candidates; the shared k-th threshold prunes the others): the same rows in
    the same order, ties (repeated values) broken by row id, NULLs last."""
    import numpy as np

    from oracle import oracle as O
    from querymode_amd.plan import FilterOp

    from ._compare import assert_rows_equal

    n = 3_000_017
    _say(f"top-K workgroups per CU: {n} rows")
    rng = np.random.default_rng(23)
    cols = {"v": rng.integers(0, 5000, n).astype(np.int64), "w": rng.random(n), "g": rng.integers(0, 9, n).astype(np.int32)}
    nulls = {"w": rng.random(n) < 0.05}
    qm.register_table("tk", cols, nulls)
    ot = O.Table(cols, nulls)
    descs = [QueryDescriptor("tk", projections=["v", "g"], sortColumn="v", sortDirection="desc", limit=k)
             for k in (10, 100, 1000)]
    descs.append(QueryDescriptor("tk", projections=["w"], sortColumn="w", sortDirection="asc", limit=500, offset=3))
    descs.append(QueryDescriptor("tk", filters=[FilterOp("g", "eq", 4)], projections=["w", "v"], sortColumn="w",
                                 sortDirection="desc", limit=100))
    dev = qm.device
    old = dev.tune_get("topk_bpc")
    ex = qm.executor
    try:
        want = None
        for bpc in (2, 1, 8):
            dev.tune_set("topk_bpc", bpc)
            got = [bytes(ex.execute(d).columnarData) for d in descs]
            if want is None:
                want = got
                for d in descs:
                    assert_rows_equal(ex.execute(d).rows, O.run(d, {"tk": ot}), ctx=f"top-K bpc {bpc}")
            assert got == want, bpc
    finally:
        dev.tune_set("topk_bpc", old)


PR_SETTINGS = [
    {"pr_max_s": 5}, {"pr_max_s": 4}, {"pr_b2": 7}, {"pr_b2": 10}, {"pr_max_s": 5, "pr_b2": 9},
    {"pr_scatter_bpc": 1}, {"pr_scatter_bpc": 3}, {"shared_waves": 1}, {"shared_waves": 16},
]


def test_partitioned_runs_knobs_keep_results(qm):
    """The batched path over a key in any row order under every partition
    shape the tuner may pick (keys per final partition, the split between the
    two scatter passes, scatter workgroups per CU) and the shared-interval
    pass's wave rounds: the same bytes as the default, which equal the oracle."""
    import numpy as np

    from oracle import oracle as O

    from ._compare import assert_groups_equal

    n = 3_000_000
    _say(f"partitioned-runs knobs: {n} rows, 64 queries, {len(PR_SETTINGS)} settings")
    cols = datagen.batch_table(n=n, seed=9)
    cols["k2"] = datagen.uniform_int(9, 72, 0, n, 0, 60_000, np.int64)
    qm.register_table("pk", cols)
    descs = [QueryDescriptor("pk", filters=d.filters, groupBy=d.groupBy, aggregates=d.aggregates)
             for d in datagen.batch_queries(64, seed=9, min_sel=1e-3)]
    dev = qm.device
    names = sorted({k for s in PR_SETTINGS for k in s})
    defaults = {k: dev.tune_get(k) for k in names}
    ex = qm.executor
    base = ex.execute_batch(descs)
    k2 = [i for i, d in enumerate(descs) if d.groupBy == ["k2"]]
    assert all(base[i].stats["strategy"] == "part_runs" for i in k2)
    ot = O.Table(cols)
    for i in k2[:4] + [i for i in range(len(descs)) if i not in k2][:2]:
        assert_groups_equal(base[i].rows, O.run(descs[i], {"pk": ot}), descs[i].groupBy, ctx=f"default q{i}")
    want = _bytes(base)
    try:
        for setting in PR_SETTINGS:
            _say(f"  setting {setting}")
            for k, v in setting.items():
                dev.tune_set(k, v)
            got = ex.execute_batch(descs)
            for i, (g, w) in enumerate(zip(got, base)):
                if g.stats["strategy"] == w.stats["strategy"]:
                    assert bytes(g.columnarData) == want[i], (setting, i)
                else:  # the setting left no partition shape for this launch: another path, same groups
                    assert_groups_equal(g.rows, w.rows, descs[i].groupBy, ctx=f"{setting} q{i}")
            for k in setting:
                dev.tune_set(k, defaults[k])
    finally:
        for k, v in defaults.items():
            dev.tune_set(k, v)
