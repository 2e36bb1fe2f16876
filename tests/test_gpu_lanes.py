"""Calls on one device context from several threads at once: each call holds
one execution lane (a stream of its own) for its duration, so up to four run
on the device together. Every result must equal the same query run alone,
and buffers one call releases must never reach another call while its
kernels may still use them (checked by repeating every query many times
under contention and comparing each result)."""
from __future__ import annotations

import threading

import pytest

from oracle import oracle as O
from querymode_amd import datagen
from querymode_amd.plan import AggregateOp, FilterOp, JoinDescriptor, QueryDescriptor

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu


def _queries():
    qs = datagen.batch_queries(12, seed=5, min_sel=1e-3)
    qs += [QueryDescriptor("lanes_t", sortColumn="v2", sortDirection="desc", limit=k, projections=["k1", "k2", "v2"])
           for k in (10, 100, 1000)]
    qs.append(QueryDescriptor("lanes_t", filters=[FilterOp("v2", "lt", 0.001)], projections=["k2", "v1"]))
    qs.append(QueryDescriptor("lanes_p", join=JoinDescriptor(QueryDescriptor("lanes_b"), "fk", "id"),
                              groupBy=["region"], computedColumns=[("aw", "amount", "weight")],
                              aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*")]))
    return qs


def _check(got, want, d, ctx):
    if d.aggregates:
        assert_groups_equal(got.rows, want.rows, d.groupBy, ctx=ctx)
    else:
        assert_rows_equal(got.rows, want.rows, ctx=ctx)


def test_concurrent_calls_equal_alone(qm):
    print("lanes: 2M-row batch table (unsorted k2) + 1M x 100K join, 4 threads x 17 queries x 3", flush=True)
    cols = datagen.batch_table(n=2_000_000, seed=4)
    qm.register_table("lanes_t", cols)
    bcols = datagen.join_build(n=100_000, seed=2)
    pcols = datagen.join_probe(n=1_000_000, nids=100_000, seed=2)
    qm.register_table("lanes_b", bcols)
    qm.register_table("lanes_p", pcols)
    qs = [d if d.table != "batch" else _retable(d) for d in _queries()]
    alone = [qm.executor.execute(d) for d in qs]
    # the alone results themselves against the oracle, for a few of them
    tabs = {"lanes_t": O.Table(cols), "lanes_b": O.Table(bcols), "lanes_p": O.Table(pcols)}
    for i in (0, 12, 16):
        want = O.run(qs[i], tabs)
        if qs[i].aggregates:
            assert_groups_equal(alone[i].rows, want, qs[i].groupBy, ctx=f"oracle {i}")
        else:
            assert_rows_equal(alone[i].rows, want, ctx=f"oracle {i}")
    errors = []

    def worker(w):
        try:
            for rep in range(3):
                order = list(range(len(qs)))[w::2] + list(range(len(qs)))[(w + 1) % 2::2]
                for i in order:
                    _check(qm.executor.execute(qs[i]), alone[i], qs[i], f"thread {w} rep {rep} query {i}")
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[0]


def test_batch_beside_single_queries(qm):
    print("lanes: a fused batch on one lane while single queries run on the others", flush=True)
    cols = datagen.batch_table(n=3_000_000, seed=6)
    qm.register_table("lanes_t", cols)
    batch = [_retable(d) for d in datagen.batch_queries(64, seed=8, min_sel=1e-3)]
    solo = [QueryDescriptor("lanes_t", sortColumn="v2", sortDirection="desc", limit=100, projections=["k2", "v2"]),
            batch[0], batch[1]]
    want_batch = qm.executor.execute_batch(batch)
    want_solo = [qm.executor.execute(d) for d in solo]
    errors = []

    def singles():
        try:
            for rep in range(10):
                for d, w in zip(solo, want_solo):
                    _check(qm.executor.execute(d), w, d, f"single rep {rep}")
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    ts = [threading.Thread(target=singles) for _ in range(2)]
    for t in ts:
        t.start()
    for rep in range(3):
        got = qm.executor.execute_batch(batch)
        for i, (g, w, d) in enumerate(zip(got, want_batch, batch)):
            _check(g, w, d, f"batch rep {rep} query {i}")
    for t in ts:
        t.join()
    assert not errors, errors[0]


def _retable(d: QueryDescriptor) -> QueryDescriptor:
    from dataclasses import replace
    return replace(d, table="lanes_t")
