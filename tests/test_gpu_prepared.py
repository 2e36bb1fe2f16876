"""Prepared batches (qm_batch_prepare / qm_prepared_run): the batch captured
into a HIP graph and replayed. Every replay must give byte-identical QMCB to
qm_batch_exec of the same plans, run after run, with other device work (the
program's own eager launches, another batch, single queries) between the
replays; the batch results themselves are checked against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from querymode_amd import _native as N
from querymode_amd import datagen, qmcb
from querymode_amd.errors import QueryModeError
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

from ._compare import assert_groups_equal

pytestmark = pytest.mark.gpu


def _bytes(results):
    return [bytes(r.columnarData) for r in results]


def _same(got, want, wb, descs, ctx):
    """Every result holds the same groups as qm_batch_exec's; results of the
    key-ordered paths (DIRECT tables: GROUP BY k1 here) are byte-identical.
    Hash-table results may list their groups in another order (slot order
    follows the order of concurrent inserts). Returns how many results were
    byte-identical."""
    gb = _bytes(got)
    same = 0
    for i, d in enumerate(descs):
        if gb[i] == wb[i]:
            same += 1
            continue
        assert want[i].stats["strategy"] in ("hash_lds", "hash_global", "part_hash"), \
            f"{ctx}: query {i} ({want[i].stats['strategy']}) bytes differ from qm_batch_exec"
        _same_groups(gb[i], wb[i], d.groupBy[0], f"{ctx}: query {i}")
    return same


def _same_groups(a: bytes, b: bytes, key: str, ctx: str) -> None:
    """Two QMCB results of one GROUP BY (one key column, unique per row) hold
    the same rows, bit for bit, in any order."""
    ba, bb = qmcb.decode(a), qmcb.decode(b)
    assert ba.row_count == bb.row_count, f"{ctx}: {ba.row_count} vs {bb.row_count} groups"
    oa = np.argsort(np.asarray(ba.column(key).data), kind="stable")
    ob = np.argsort(np.asarray(bb.column(key).data), kind="stable")
    n = ba.row_count
    for ca, cb in zip(ba.columns, bb.columns):
        assert ca.name == cb.name and ca.dtype == cb.dtype, ctx
        da = np.asarray(ca.data)[:n][oa]
        db = np.asarray(cb.data)[:n][ob]
        assert da.tobytes() == db.tobytes(), f"{ctx}: column {ca.name} differs"
        na = qmcb.unpack_bits(ca.null_bitmap, n)[oa] if ca.null_bitmap is not None else np.zeros(n, bool)
        nb = qmcb.unpack_bits(cb.null_bitmap, n)[ob] if cb.null_bitmap is not None else np.zeros(n, bool)
        assert (na == nb).all(), f"{ctx}: NULLs of column {ca.name} differ"


@pytest.mark.parametrize("nq,rows", [(7, 100_003), (96, 3_000_000)])
def test_replay_equals_batch_exec(qm, nq, rows):
    print(f"prepared: {nq} queries over {rows} unsorted rows", flush=True)
    cols = datagen.batch_table(n=rows, seed=21)
    qm.register_table("prep_t", cols)
    from dataclasses import replace
    descs = [replace(d, table="prep_t") for d in datagen.batch_queries(nq, seed=22, min_sel=1e-3)]
    want = qm.executor.execute_batch(descs)
    tab = {"prep_t": O.Table(cols)}
    for i in (0, 1, nq - 1):
        assert_groups_equal(want[i].rows, O.run(descs[i], tab), descs[i].groupBy, ctx=f"oracle {i}")
    wb = _bytes(want)
    print("prepared: strategies", sorted({r.stats["strategy"] for r in want}), flush=True)
    print("prepared: qm_batch_prepare", flush=True)
    prep = qm.executor.prepare_batch(descs)
    print("prepared: captured", prep.info(), flush=True)
    try:
        info = prep.info()
        assert info["plans"] == nq and info["graph_nodes"] > 0
        other = [QueryDescriptor("prep_t", sortColumn="v2", sortDirection="desc", limit=100),
                 QueryDescriptor("prep_t", filters=[FilterOp("v2", "lt", 0.25)], groupBy=["k1"],
                                 aggregates=[AggregateOp("count", "*")])]
        for rep in range(4):
            print(f"prepared: replay {rep}", flush=True)
            got = prep.run()
            same = _same(got, want, wb, descs, f"replay {rep}")
            print(f"prepared: replay {rep}: {same} of {nq} results byte-identical, the rest the same groups",
                  flush=True)
            for r in got:
                assert r.stats["kernel_launches"] == 1
            # eager work between replays, reusing pool buffers the graph must not share
            qm.executor.execute(other[rep % 2])
            if rep == 1:
                _same(qm.executor.execute_batch(descs), want, wb, descs, "eager batch between replays")
    finally:
        prep.close()


def test_prepared_batch_pins_its_table(qm):
    print("prepared: table drop refused while a prepared batch is open", flush=True)
    cols = datagen.batch_table(n=200_000, seed=23)
    t = qm.register_table("prep_pin", cols)
    d = QueryDescriptor("prep_pin", filters=[FilterOp("v2", "lt", 0.5)], groupBy=["k1"],
                        aggregates=[AggregateOp("sum", "v1"), AggregateOp("count", "*")])
    prep = qm.executor.prepare_batch([d, d])
    with pytest.raises(QueryModeError):
        N.check(qm.device.lib.qm_table_drop(t.handle))
    a, b = prep.run()
    assert a.rows == b.rows
    prep.close()
    t.drop()
