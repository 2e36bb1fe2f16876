"""The serving loop on the device: a concurrent mix of group-by, join and
top-K requests arriving while others run, several launches on the device at
once (one per execution lane); each result must equal the same request run
alone (bit for bit; FP64 sums are exact and order-independent)."""
from __future__ import annotations

import random
import threading

import numpy as np
import pytest

from querymode_amd import datagen
from querymode_amd.plan import AggregateOp, JoinDescriptor, QueryDescriptor
from querymode_amd.serve import Server, percentile

from ._compare import assert_groups_equal, assert_rows_equal

pytestmark = pytest.mark.gpu


def test_mix_alone_equals_in_mix(qm):
    print("serve: 2M-row batch table + 1M x 100K join tables, 90 mixed requests", flush=True)
    cols = datagen.batch_table(n=2_000_000, seed=1, k2_order=True)
    cols["k2"] = datagen.sorted_uniform_i64(1, 71, 2_000_000, 0, 20_000)
    qm.register_table("batch", cols)
    qm.register_table("build", datagen.join_build(n=100_000, seed=1))
    qm.register_table("probe", datagen.join_probe(n=1_000_000, nids=100_000, seed=1))
    join = QueryDescriptor("probe", join=JoinDescriptor(QueryDescriptor("build"), "fk", "id"), groupBy=["region"],
                           computedColumns=[("aw", "amount", "weight")],
                           aggregates=[AggregateOp("sum", "aw"), AggregateOp("count", "*")])
    rng = random.Random(7)
    groupbys = datagen.batch_queries(64, seed=3, min_sel=1e-3)
    mix = []
    for i in range(90):
        u = rng.random()
        if u < 0.4:
            mix.append(groupbys[i % len(groupbys)])
        elif u < 0.7:
            mix.append(join)
        else:
            mix.append(QueryDescriptor("batch", sortColumn="v2", sortDirection="desc", limit=rng.choice([10, 100, 1000]),
                                       projections=["k1", "k2", "v1", "v2"]))
    srv = Server(qm)
    reqs = []
    lock = threading.Lock()

    def client(part):
        for d in part:
            r = srv.submit(d)
            with lock:
                reqs.append(r)

    threads = [threading.Thread(target=client, args=(mix[k::3],)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for r in reqs:
        r.wait(300)
    srv.close()
    assert srv.launches < len(reqs)  # requests were batched together
    assert srv.batched > 0
    assert srv.max_overlap >= 2      # launches ran beside each other
    for i, r in enumerate(reqs):
        alone = qm.executor.execute(r.desc)
        got = r.result
        if r.desc.aggregates:
            assert_groups_equal(got.rows, alone.rows, r.desc.groupBy, ctx=f"request {i}")
        else:
            assert_rows_equal(got.rows, alone.rows, ctx=f"request {i}")
    lat = [r.latency_ms for r in reqs]
    print(f"p50 {percentile(lat, 50):.2f} ms  p99 {percentile(lat, 99):.2f} ms  launches {srv.launches}  "
          f"max concurrent {srv.max_overlap}  overlap share {srv.overlap_share:.2f}", flush=True)
