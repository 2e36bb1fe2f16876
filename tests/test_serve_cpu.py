"""Scheduler logic of the serving loop (querymode_amd/serve.py) with a fake
executor: admission by the HBM budget, continuous batching (aggregates that
arrive while a batch over their table runs go out together as the next
batch), fused aggregate batches, and short requests running on other lanes
beside a running batch."""
from __future__ import annotations

import threading
import time

import pytest

from querymode_amd import serve
from querymode_amd.errors import QueryModeError
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor


class FakeDev:
    def __init__(self, budget):
        self.budget = budget

    def info(self):
        return {"budget": self.budget, "used": 0}


class FakeExec:
    def __init__(self, delay=0.05):
        self.delay = delay
        self.calls = []
        self.gate = threading.Event()        # holds every call
        self.gate.set()
        self.agg_gate = threading.Event()    # holds aggregate calls only
        self.agg_gate.set()
        self.agg_started = threading.Event()

    def execute(self, d):
        self.gate.wait()
        if d.aggregates:
            self.agg_started.set()
            self.agg_gate.wait()
        time.sleep(self.delay)
        self.calls.append(("one", d.table))
        return f"one:{d.table}:{d.limit}"

    def execute_batch(self, ds):
        self.gate.wait()
        self.agg_started.set()
        self.agg_gate.wait()
        time.sleep(self.delay)
        self.calls.append(("batch", len(ds)))
        return [f"batch:{d.table}:{d.limit}" for d in ds]


class FakeQM:
    def __init__(self, budget, delay=0.05):
        self.device = FakeDev(budget)
        self.executor = FakeExec(delay)


def _agg(limit=None, table="t"):
    return QueryDescriptor(table, filters=[FilterOp("v", "lt", 1)], groupBy=["k"], aggregates=[AggregateOp("count", "*")],
                           limit=limit)


def _topk(k):
    return QueryDescriptor("t", sortColumn="v", sortDirection="desc", limit=k)


@pytest.fixture
def server(monkeypatch):
    made = []

    def make(budget=10_000, sizes=None, delay=0.05, slots=256):
        qm = FakeQM(budget, delay)
        monkeypatch.setattr(serve.Server, "_estimate", lambda self, r: (sizes or {}).get(r.desc.limit, 100))
        s = serve.Server(qm, max_slots=slots, hbm_reserve=0.0)
        made.append(s)
        return s, qm

    yield make
    for s in made:
        s.close()


def test_aggregates_of_one_cycle_are_fused(server):
    s, qm = server()
    qm.executor.gate.clear()                 # hold the first cycle
    first = s.submit(_agg(1))
    time.sleep(0.05)
    rest = [s.submit(_agg(i)) for i in range(2, 12)] + [s.submit(_topk(10))]
    qm.executor.gate.set()
    for r in [first] + rest:
        r.wait(5)
    # first cycle: the lone request; second cycle: ten fused aggregates, then the top-K
    assert qm.executor.calls[0] == ("one", "t")
    assert ("batch", 10) in qm.executor.calls
    assert rest[-1].result == "one:t:10"
    assert all(r.latency_ms > 0 for r in [first] + rest)


def test_admission_respects_the_budget(server):
    # requests of 400 bytes each against a budget of 1000: at most two per cycle
    s, qm = server(budget=1000, sizes={i: 400 for i in range(10)})
    qm.executor.gate.clear()
    rs = [s.submit(_agg(i)) for i in range(6)]
    time.sleep(0.05)
    qm.executor.gate.set()
    for r in rs:
        r.wait(5)
    sizes = [c[1] if c[0] == "batch" else 1 for c in qm.executor.calls]
    assert max(sizes) <= 2 and sum(sizes) == 6


def test_request_larger_than_budget_fails_alone(server):
    s, qm = server(budget=1000, sizes={7: 5000})
    big = s.submit(_agg(7))
    ok = s.submit(_agg(1))
    with pytest.raises(QueryModeError) as e:
        big.wait(5)
    assert e.value.code == "MEMORY_EXCEEDED"
    assert ok.wait(5).startswith("one:") or ok.wait(5).startswith("batch:")


def test_slots_bound_the_cycle(server):
    s, qm = server(slots=3)
    qm.executor.gate.clear()
    rs = [s.submit(_agg(i)) for i in range(7)]
    time.sleep(0.05)
    qm.executor.gate.set()
    for r in rs:
        r.wait(5)
    sizes = [c[1] if c[0] == "batch" else 1 for c in qm.executor.calls]
    assert max(sizes) <= 3 and sum(sizes) == 7


def test_short_requests_run_beside_a_running_batch(server):
    # a batch holds its lane; top-K requests that arrive meanwhile finish on the other lanes
    s, qm = server(delay=0.0)
    qm.executor.agg_gate.clear()
    aggs = [s.submit(_agg(i)) for i in range(3)]
    assert qm.executor.agg_started.wait(5)
    tops = [s.submit(_topk(k)) for k in (10, 100, 1000)]
    for r in tops:
        assert r.wait(5).startswith("one:")
    assert not any(r._event.is_set() for r in aggs)   # the batch is still on its lane
    qm.executor.agg_gate.set()
    for r in aggs:
        r.wait(5)
    assert s.max_overlap >= 2
    assert all(r.overlap >= 1 for r in tops)


def test_aggregates_arriving_during_a_batch_form_the_next_one(server):
    s, qm = server(delay=0.0)
    qm.executor.agg_gate.clear()
    first = s.submit(_agg(0))
    assert qm.executor.agg_started.wait(5)
    later = [s.submit(_agg(i)) for i in range(1, 9)]
    time.sleep(0.1)
    assert not any(r._event.is_set() for r in later)   # they wait for the table's batch
    qm.executor.agg_gate.set()
    for r in [first] + later:
        r.wait(5)
    assert qm.executor.calls[0] == ("one", "t")
    assert ("batch", 8) in qm.executor.calls
    assert all(r.batch_size == 8 for r in later)


def test_tables_batch_independently(server):
    # a batch over table t does not hold back aggregates over table u
    s, qm = server(delay=0.0)
    qm.executor.agg_gate.clear()
    a = s.submit(_agg(0, table="t"))
    assert qm.executor.agg_started.wait(5)
    b = [s.submit(_agg(i, table="u")) for i in range(1, 4)]
    time.sleep(0.1)
    assert s.max_overlap == 2   # the u batch runs on a second lane
    qm.executor.agg_gate.set()
    for r in [a] + b:
        r.wait(5)


def test_close_finishes_queued_requests(server):
    s, qm = server(delay=0.01)
    rs = [s.submit(_agg(i)) for i in range(5)] + [s.submit(_topk(10)) for _ in range(5)]
    s.close()
    assert all(r._event.is_set() and r.error is None for r in rs)


def test_percentile():
    v = list(range(1, 101))
    assert serve.percentile(v, 50) == 50 and serve.percentile(v, 99) == 99 and serve.percentile(v, 100) == 100
    assert serve.percentile([5.0], 99) == 5.0
