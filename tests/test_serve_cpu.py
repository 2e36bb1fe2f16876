"""Scheduler logic of the serving loop (querymode_amd/serve.py) with a fake
executor: admission by the HBM budget, continuous batching (requests that
arrive during a cycle run in the next one), fused aggregate batches."""
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
        self.gate = threading.Event()
        self.gate.set()

    def execute(self, d):
        self.gate.wait()
        time.sleep(self.delay)
        self.calls.append(("one", d.table))
        return f"one:{d.table}:{d.limit}"

    def execute_batch(self, ds):
        self.gate.wait()
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


def test_percentile():
    v = list(range(1, 101))
    assert serve.percentile(v, 50) == 50 and serve.percentile(v, 99) == 99 and serve.percentile(v, 100) == 100
    assert serve.percentile([5.0], 99) == 5.0
