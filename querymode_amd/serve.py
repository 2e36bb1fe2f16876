"""Serving loop: continuous batching of concurrent requests over resident tables.

Requests (QueryDescriptors) arrive from any thread through ``Server.submit``.
One scheduler thread owns the device context:

* **Admission.** Each request's device memory beyond the resident tables is
  estimated from the same binding the execution uses
  (``qm_plan_device_bytes``); a cycle admits waiting requests in arrival order
  while the sum stays inside the HBM budget left over by the resident tables,
  up to ``max_slots`` requests.
* **Continuous batching.** While one cycle runs on the GPU, new requests wait
  in the queue and are admitted at the next cycle, as soon as the slots and
  memory of the finished ones are free. Admitted aggregate requests over the
  same table run as one fused batch (``qm_batch_exec``: they share one pass
  over the table); the cycle's short requests (joins, top-K, row queries) run
  first, on the same stream, back to back.
* **Latency.** Each request records arrival, start and completion times;
  ``latency_ms`` is completion minus arrival.

Tables, group-table buffers (the context's pool) and join tables stay
device-resident across requests. A result is the same whether the request ran
alone or in a mix: batching changes which launch computes it, not the values
(the aggregates are exact and order-independent).
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from collections import deque
from dataclasses import dataclass, field

from . import _native as N
from .client import QueryMode, QueryResult, _Plan
from .errors import QueryModeError
from .plan import QueryDescriptor


@dataclass
class Request:
    desc: QueryDescriptor
    arrival: float = field(default_factory=time.perf_counter)
    start: float = 0.0
    done: float = 0.0
    result: QueryResult | None = None
    error: Exception | None = None
    device_bytes: int = 0
    _event: threading.Event = field(default_factory=threading.Event, repr=False)

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3

    def wait(self, timeout: float | None = None) -> QueryResult:
        if not self._event.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.result


class Server:
    """Continuous-batching request loop over one device context."""

    def __init__(self, qm: QueryMode, max_slots: int = 256, hbm_reserve: float = 0.05):
        self.qm = qm
        self.ex = qm.executor
        self.dev = qm.device
        self.max_slots = max_slots
        self.hbm_reserve = hbm_reserve
        self._q: deque[Request] = deque()
        self._cv = threading.Condition()
        self._stop = False
        self.cycles = 0
        self.batched = 0
        self._thread = threading.Thread(target=self._loop, name="qm-serve", daemon=True)
        self._thread.start()

    # ---- client side ----
    def submit(self, desc: QueryDescriptor) -> Request:
        r = Request(desc)
        with self._cv:
            if self._stop:
                raise QueryModeError("QUERY_FAILED", "server is closed")
            self._q.append(r)
            self._cv.notify()
        return r

    def close(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()

    # ---- scheduler ----
    def _estimate(self, r: Request) -> int:
        d = r.desc
        left = self.ex._table(d.table)
        right = self.ex._table(d.join.right.table) if d.join is not None else None
        p = _Plan(self.dev, d)
        try:
            b = C.c_uint64()
            N.check(self.dev.lib.qm_plan_device_bytes(p.handle, left.handle, right.handle if right else None,
                                                      C.byref(b)))
            return int(b.value)
        finally:
            p.close()

    def _admit(self) -> list[Request]:
        info = self.dev.info()
        free = int(info["budget"] * (1.0 - self.hbm_reserve)) - int(info["used"])
        take: list[Request] = []
        total = 0
        with self._cv:
            while self._q and len(take) < self.max_slots:
                r = self._q[0]
                if not r.device_bytes:
                    try:
                        r.device_bytes = max(1, self._estimate(r))
                    except QueryModeError as e:  # unknown table / column: fails alone
                        self._q.popleft()
                        r.error, r.done = e, time.perf_counter()
                        r._event.set()
                        continue
                if take and total + r.device_bytes > free:
                    break  # the rest waits for memory to free up
                if not take and r.device_bytes > free:
                    self._q.popleft()
                    r.error, r.done = QueryModeError(
                        "MEMORY_EXCEEDED", f"request needs {r.device_bytes} device bytes, {free} free"), \
                        time.perf_counter()
                    r._event.set()
                    continue
                self._q.popleft()
                take.append(r)
                total += r.device_bytes
        return take

    def _run_cycle(self, reqs: list[Request]) -> None:
        now = time.perf_counter()
        for r in reqs:
            r.start = now
        # aggregate requests per table: one fused batch each
        fused: dict[str, list[Request]] = {}
        rest: list[Request] = []
        for r in reqs:
            d = r.desc
            if d.aggregates and d.join is None:
                fused.setdefault(d.table, []).append(r)
            else:
                rest.append(r)
        # short requests (top-K, joins, row queries) first, then the fused batches
        for r in rest:
            self._run_one(r)
        for table, group in fused.items():
            try:
                if len(group) == 1:
                    results = [self.ex.execute(group[0].desc)]
                else:
                    results = self.ex.execute_batch([r.desc for r in group])
                    self.batched += len(group)
                t = time.perf_counter()
                for r, res in zip(group, results):
                    r.result, r.done = res, t
                    r._event.set()
            except Exception:  # noqa: BLE001 - a failing batch is retried request by request
                for r in group:
                    self._run_one(r)

    def _run_one(self, r: Request) -> None:
        try:
            r.result = self.ex.execute(r.desc)
        except Exception as e:  # noqa: BLE001 - reported to the request's owner
            r.error = e
        r.done = time.perf_counter()
        r._event.set()

    def _loop(self) -> None:
        while True:
            with self._cv:
                while not self._q and not self._stop:
                    self._cv.wait()
                if self._stop and not self._q:
                    return
            reqs = self._admit()
            if reqs:
                self.cycles += 1
                self._run_cycle(reqs)


def percentile(values: list[float], p: float) -> float:
    """Nearest-rank percentile (p in [0, 100])."""
    if not values:
        return float("nan")
    import math
    v = sorted(values)
    k = max(0, min(len(v) - 1, math.ceil(p / 100.0 * len(v)) - 1))
    return v[k]


__all__ = ["Server", "Request", "percentile"]
