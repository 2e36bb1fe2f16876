"""Serving loop: continuous batching of concurrent requests over resident tables.

Requests (QueryDescriptors) arrive from any thread through ``Server.submit``.
A pool of worker threads, one per execution lane of the device context (the
library runs up to that many calls at once, each on a stream of its own),
takes work as soon as a lane is free:

* **Admission.** Each request's device memory beyond the resident tables is
  estimated when it is submitted, from the same binding the execution uses
  (``qm_plan_device_bytes``). A worker admits waiting requests in arrival
  order while their bytes, plus those of the work already on the device, fit
  the HBM budget (less a reserve); a request that cannot fit waits until
  running work finishes and frees memory, and fails with ``MEMORY_EXCEEDED``
  only when it cannot fit with nothing else running.
* **Continuous batching.** Aggregate requests over one table are fused into
  one ``qm_batch_exec`` launch set (one shared pass over the table for all of
  them, up to ``max_slots``). While such a batch runs, new aggregate requests
  over the same table accumulate and go out together as the next batch the
  moment it finishes; joins, top-K and row queries never wait for it: any
  free lane takes them at once, so their kernels run beside the batch's on
  the device.
* **Latency.** Each request records arrival, start and completion times;
  ``latency_ms`` is completion minus arrival.

Tables, group-table buffers (the context's pool) and join tables stay
device-resident across requests. A result is the same whether the request ran
alone or in a mix: batching and overlap change which launch computes it and
when, not the values (the aggregates are exact and order-independent).
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

LANES = 4  # execution lanes of a device context (streams the library runs calls on)


@dataclass
class Request:
    desc: QueryDescriptor
    arrival: float = field(default_factory=time.perf_counter)
    start: float = 0.0
    done: float = 0.0
    result: QueryResult | None = None
    error: Exception | None = None
    device_bytes: int = 0
    batch_size: int = 0   # requests in the launch that served this one
    overlap: int = 0      # most other launches on the device while this one ran
    _event: threading.Event = field(default_factory=threading.Event, repr=False)

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3

    @property
    def fused(self) -> bool:
        d = self.desc
        return bool(d.aggregates) and d.join is None

    def wait(self, timeout: float | None = None) -> QueryResult:
        if not self._event.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.result


class _Unit:
    """One launch on a lane: a fused aggregate batch or a single request."""

    def __init__(self, reqs: list[Request], table: str | None):
        self.reqs = reqs
        self.table = table            # set for fused aggregate batches
        self.bytes = sum(r.device_bytes for r in reqs)
        self.peak_others = 0


class Server:
    """Continuous-batching request loop over one device context."""

    def __init__(self, qm: QueryMode, max_slots: int = 256, hbm_reserve: float = 0.05, lanes: int = LANES):
        self.qm = qm
        self.ex = qm.executor
        self.dev = qm.device
        self.max_slots = max_slots
        self.hbm_reserve = hbm_reserve
        self._q: deque[Request] = deque()
        self._cv = threading.Condition()
        self._stop = False
        self._running: list[_Unit] = []
        self._inflight = 0             # estimated device bytes of the running units
        # statistics
        self.launches = 0              # units sent to the device
        self.batched = 0               # requests served by fused batches of >= 2
        self.max_overlap = 0           # most units on the device at once
        self._busy_t = 0.0             # time with >= 1 unit running
        self._overlap_t = 0.0          # time with >= 2 units running
        self._mark = time.perf_counter()
        self._workers = [threading.Thread(target=self._worker, name=f"qm-serve-{i}", daemon=True)
                         for i in range(max(1, lanes))]
        for w in self._workers:
            w.start()

    # kept for callers that count scheduling rounds
    @property
    def cycles(self) -> int:
        return self.launches

    @property
    def overlap_share(self) -> float:
        """Share of the busy time during which two or more launches ran at once."""
        return self._overlap_t / self._busy_t if self._busy_t > 0 else 0.0

    # ---- client side ----
    def submit(self, desc: QueryDescriptor) -> Request:
        r = Request(desc)
        try:
            r.device_bytes = max(1, self._estimate(r))
        except QueryModeError as e:  # unknown table / column: fails alone, at once
            r.error, r.done = e, time.perf_counter()
            r._event.set()
            return r
        with self._cv:
            if self._stop:
                raise QueryModeError("QUERY_FAILED", "server is closed")
            self._q.append(r)
            self._cv.notify()
        return r

    def close(self) -> None:
        """Finish every submitted request, then stop the workers."""
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()

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

    def _tick(self) -> None:
        """Accumulate busy / overlap time up to now (caller holds the lock)."""
        now = time.perf_counter()
        dt = now - self._mark
        self._mark = now
        if len(self._running) >= 1:
            self._busy_t += dt
        if len(self._running) >= 2:
            self._overlap_t += dt

    def _fail(self, r: Request, err: Exception) -> None:
        r.error, r.done = err, time.perf_counter()
        r._event.set()

    def _pick(self) -> _Unit | None:
        """The next unit to run, or None (caller holds the lock). Requests are
        considered in arrival order; aggregates over a table whose batch is
        running stay queued for the next batch."""
        if not self._q:
            return None
        info = self.dev.info()
        free = int(info["budget"] * (1.0 - self.hbm_reserve)) - int(info["used"]) - self._inflight
        busy_tables = {u.table for u in self._running if u.table is not None}
        for r in list(self._q):
            if r.fused and r.desc.table in busy_tables:
                continue
            if r.device_bytes > free:
                if not self._running:  # cannot fit even alone
                    self._q.remove(r)
                    self._fail(r, QueryModeError(
                        "MEMORY_EXCEEDED", f"request needs {r.device_bytes} device bytes, {free} free"))
                    continue
                return None  # wait for running work to free memory (keeps arrival order)
            if not r.fused:
                self._q.remove(r)
                return _Unit([r], None)
            take, total = [], 0
            for q in list(self._q):
                if len(take) >= self.max_slots:
                    break
                if not q.fused or q.desc.table != r.desc.table:
                    continue
                if take and total + q.device_bytes > free:
                    break
                self._q.remove(q)
                take.append(q)
                total += q.device_bytes
            return _Unit(take, r.desc.table)
        return None

    def _worker(self) -> None:
        while True:
            with self._cv:
                while True:
                    unit = self._pick()
                    if unit is not None:
                        break
                    if self._stop and not self._q and not self._running:
                        self._cv.notify_all()
                        return
                    self._cv.wait(0.05 if self._stop else None)
                self._tick()
                self._running.append(unit)
                for u in self._running:
                    u.peak_others = max(u.peak_others, len(self._running) - 1)
                self._inflight += unit.bytes
                self.launches += 1
                self.max_overlap = max(self.max_overlap, len(self._running))
            self._run(unit)
            with self._cv:
                self._tick()
                self._running.remove(unit)
                self._inflight -= unit.bytes
                self._cv.notify_all()

    def _run(self, unit: _Unit) -> None:
        now = time.perf_counter()
        for r in unit.reqs:
            r.start = now
            r.batch_size = len(unit.reqs)
        if len(unit.reqs) == 1:
            self._run_one(unit.reqs[0], unit)
            return
        try:
            results = self.ex.execute_batch([r.desc for r in unit.reqs])
        except Exception:  # noqa: BLE001 - a failing batch is retried request by request
            for r in unit.reqs:
                self._run_one(r, unit)
            return
        self.batched += len(unit.reqs)
        t = time.perf_counter()
        for r, res in zip(unit.reqs, results):
            r.result, r.done, r.overlap = res, t, unit.peak_others
            r._event.set()

    def _run_one(self, r: Request, unit: _Unit) -> None:
        try:
            r.result = self.ex.execute(r.desc)
        except Exception as e:  # noqa: BLE001 - reported to the request's owner
            r.error = e
        r.done = time.perf_counter()
        r.overlap = unit.peak_others
        r._event.set()


def percentile(values: list[float], p: float) -> float:
    """Nearest-rank percentile (p in [0, 100])."""
    if not values:
        return float("nan")
    import math
    v = sorted(values)
    k = max(0, min(len(v) - 1, math.ceil(p / 100.0 * len(v)) - 1))
    return v[k]


__all__ = ["Server", "Request", "percentile", "LANES"]
