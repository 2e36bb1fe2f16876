"""Pull-based operator interface (src/operators.ts:51-90, 3155-3179).

``Operator.next() -> RowBatch | None`` (a list of row dicts),
``Operator.next_columnar() -> ColumnarBatch | None`` and ``close()``; the
terminal GPU operator runs the whole hot path on the device at its first pull
and hands the result out in batches.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Protocol

import numpy as np

from . import qmcb

RowBatch = list  # list[dict]


@dataclass
class PipelineBatch:
    """ColumnarBatch of the pipeline (src/operators.ts:57-64): columns by name,
    row count and an optional selection vector of active row indices."""
    columns: dict
    rowCount: int
    selection: np.ndarray | None = None


class Operator(Protocol):
    def next(self) -> RowBatch | None: ...
    def next_columnar(self) -> PipelineBatch | None: ...
    def close(self) -> None: ...


def materialize_rows(batch: PipelineBatch) -> list[dict]:
    """materializeRows (src/operators.ts:78-90)."""
    idx = batch.selection if batch.selection is not None else np.arange(batch.rowCount)
    rows = []
    for i in idx:
        row = {}
        for name, vals in batch.columns.items():
            v = vals[int(i)]
            row[name] = None if v is None else (v.item() if hasattr(v, "item") else v)
        rows.append(row)
    return rows


class GpuQueryOperator:
    """Terminal operator over a device result. The device work happens on the
    first pull (like AggregateOperator / TopKOperator draining upstream at their
    first next(), src/operators.ts:1656-1707, 1819-1824); the result then stays
    in device memory and every pull encodes the next `batch_size` rows as QMCB
    on the device and hands them over (no host copy of the whole result).

    `run` returns either a device result (an object with ``rows``,
    ``fetch(row0, n)`` and ``close()``) or QMCB bytes of the whole result."""

    def __init__(self, run, batch_size: int = 4096):
        self._run = run
        self._batch_size = max(1, int(batch_size))
        self._src = None
        self._batch: qmcb.ColumnarBatch | None = None
        self._pos = 0
        self._closed = False

    def _ensure(self) -> None:
        if self._src is None and self._batch is None:
            r = self._run()
            if hasattr(r, "fetch"):
                self._src = r
            else:
                self._batch = qmcb.decode(r)

    def _slice(self) -> qmcb.ColumnarBatch | None:
        if self._src is not None:
            if self._pos >= self._src.rows:
                return None
            b = qmcb.decode(self._src.fetch(self._pos, self._batch_size))
            self._pos += b.row_count
            return b
        b = self._batch
        if self._pos >= b.row_count:
            return None
        lo, hi = self._pos, min(self._pos + self._batch_size, b.row_count)
        self._pos = hi
        cols = []
        for c in b.columns:
            nb = None
            if c.null_bitmap is not None:
                nb = qmcb.pack_bits(qmcb.unpack_bits(c.null_bitmap, b.row_count)[lo:hi])
            cols.append(qmcb.ColumnarColumn(c.name, c.dtype, c.data[lo:hi] if c.data is not None else None, hi - lo, nb))
        return qmcb.ColumnarBatch(cols, hi - lo)

    def next_columnar(self) -> PipelineBatch | None:
        if self._closed:
            return None
        self._ensure()
        b = self._slice()
        if b is None:
            return None
        cols = {}
        for c in b.columns:
            vals = list(c.data) if c.data is not None else [None] * b.row_count
            if c.null_bitmap is not None:
                nulls = qmcb.unpack_bits(c.null_bitmap, b.row_count)
                vals = [None if nulls[i] else v for i, v in enumerate(vals)]
            cols[c.name] = vals
        return PipelineBatch(cols, b.row_count)

    def next(self) -> RowBatch | None:
        b = self.next_columnar()
        return None if b is None else materialize_rows(b)

    def close(self) -> None:
        self._closed = True
        self._batch = None
        if self._src is not None:
            self._src.close()
            self._src = None


def drain_pipeline(op) -> list[dict]:
    """drainPipeline (src/operators.ts:3155-3179): prefer nextColumnar, close in finally."""
    rows: list[dict] = []
    try:
        if hasattr(op, "next_columnar"):
            while True:
                b = op.next_columnar()
                if b is None:
                    break
                rows.extend(materialize_rows(b))
        else:
            while True:
                b = op.next()
                if b is None:
                    break
                rows.extend(b)
    finally:
        op.close()
    return rows
