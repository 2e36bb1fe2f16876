"""DataFrame API, the GPU QueryExecutor and QueryResult.

Mirrors the reference's builder (src/client.ts:176-371, 491-546, 876-901),
QueryExecutor (src/client.ts:1280-1304) and QueryMode.local()/table()/sql()
(src/local.ts:23-70); the executor drives the device library through its
C-ABI instead of LocalExecutor's page-by-page JS/WASM pipeline.
"""
from __future__ import annotations

import ctypes as C
import time
import weakref
from dataclasses import dataclass, field, replace
from typing import Any, Iterable

import numpy as np

from . import _native as N
from . import qmcb
from .errors import QueryModeError
from .operators import GpuQueryOperator, drain_pipeline
from .plan import AggregateOp, FilterOp, JoinDescriptor, QueryDescriptor, encode_plan

_NP_TO_QM = {
    np.dtype(np.float64): N.QM_F64, np.dtype(np.int64): N.QM_I64, np.dtype(np.int32): N.QM_I32,
    np.dtype(np.float32): N.QM_F32, np.dtype(np.int8): N.QM_I8, np.dtype(np.int16): N.QM_I16,
    np.dtype(np.bool_): N.QM_BOOL,
}


class QueryResult:
    """src/types.ts:262-295 (the fields this path fills). The result arrives
    as QMCB (``columnarData``); ``rows`` materialises one dict per row on first
    access only, so a large result is never turned into Python objects unless
    a caller asks for them."""

    def __init__(self, rowCount: int, columns: list[str], columnarData: bytes | memoryview | None = None,
                 durationMs: float = 0.0, stats: dict | None = None, rows: list[dict] | None = None,
                 bytesRead: int = 0, pagesSkipped: int = 0):
        self.rowCount = rowCount
        self.columns = columns
        self.columnarData = columnarData
        self.durationMs = durationMs
        self.stats = stats or {}
        self.bytesRead = bytesRead
        self.pagesSkipped = pagesSkipped
        self._rows = rows

    @property
    def rows(self) -> list[dict]:
        if self._rows is None:
            self._rows = qmcb.to_rows(self.batch()) if self.columnarData is not None else []
        return self._rows

    def batch(self) -> qmcb.ColumnarBatch:
        return qmcb.decode(self.columnarData)

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"QueryResult(rowCount={self.rowCount}, columns={self.columns})"


def _qmcb_header(data: bytes) -> tuple[int, list[str]]:
    """Row count and column names from a QMCB header (src/columnar.ts:438-470)."""
    import struct
    magic, rows, ncols = struct.unpack_from("<IIH", data, 0)
    if magic != qmcb.QMCB_MAGIC:
        raise QueryModeError("INVALID_FORMAT", "device result is not QMCB")
    pos, names = 10, []
    for _ in range(ncols):
        (nl,) = struct.unpack_from("<H", data, pos)
        names.append(bytes(data[pos + 2:pos + 2 + nl]).decode("utf-8"))
        pos += 2 + nl + 2
    return rows, names


class Device:
    """One GPU context (qm_init): resident tables + an HBM budget."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = N.load()
        h = C.c_void_p()
        N.check(self.lib.qm_init(device, hbm_budget_bytes, C.byref(h)))
        self.handle = h
        self.tables: dict[str, GpuTable] = {}

    def info(self) -> dict:
        name = C.create_string_buffer(128)
        cus, budget, used = C.c_uint32(), C.c_uint64(), C.c_uint64()
        N.check(self.lib.qm_device_info(self.handle, name, 128, C.byref(cus), C.byref(budget), C.byref(used)))
        return {"arch": name.value.decode(), "cus": cus.value, "budget": budget.value, "used": used.value}

    # ---- planner tuning knobs (qm_tune_set / qm_tune_get) ----
    def tune_get(self, name: str) -> float:
        v = C.c_double()
        N.check(self.lib.qm_tune_get(self.handle, name.encode(), C.byref(v)))
        return v.value

    def tune_set(self, name: str, value: float) -> None:
        N.check(self.lib.qm_tune_set(self.handle, name.encode(), float(value)))

    def register(self, name: str, columns: dict[str, np.ndarray],
                 nulls: dict[str, np.ndarray] | None = None) -> "GpuTable":
        """Upload decoded columns once; they stay resident in HBM."""
        nulls = nulls or {}
        arrays, descs, keep = [], [], []
        nrows = None
        for cname, arr in columns.items():
            a = np.ascontiguousarray(arr)
            if a.dtype not in _NP_TO_QM:
                raise QueryModeError("UNSUPPORTED", f"column {cname}: dtype {a.dtype} is not on the device path")
            if nrows is None:
                nrows = len(a)
            elif len(a) != nrows:
                raise QueryModeError("SCHEMA_MISMATCH", f"column {cname} has {len(a)} rows, expected {nrows}")
            if a.dtype == np.bool_:
                a = a.astype(np.uint8)
            arrays.append(a)
            nb = None
            if cname in nulls and nulls[cname] is not None:
                nb = qmcb.pack_bits(np.asarray(nulls[cname], dtype=bool))
                keep.append(nb)
            d = N.ColumnDesc(cname.encode(), _NP_TO_QM[np.dtype(arr.dtype)], a.ctypes.data_as(C.c_void_p),
                             nb.ctypes.data_as(C.c_void_p) if nb is not None else None)
            descs.append(d)
        arr_t = (N.ColumnDesc * len(descs))(*descs)
        h = C.c_void_p()
        N.check(self.lib.qm_table_create(self.handle, name.encode(), nrows or 0, len(descs), arr_t, C.byref(h)))
        t = GpuTable(self, name, h, nrows or 0, list(columns.keys()))
        if name in self.tables:
            self.tables[name].drop()
        self.tables[name] = t
        return t

    def register_qmcb(self, name: str, data: bytes) -> "GpuTable":
        b = qmcb.decode(data)
        if b is None:
            raise QueryModeError("INVALID_FORMAT", "not a QMCB buffer")
        cols, nulls = {}, {}
        for c in b.columns:
            if c.dtype not in (qmcb.DTYPE_F64, qmcb.DTYPE_I64, qmcb.DTYPE_I32, qmcb.DTYPE_F32, qmcb.DTYPE_BOOL):
                raise QueryModeError("UNSUPPORTED", f"column {c.name}: QMCB dtype {c.dtype} not on the device path")
            cols[c.name] = c.data
            if c.null_bitmap is not None:
                nulls[c.name] = qmcb.unpack_bits(c.null_bitmap, b.row_count)
        return self.register(name, cols, nulls)

    def register_parquet(self, name: str, path: str, columns: list[str] | None = None) -> "GpuTable":
        """Host-side Parquet decode (decode_parquet), then one upload."""
        cols, nulls = decode_parquet(path, columns)
        return self.register(name, cols, nulls)

    def close(self) -> None:
        for t in list(self.tables.values()):
            t.drop()
        if self.handle:
            self.lib.qm_shutdown(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


_PQ_TYPES = {"double": np.float64, "float": np.float32, "int64": np.int64, "int32": np.int32, "int16": np.int16,
             "int8": np.int8, "bool": np.bool_, "uint8": np.int16, "uint16": np.int32, "uint32": np.int64}


def decode_parquet(path: str, columns: list[str] | None = None) -> tuple[dict, dict]:
    """Parquet pages -> typed host columns and NULL masks (the host decode of
    src/parquet-decode.ts:280-350, by pyarrow): numeric and boolean columns;
    unsigned widths widen to the next signed width; NULL slots hold 0.
    Other physical types raise UNSUPPORTED."""
    import pyarrow.parquet as pq
    tbl = pq.read_table(path, columns=columns)
    cols, nulls = {}, {}
    for cname in tbl.column_names:
        col = tbl.column(cname).combine_chunks()
        t = str(col.type)
        if t not in _PQ_TYPES:
            raise QueryModeError("UNSUPPORTED", f"column {cname}: Parquet type {t} is not on the device path")
        mask = np.asarray(col.is_null(), dtype=bool) if col.null_count else None
        arr = col.fill_null(False if t == "bool" else 0).to_numpy(zero_copy_only=False) if col.null_count \
            else col.to_numpy(zero_copy_only=False)
        cols[cname] = np.ascontiguousarray(np.asarray(arr).astype(_PQ_TYPES[t], copy=False))
        if mask is not None:
            nulls[cname] = mask
    return cols, nulls


class GpuTable:
    def __init__(self, dev: Device, name: str, handle, rows: int, columns: list[str]):
        self.dev, self.name, self.handle, self.rows, self.columns = dev, name, handle, rows, columns

    def drop(self) -> None:
        if self.handle:
            self.dev.lib.qm_table_drop(self.handle)
            self.handle = None
            self.dev.tables.pop(self.name, None)


class _Plan:
    def __init__(self, dev: Device, desc: QueryDescriptor):
        self.dev = dev
        blob = encode_plan(desc)
        h = C.c_void_p()
        N.check(dev.lib.qm_plan_compile(dev.handle, blob, len(blob), C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self.dev.lib.qm_plan_close(self.handle)
            self.handle = None


def pinned_qmcb(lib, rh, row0: int = 0, nrows: int = 2 ** 63) -> memoryview:
    """QMCB of rows [row0, row0 + nrows) of a device result, encoded on the
    device into pinned host memory (qm_result_qmcb_pinned) and wrapped without
    a copy; the buffer goes back to the library when the last view is gone."""
    ptr, size = C.c_void_p(), C.c_uint64()
    N.check(lib.qm_result_qmcb_pinned(rh, row0, nrows, C.byref(ptr), C.byref(size)))
    arr = (C.c_uint8 * size.value).from_address(ptr.value)
    weakref.finalize(arr, lib.qm_qmcb_release, C.c_void_p(ptr.value))
    return memoryview(arr).cast("B")


def _read_result(lib, rh) -> tuple[memoryview, dict]:
    try:
        data = pinned_qmcb(lib, rh)
        st = N.Stats()
        N.check(lib.qm_result_stats(rh, C.byref(st)))
        return data, st.as_dict()
    finally:
        lib.qm_result_close(rh)


class DeviceResult:
    """A result kept in device memory; ``fetch(row0, n)`` encodes one range of
    it as QMCB (the pull of a cursor). Closed with ``close()``."""

    def __init__(self, lib, handle):
        self._lib, self._h = lib, handle
        n = C.c_uint64()
        N.check(lib.qm_result_rows(handle, C.byref(n)))
        self.rows = int(n.value)

    def fetch(self, row0: int, nrows: int) -> memoryview:
        if not self._h:
            raise QueryModeError("INVALID_ARGUMENT", "the result is closed")
        return pinned_qmcb(self._lib, self._h, row0, nrows)

    def stats(self) -> dict:
        st = N.Stats()
        N.check(self._lib.qm_result_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if self._h:
            self._lib.qm_result_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def _batch_results(lib, outs, n: int, t0: float) -> list[QueryResult]:
    res = []
    for i in range(n):
        data, st = _read_result(lib, C.c_void_p(outs[i]))
        nrow, names = _qmcb_header(data)
        res.append(QueryResult(rowCount=nrow, columns=names,
                               columnarData=data, durationMs=(time.perf_counter() - t0) * 1e3, stats=st,
                               pagesSkipped=int(st.get("pages_skipped", 0))))
    return res


class PendingBatch:
    """A batch running on the library's worker thread (qm_batch_submit). The
    plans and the table stay referenced until the batch is closed."""

    def __init__(self, lib, handle, plans, table, t0: float):
        self._lib, self._h, self._plans, self._table, self._t0 = lib, handle, plans, table, t0
        self.size = len(plans)

    def done(self) -> bool:
        d = C.c_int()
        N.check(self._lib.qm_batch_poll(self._h, C.byref(d)))
        return bool(d.value)

    def wait(self) -> list[QueryResult]:
        if not self._h:
            raise QueryModeError("INVALID_ARGUMENT", "the batch is closed")
        outs = (C.c_void_p * self.size)()
        try:
            N.check(self._lib.qm_batch_wait(self._h, outs))
            return _batch_results(self._lib, outs, self.size, self._t0)
        finally:
            self.close()

    def close(self) -> None:
        """Wait for the batch if it still runs, then release it and its plans."""
        if self._h:
            self._lib.qm_batch_close(self._h)
            self._h = None
            for p in self._plans:
                p.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class PreparedBatch:
    """A batch captured into a HIP graph (qm_batch_prepare): ``run()`` replays
    every pass over the table with one host launch. The results of a run are
    read out (QMCB) before the next run reuses the prepared buffers."""

    def __init__(self, dev: "Device", handle, table: "GpuTable", size: int):
        self._dev, self._h, self._table, self.size = dev, handle, table, size

    def info(self) -> dict:
        n, nodes, b = C.c_uint32(), C.c_uint64(), C.c_uint64()
        N.check(self._dev.lib.qm_prepared_info(self._h, C.byref(n), C.byref(nodes), C.byref(b)))
        return {"plans": n.value, "graph_nodes": nodes.value, "device_bytes": b.value}

    def run_handles(self, outs) -> None:
        """Replay into `outs` (nplans result handles; the caller closes them)."""
        if not self._h:
            raise QueryModeError("INVALID_ARGUMENT", "the prepared batch is closed")
        N.check(self._dev.lib.qm_prepared_run(self._h, outs))

    def run(self) -> list[QueryResult]:
        outs = (C.c_void_p * self.size)()
        t0 = time.perf_counter()
        self.run_handles(outs)
        return _batch_results(self._dev.lib, outs, self.size, t0)

    def close(self) -> None:
        if self._h:
            N.check(self._dev.lib.qm_prepared_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class GpuExecutor:
    """QueryExecutor (src/client.ts:1280-1304) backed by the device library."""

    def __init__(self, dev: Device):
        self.dev = dev

    def _table(self, name: str) -> GpuTable:
        t = self.dev.tables.get(name)
        if t is None or not t.handle:
            raise QueryModeError("TABLE_NOT_FOUND", f"Table not found: {name}")
        return t

    def run_device(self, desc: QueryDescriptor) -> DeviceResult:
        """Execute and keep the result on the device (read it with fetch)."""
        left = self._table(desc.table)
        right = self._table(desc.join.right.table) if desc.join is not None else None
        plan = _Plan(self.dev, desc)
        try:
            rh = C.c_void_p()
            N.check(self.dev.lib.qm_exec(plan.handle, left.handle, right.handle if right else None, C.byref(rh)))
        finally:
            plan.close()
        return DeviceResult(self.dev.lib, rh)

    def cursor(self, desc: QueryDescriptor, batch_size: int = 1000):
        """QueryExecutor.cursor (src/client.ts:1300): the result stays in device
        memory and each batch of `batch_size` rows is encoded and handed over
        when it is pulled. Yields row lists."""
        op = GpuQueryOperator(lambda: self.run_device(desc), batch_size)
        try:
            while True:
                b = op.next()
                if b is None:
                    break
                yield b
        finally:
            op.close()

    def run_qmcb(self, desc: QueryDescriptor) -> tuple[bytes, dict]:
        left = self._table(desc.table)
        right = self._table(desc.join.right.table) if desc.join is not None else None
        plan = _Plan(self.dev, desc)
        try:
            rh = C.c_void_p()
            N.check(self.dev.lib.qm_exec(plan.handle, left.handle, right.handle if right else None, C.byref(rh)))
        finally:
            plan.close()
        return _read_result(self.dev.lib, rh)

    def execute(self, desc: QueryDescriptor) -> QueryResult:
        t0 = time.perf_counter()
        data, st = self.run_qmcb(desc)
        rows, names = _qmcb_header(data)
        return QueryResult(rowCount=rows, columns=names, columnarData=data,
                           durationMs=(time.perf_counter() - t0) * 1e3, stats=st,
                           pagesSkipped=int(st.get("pages_skipped", 0)))

    def _batch_table(self, descs: list[QueryDescriptor]) -> GpuTable:
        tname = descs[0].table
        if any(d.table != tname for d in descs):
            raise QueryModeError("INVALID_ARGUMENT", "a batch runs over one table")
        return self._table(tname)

    def execute_batch(self, descs: list[QueryDescriptor]) -> list[QueryResult]:
        """Many queries over one resident table in fused launches (qm_batch_exec)."""
        if not descs:
            return []
        table = self._batch_table(descs)
        plans = [_Plan(self.dev, d) for d in descs]
        try:
            arr = (C.c_void_p * len(plans))(*[p.handle for p in plans])
            outs = (C.c_void_p * len(plans))()
            t0 = time.perf_counter()
            N.check(self.dev.lib.qm_batch_exec(self.dev.handle, arr, len(plans), table.handle, outs))
        finally:
            for p in plans:
                p.close()
        return _batch_results(self.dev.lib, outs, len(descs), t0)

    def prepare_batch(self, descs: list[QueryDescriptor]) -> PreparedBatch:
        """Run the batch once and capture its launches into a HIP graph
        (qm_batch_prepare); ``PreparedBatch.run()`` replays it."""
        if not descs:
            raise QueryModeError("INVALID_ARGUMENT", "an empty batch")
        table = self._batch_table(descs)
        plans = [_Plan(self.dev, d) for d in descs]
        try:
            arr = (C.c_void_p * len(plans))(*[p.handle for p in plans])
            h = C.c_void_p()
            N.check(self.dev.lib.qm_batch_prepare(self.dev.handle, arr, len(plans), table.handle, C.byref(h)))
        finally:
            for p in plans:
                p.close()
        return PreparedBatch(self.dev, h, table, len(descs))

    def submit_batch(self, descs: list[QueryDescriptor]) -> "PendingBatch":
        """Start a batch on the library's worker thread (qm_batch_submit) and
        return at once; ``PendingBatch.wait()`` gives the results."""
        if not descs:
            raise QueryModeError("INVALID_ARGUMENT", "an empty batch")
        table = self._batch_table(descs)
        plans = [_Plan(self.dev, d) for d in descs]
        try:
            arr = (C.c_void_p * len(plans))(*[p.handle for p in plans])
            h = C.c_void_p()
            t0 = time.perf_counter()
            N.check(self.dev.lib.qm_batch_submit(self.dev.handle, arr, len(plans), table.handle, C.byref(h)))
        except BaseException:
            for p in plans:
                p.close()
            raise
        return PendingBatch(self.dev.lib, h, plans, table, t0)

    def explain(self, desc: QueryDescriptor) -> dict:
        ops = ["GpuScan(filter)" if desc.filters or desc.filterGroups else "GpuScan"]
        if desc.join is not None:
            ops.append(f"GpuHashJoin(build={desc.join.right.table}, {desc.join.leftKey}={desc.join.rightKey})")
        if desc.aggregates:
            ops.append(f"GpuHashAggregate(groupBy={desc.groupBy})")
        if desc.sortColumn:
            ops.append("GpuTopK" if desc.limit is not None else "GpuSort")
        return {"table": desc.table, "operators": ops, "device": self.dev.info()}

    def to_operator(self, desc: QueryDescriptor) -> GpuQueryOperator:
        return GpuQueryOperator(lambda: self.run_device(desc))

    # reference-style camelCase alias
    toOperator = to_operator

    def count(self, desc: QueryDescriptor) -> int:
        d = replace(desc, aggregates=[AggregateOp("count", "*", "count")], groupBy=[], sortColumn=None,
                    limit=None, offset=None, projections=[])
        return int(self.execute(d).rows[0]["count"])


def _parse_product(expr) -> tuple[str, str]:
    if isinstance(expr, (tuple, list)) and len(expr) == 3 and expr[0] in ("*", "multiply"):
        return expr[1], expr[2]
    if isinstance(expr, str):
        s = expr.strip()
        if s.startswith("multiply(") and s.endswith(")"):
            a, b = s[9:-1].split(",")
            return a.strip(), b.strip()
        if "*" in s:
            a, b = s.split("*")
            return a.strip(), b.strip()
    raise QueryModeError("UNSUPPORTED", "computed columns on the device are products of two columns: 'a * b'")


class DataFrame:
    """Immutable lazy builder (src/client.ts:88-921)."""

    def __init__(self, table: str, executor: GpuExecutor, desc: QueryDescriptor | None = None):
        self._executor = executor
        self._desc = desc or QueryDescriptor(table=table)

    def _derive(self, **kw) -> "DataFrame":
        return DataFrame(self._desc.table, self._executor, replace(self._desc, **kw))

    # ---- filters (src/client.ts:176-230) ----
    def filter(self, column: str, op: str, value: Any = None) -> "DataFrame":
        return self._derive(filters=self._desc.filters + [FilterOp(column, op, value)])

    where = filter

    def whereBetween(self, column: str, low, high) -> "DataFrame":
        return self.filter(column, "between", [low, high])

    def whereNotBetween(self, column: str, low, high) -> "DataFrame":
        return self.filter(column, "not_between", [low, high])

    def whereIn(self, column: str, values: Iterable) -> "DataFrame":
        return self.filter(column, "in", list(values))

    def whereNotIn(self, column: str, values: Iterable) -> "DataFrame":
        return self.filter(column, "not_in", list(values))

    def whereNull(self, column: str) -> "DataFrame":
        return self.filter(column, "is_null")

    def whereNotNull(self, column: str) -> "DataFrame":
        return self.filter(column, "is_not_null")

    def whereOr(self, *groups) -> "DataFrame":
        gs = [[f if isinstance(f, FilterOp) else FilterOp(*f) for f in g] for g in groups]
        return self._derive(filterGroups=self._desc.filterGroups + gs)

    # ---- shape ----
    def select(self, *columns: str) -> "DataFrame":
        return self._derive(projections=list(columns))

    def sort(self, column: str, direction: str = "asc") -> "DataFrame":
        return self._derive(sortColumn=column, sortDirection=direction)

    orderBy = sort

    def sortBy(self, *keys) -> "DataFrame":
        """Multi-column ORDER BY: keys are column names or (column, "asc" | "desc")
        pairs; stable, NULLs last (the SQL wrapper's order, src/sql/executor.ts:114-141)."""
        ks = [(k, "asc") if isinstance(k, str) else (k[0], k[1]) for k in keys]
        return self._derive(orderBy=ks, sortColumn=None, sortDirection=None)

    def distinct(self, *columns: str) -> "DataFrame":
        """DISTINCT over `columns` (none: every output column), src/client.ts:410-413."""
        return self._derive(distinct=list(columns))

    def having(self, expr: tuple) -> "DataFrame":
        """HAVING over the aggregate output: a tree of querymode_amd.plan.HAVING_NODES."""
        return self._derive(having=expr)

    def limit(self, n: int) -> "DataFrame":
        return self._derive(limit=n)

    def offset(self, n: int) -> "DataFrame":
        return self._derive(offset=n)

    def aggregate(self, fn: str, column: str, alias: str | None = None) -> "DataFrame":
        return self._derive(aggregates=self._desc.aggregates + [AggregateOp(fn, column, alias)])

    def groupBy(self, *columns: str) -> "DataFrame":
        return self._derive(groupBy=list(columns))

    def join(self, other: "DataFrame", keys: dict, type: str = "inner") -> "DataFrame":
        return self._derive(join=JoinDescriptor(other.toDescriptor(), keys["left"], keys["right"], type))

    def computed(self, alias: str, expr) -> "DataFrame":
        a, b = _parse_product(expr)
        return self._derive(computedColumns=self._desc.computedColumns + [(alias, a, b)])

    # ---- terminal ----
    def toDescriptor(self) -> QueryDescriptor:
        return self._desc

    def collect(self) -> QueryResult:
        return self._executor.execute(self._desc)

    exec = collect

    def explain(self) -> dict:
        return self._executor.explain(self._desc)

    def toOperator(self) -> GpuQueryOperator:
        return self._executor.to_operator(self._desc)

    def count(self) -> int:
        return self._executor.count(self._desc)

    def first(self) -> dict | None:
        r = self.limit(1).collect()
        return r.rows[0] if r.rows else None

    def cursor(self, batch_size: int = 4096):
        op = GpuQueryOperator(lambda: self._executor.run_device(self._desc), batch_size)
        try:
            while True:
                b = op.next()
                if b is None:
                    break
                yield b
        finally:
            op.close()


class QueryMode:
    """Entry point (src/local.ts:23-70): QueryMode.gpu() instead of QueryMode.local()."""

    def __init__(self, dev: Device):
        self.device = dev
        self.executor = GpuExecutor(dev)

    @staticmethod
    def gpu(device: int = 0, hbm_budget_bytes: int = 0) -> "QueryMode":
        return QueryMode(Device(device, hbm_budget_bytes))

    def register_table(self, name: str, columns: dict, nulls: dict | None = None) -> GpuTable:
        return self.device.register(name, columns, nulls)

    def register_qmcb(self, name: str, data: bytes) -> GpuTable:
        return self.device.register_qmcb(name, data)

    def register_parquet(self, name: str, path: str) -> GpuTable:
        return self.device.register_parquet(name, path)

    def table(self, name: str) -> DataFrame:
        return DataFrame(name, self.executor)

    def sql(self, query: str) -> DataFrame:
        from .sql import sql_to_descriptor
        return DataFrame("", self.executor, sql_to_descriptor(query))

    def close(self) -> None:
        self.device.close()


__all__ = ["QueryMode", "DataFrame", "GpuExecutor", "Device", "GpuTable", "QueryResult", "drain_pipeline"]
