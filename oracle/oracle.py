"""Python driver of the CPU oracle (libqmoracle.so) — TEST INFRASTRUCTURE ONLY.

Runs a QueryDescriptor through a CPU restatement of the reference pipeline
(filter -> join -> aggregate -> sort/limit -> project) and returns rows in the
reference's shape. Only tests, __graft_entry__.smoke() and bench.py's CPU
baseline use this module, and only as the checker / the timed baseline.

Function-by-function anchors are in qm_oracle.c. Results:
  * group-by output in first-seen order (Map insertion order, src/partial-agg.ts);
  * join pairs in probe order, then build insertion order;
  * top-K under the declared total order (value, then row id), NULL/NaN last.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

try:
    from . import sql_post as SP
except ImportError:  # loaded as a top-level module
    import sql_post as SP

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libqmoracle.so"

DT = {np.dtype(np.float64): 0, np.dtype(np.int64): 1, np.dtype(np.bool_): 3, np.dtype(np.uint8): 3,
      np.dtype(np.int32): 6, np.dtype(np.float32): 7, np.dtype(np.int8): 16, np.dtype(np.int16): 17}
OPS = {"eq": 0, "neq": 1, "lt": 2, "lte": 3, "gt": 4, "gte": 5, "between": 6, "not_between": 7, "in": 8,
       "not_in": 9, "is_null": 10, "is_not_null": 11}
AGG = {"sum": 0, "avg": 1, "min": 2, "max": 3, "count": 4}
MAX_SAFE = 2 ** 53 - 1


class OcCol(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nulls", C.c_void_p), ("dtype", C.c_int)]


class OcPred(C.Structure):
    _fields_ = [("col", C.c_int), ("op", C.c_int), ("group", C.c_int), ("kind", C.c_int), ("f", C.c_double),
                ("i", C.c_int64), ("n", C.c_int), ("lkinds", C.c_void_p), ("lf", C.c_void_p), ("li", C.c_void_p)]


class OcAgg(C.Structure):
    _fields_ = [("fn", C.c_int), ("col_a", C.c_int), ("col_b", C.c_int)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        u64p, vp = C.POINTER(C.c_uint64), C.c_void_p
        lib.oc_filter.restype = C.c_uint64
        lib.oc_filter.argtypes = [vp, C.c_uint64, vp, C.c_int, C.c_int, C.c_uint32, vp]
        lib.oc_groupby.restype = C.c_int64
        lib.oc_groupby.argtypes = [vp, vp, vp, vp, C.c_uint64, vp, C.c_int, vp, C.c_int, C.c_uint64,
                                   vp, vp, vp, vp, vp, vp]
        lib.oc_join.restype = C.c_int64
        lib.oc_join.argtypes = [vp, vp, C.c_uint64, vp, vp, C.c_uint64, C.c_uint64, vp, vp]
        lib.oc_topk_exact.restype = C.c_uint64
        lib.oc_topk_exact.argtypes = [vp, vp, C.c_uint64, C.c_int, C.c_uint64, vp]
        lib.oc_topk_heap.restype = C.c_uint64
        lib.oc_topk_heap.argtypes = [vp, vp, C.c_uint64, C.c_int, C.c_uint64, vp]
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _lit(v):
    """(kind, f, i): JS number (1) or bigint (2), as the plan encoder types it."""
    from querymode_amd.plan import BigInt
    if isinstance(v, BigInt) or (isinstance(v, int) and not isinstance(v, bool) and abs(v) > MAX_SAFE):
        return 2, 0.0, int(v)
    return 1, float(v), 0


class Table:
    """Host columns (+ optional boolean null masks)."""

    def __init__(self, columns: dict, nulls: dict | None = None):
        self.names = list(columns.keys())
        self.arrays = [np.ascontiguousarray(columns[n]) for n in self.names]
        self.rows = len(self.arrays[0]) if self.arrays else 0
        nulls = nulls or {}
        self.null_bits = []
        for n, a in zip(self.names, self.arrays):
            m = nulls.get(n)
            self.null_bits.append(None if m is None else np.packbits(np.asarray(m, dtype=np.uint8),
                                                                     bitorder="little"))

    def col(self, name: str) -> int:
        return self.names.index(name)

    def occol(self, i: int) -> OcCol:
        a = self.arrays[i]
        return OcCol(_ptr(a), _ptr(self.null_bits[i]), DT[a.dtype])

    def value(self, i: int, r: int):
        nb = self.null_bits[i]
        if nb is not None and (nb[r >> 3] >> (r & 7)) & 1:
            return None
        v = self.arrays[i][r]
        return float(v) if self.arrays[i].dtype.kind == "f" else int(v)


def filter_rows(t: Table, filters, groups=(), page: int = 10000) -> np.ndarray:
    """Ascending row ids passing AND filters and OR groups (oc_filter)."""
    lib = load()
    cols = (OcCol * max(1, len(t.names)))(*[t.occol(i) for i in range(len(t.names))])
    preds, keep = [], []
    for g, fl in [(0, filters)] + [(i + 1, gr) for i, gr in enumerate(groups)]:
        for f in fl:
            p = OcPred(t.col(f.column), OPS[f.op], g, 0, 0.0, 0, 0, None, None, None)
            if f.op in ("between", "not_between", "in", "not_in"):
                vals = list(f.value)
                kinds = np.array([_lit(v)[0] for v in vals] or [0], dtype=np.int32)
                lf = np.array([_lit(v)[1] for v in vals] or [0], dtype=np.float64)
                li = np.array([_lit(v)[2] for v in vals] or [0], dtype=np.int64)
                keep += [kinds, lf, li]
                p.n, p.lkinds, p.lf, p.li = len(vals), _ptr(kinds), _ptr(lf), _ptr(li)
            elif f.op not in ("is_null", "is_not_null"):
                p.kind, p.f, p.i = _lit(f.value)
            preds.append(p)
    parr = (OcPred * max(1, len(preds)))(*preds)
    out = np.empty(max(1, t.rows), dtype=np.uint64)
    n = lib.oc_filter(cols, t.rows, parr, len(preds), len(groups), page, _ptr(out))
    return out[:n].copy()


def _value_cols(space: list, name: str, computed: dict) -> tuple[int, int]:
    if name in computed:
        a, b = computed[name]
        return space.index(a), space.index(b)
    if name.startswith("multiply(") and name.endswith(")"):
        a, b = [s.strip() for s in name[9:-1].split(",")]
        return space.index(a), space.index(b)
    return space.index(name), -1


def run(desc, tables: dict[str, Table]) -> list[dict]:
    """Execute a QueryDescriptor on the CPU oracle; returns rows."""
    lib = load()
    L = tables[desc.table]
    lrows = filter_rows(L, desc.filters, desc.filterGroups)
    R = None
    rrows = None
    # column space: left columns, then right columns (join)
    space = list(L.names)
    side = [0] * len(space)
    cols = [L.occol(i) for i in range(len(L.names))]
    names_out = list(L.names)
    if desc.join is not None:
        R = tables[desc.join.right.table]
        rsel = filter_rows(R, desc.join.right.filters, desc.join.right.filterGroups)
        cap = max(1, len(lrows) * 4 + 16)
        while True:
            ol = np.empty(cap, dtype=np.uint64)
            orr = np.empty(cap, dtype=np.uint64)
            lk, rk = L.occol(L.col(desc.join.leftKey)), R.occol(R.col(desc.join.rightKey))
            m = lib.oc_join(C.byref(lk), _ptr(lrows), len(lrows), C.byref(rk), _ptr(rsel), len(rsel), cap,
                            _ptr(ol), _ptr(orr))
            if m >= 0:
                break
            cap *= 4
        lrows, rrows = ol[:m].copy(), orr[:m].copy()
        for i, n in enumerate(R.names):
            if n == desc.join.rightKey:
                continue
            out_n = ("right_" + n) if n in L.names else n
            space.append(out_n)
            side.append(1)
            cols.append(R.occol(i))
            names_out.append(out_n)
    computed = {a: (x, y) for a, x, y in desc.computedColumns}
    carr = (OcCol * len(cols))(*cols)
    sarr = np.array(side, dtype=np.int32)

    if desc.aggregates:
        keycols = np.array([space.index(g) for g in desc.groupBy] or [0], dtype=np.int32)
        aggs = []
        for a in desc.aggregates:
            if a.column == "*":
                aggs.append(OcAgg(AGG[a.fn], -1, -1))
            else:
                ca, cb = _value_cols(space, a.column, computed)
                aggs.append(OcAgg(AGG[a.fn], ca, cb))
        aarr = (OcAgg * len(aggs))(*aggs)
        n = len(lrows)
        cap = max(16, min(n + 1, 1 << 26))
        nk, na = len(desc.groupBy), len(aggs)
        while True:
            ok = np.empty(cap * max(1, nk), dtype=np.int64)
            okn = np.empty(cap * max(1, nk), dtype=np.uint8)
            oi = np.empty(cap * na, dtype=np.int64)
            of = np.empty(cap * na, dtype=np.float64)
            on = np.empty(cap * na, dtype=np.uint8)
            oint = np.empty(na, dtype=np.uint8)
            g = lib.oc_groupby(carr, _ptr(sarr), _ptr(lrows), _ptr(rrows), n, _ptr(keycols), nk, aarr, na, cap,
                               _ptr(ok), _ptr(okn), _ptr(oi), _ptr(of), _ptr(on), _ptr(oint))
            if g >= 0:
                break
            cap *= 4
        rows = []
        for gi in range(g):
            row = {}
            for k, gname in enumerate(desc.groupBy):
                ci = space.index(gname)
                arr_dt = (L.arrays if side[ci] == 0 else R.arrays)
                src = arr_dt[ci if side[ci] == 0 else (R.names.index(gname if gname in R.names
                                                                    else gname[6:]))]
                if okn[gi * nk + k]:
                    row[gname] = None
                elif src.dtype.kind == "f":
                    row[gname] = float(np.int64(ok[gi * nk + k]).view(np.float64))
                else:
                    row[gname] = int(ok[gi * nk + k])
            for j, a in enumerate(desc.aggregates):
                o = gi * na + j
                if on[o]:
                    row[a.out_name()] = None
                elif oint[j]:
                    row[a.out_name()] = int(oi[o])
                else:
                    row[a.out_name()] = float(of[o])
            rows.append(row)
        return _post_aggregate(rows, desc)

    # row-returning query
    proj = list(desc.projections) if desc.projections else names_out

    def row_at(p_: int, names) -> dict:
        row = {}
        for name in names:
            ci = space.index(name)
            r_ = int(lrows[p_]) if side[ci] == 0 else int(rrows[p_])
            if side[ci] == 0:
                row[name] = L.value(ci, r_)
            else:
                base = name[6:] if name.startswith("right_") and name[6:] in R.names else name
                row[name] = R.value(R.col(base), r_)
        return row

    dist = getattr(desc, "distinct", None)
    order = list(getattr(desc, "orderBy", None) or [])
    if dist is not None or order:
        # SQL post-stages over every selected row in scan order, then OFFSET /
        # LIMIT and the projection (src/sql/executor.ts:63-84, 114-170)
        names = list(proj)
        for c in list(dist or []) + [k for k, _ in order] + ([desc.sortColumn] if desc.sortColumn else []):
            if c not in names:
                names.append(c)
        rows = [row_at(p_, names) for p_ in range(len(lrows))]
        if dist is not None:
            rows = SP.distinct(rows, list(dist) or list(proj))
        if order:
            rows = SP.order_by(rows, order)
            rows = _slice(rows, desc)
        else:
            rows = _sort_limit_rows(rows, desc)
        return [{k: r[k] for k in proj} for r in rows]

    idx = np.arange(len(lrows))
    if desc.sortColumn:
        ci = space.index(desc.sortColumn)
        src_rows = lrows if side[ci] == 0 else rrows
        # sort positions by (value, position) with NULL/NaN last: ties keep scan order
        oc = cols[ci]
        k = len(idx) if desc.limit is None else min(len(idx), int(desc.limit) + int(desc.offset or 0))
        out = np.empty(max(1, k), dtype=np.uint64)
        m = lib.oc_topk_exact(C.byref(oc), _ptr(np.ascontiguousarray(src_rows)), len(src_rows),
                              1 if desc.sortDirection == "desc" else 0, k, _ptr(out))
        # map sorted source row ids back to positions (stable: positions ascending per id)
        pos_of = {}
        for p_, r_ in enumerate(src_rows.tolist()):
            pos_of.setdefault(r_, []).append(p_)
        order = []
        for r_ in out[:m].tolist():
            order.append(pos_of[r_].pop(0))
        idx = np.array(order, dtype=np.int64)
        idx = idx[int(desc.offset or 0):]
    else:
        off = int(desc.offset or 0)
        end = None if desc.limit is None else off + int(desc.limit)
        idx = idx[off:end]
    rows = []
    for p_ in idx.tolist():
        row = {}
        for name in proj:
            ci = space.index(name)
            r_ = int(lrows[p_]) if side[ci] == 0 else int(rrows[p_])
            if side[ci] == 0:
                row[name] = L.value(ci, r_)
            else:
                base = name[6:] if name.startswith("right_") and name[6:] in R.names else name
                row[name] = R.value(R.col(base), r_)
        rows.append(row)
    return rows


def _sort_key(v, desc):
    if v is None or (isinstance(v, float) and math.isnan(v)):
        return (1, 0)
    return (0, -v if desc else v)


def _slice(rows, desc):
    off = int(desc.offset or 0)
    end = None if desc.limit is None else off + int(desc.limit)
    return rows[off:end]


def _sort_limit_rows(rows, desc):
    if desc.sortColumn:
        d = desc.sortDirection == "desc"
        rows = sorted(rows, key=lambda r: _sort_key(r[desc.sortColumn], d))  # stable
    return _slice(rows, desc)


def _post_aggregate(rows, desc):
    """Group rows -> HAVING -> ORDER BY (one column: the descriptor's sort;
    several: the SQL wrapper's comparator) -> OFFSET / LIMIT -> the SELECT
    columns when HAVING added aggregates (src/sql/executor.ts:108-170)."""
    hv = getattr(desc, "having", None)
    if hv is not None:
        rows = SP.having(rows, hv)
    order = list(getattr(desc, "orderBy", None) or [])
    if order:
        rows = _slice(SP.order_by(rows, order), desc)
    else:
        rows = _sort_limit_rows(rows, desc)
    if hv is not None and desc.projections:
        rows = [{k: r[k] for k in desc.projections if k in r} for r in rows]
    return rows


def topk_heap_rows(t: Table, col: str, desc: bool, k: int, rows=None) -> np.ndarray:
    """The reference's heap TopK (timing baseline)."""
    lib = load()
    oc = t.occol(t.col(col))
    n = t.rows if rows is None else len(rows)
    out = np.empty(max(1, k), dtype=np.uint64)
    m = lib.oc_topk_heap(C.byref(oc), _ptr(rows), n, 1 if desc else 0, k, _ptr(out))
    return out[:m].copy()
