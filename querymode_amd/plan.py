"""QueryDescriptor (src/client.ts:1254-1277) and its binary plan encoding.

The descriptor keeps the reference's field names. Literal typing follows
JavaScript: a Python ``float`` (or an ``int`` within 2^53) is a JS *number*;
``BigInt(v)`` (or an ``int`` beyond 2^53) is a JS *bigint*. The distinction
matters because the reference coerces number and bigint literals differently
(src/operators.ts:487-547, src/decode.ts:476-556).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Any

OPS = {"eq": 0, "neq": 1, "lt": 2, "lte": 3, "gt": 4, "gte": 5, "between": 6, "not_between": 7,
       "in": 8, "not_in": 9, "is_null": 10, "is_not_null": 11, "like": 12, "not_like": 13}
AGG_FNS = {"sum": 0, "avg": 1, "min": 2, "max": 3, "count": 4, "count_distinct": 5, "stddev": 6,
           "variance": 7, "median": 8, "percentile": 9}
MAX_SAFE = 2 ** 53 - 1


class BigInt(int):
    """A JS bigint literal (exact int64 comparison semantics)."""


def _lit(v: Any) -> bytes:
    if isinstance(v, BigInt) or (isinstance(v, int) and not isinstance(v, bool) and abs(v) > MAX_SAFE):
        iv = int(v)
        if not -(2 ** 63) <= iv < 2 ** 63:
            raise ValueError(f"bigint literal {iv} outside int64")
        return bytes([2]) + struct.pack("<q", iv)
    if isinstance(v, (int, float)) and not isinstance(v, bool):
        return bytes([1]) + struct.pack("<d", float(v))
    raise ValueError(f"unsupported literal {v!r} (numbers and bigints only)")


def _str(s: str) -> bytes:
    b = s.encode("utf-8")
    if len(b) > 0xFFFF:
        raise ValueError("name too long")
    return struct.pack("<H", len(b)) + b


@dataclass
class FilterOp:
    column: str
    op: str
    value: Any = None

    def encode(self) -> bytes:
        if self.op not in OPS:
            raise ValueError(f"unknown filter op {self.op}")
        out = _str(self.column) + bytes([OPS[self.op]])
        v = self.value
        if self.op in ("is_null", "is_not_null"):
            return out + bytes([0])
        if self.op in ("between", "not_between"):
            if not isinstance(v, (list, tuple)) or len(v) != 2:
                raise ValueError("between needs [low, high]")
            return out + bytes([4]) + _lit(v[0]) + _lit(v[1])
        if self.op in ("in", "not_in"):
            vals = list(v)
            return out + bytes([3]) + struct.pack("<I", len(vals)) + b"".join(_lit(x) for x in vals)
        if self.op in ("like", "not_like"):
            raise ValueError("LIKE needs string columns, which are not on the device path")
        lit = _lit(v)
        return out + lit  # kind byte doubles as vkind (1 = f64, 2 = i64)


@dataclass
class AggregateOp:
    fn: str
    column: str
    alias: str | None = None

    def out_name(self) -> str:
        return self.alias if self.alias is not None else f"{self.fn}_{self.column}"


@dataclass
class JoinDescriptor:
    right: "QueryDescriptor"
    leftKey: str
    rightKey: str
    type: str = "inner"


@dataclass
class QueryDescriptor:
    table: str
    filters: list[FilterOp] = field(default_factory=list)
    filterGroups: list[list[FilterOp]] = field(default_factory=list)
    projections: list[str] = field(default_factory=list)
    sortColumn: str | None = None
    sortDirection: str | None = None
    limit: int | None = None
    offset: int | None = None
    aggregates: list[AggregateOp] = field(default_factory=list)
    groupBy: list[str] = field(default_factory=list)
    join: JoinDescriptor | None = None
    # product expressions only: alias -> (left column, right column)
    computedColumns: list[tuple[str, str, str]] = field(default_factory=list)
    # SQL post-stages (src/sql/executor.ts:30-170, src/operators.ts:1177-1227):
    # DISTINCT columns (None: no DISTINCT; []: every output column), a
    # multi-column ORDER BY [(column, "asc" | "desc")] in place of sortColumn,
    # and a HAVING tree over the aggregate output (see HAVING_NODES)
    distinct: list[str] | None = None
    orderBy: list[tuple[str, str]] = field(default_factory=list)
    having: tuple | None = None


# HAVING tree nodes (three-valued, src/sql/evaluator.ts:7-155):
#   ("cmp", op, column, literal)        op: eq ne lt le gt ge
#   ("cmp_col", op, column, column)
#   ("between", column, low, high, negated)
#   ("in", column, [literal or None, ...], negated)
#   ("isnull", column, negated)         negated: IS NOT NULL
#   ("and", a, b)  ("or", a, b)  ("not", a)
HAVING_NODES = ("cmp", "cmp_col", "between", "in", "isnull", "and", "or", "not")
_HCMP = {"eq": 0, "ne": 1, "lt": 2, "le": 3, "gt": 4, "ge": 5}
_HK = {"cmp": 1, "cmp_col": 2, "between": 3, "in": 4, "isnull": 5, "and": 6, "or": 7, "not": 8}


def _having_ops(e, out: list) -> None:
    """Postfix program of a HAVING tree (qm_plan.cpp reads it back)."""
    k = e[0]
    if k not in _HK:
        raise ValueError(f"unknown HAVING node {k!r}")
    if k in ("and", "or"):
        _having_ops(e[1], out)
        _having_ops(e[2], out)
        out.append(bytes([_HK[k]]))
    elif k == "not":
        _having_ops(e[1], out)
        out.append(bytes([_HK[k]]))
    elif k == "cmp":
        out.append(bytes([_HK[k]]) + _str(e[2]) + bytes([_HCMP[e[1]]]) + _lit(e[3]))
    elif k == "cmp_col":
        out.append(bytes([_HK[k]]) + _str(e[2]) + bytes([_HCMP[e[1]]]) + _str(e[3]))
    elif k == "between":
        out.append(bytes([_HK[k]]) + _str(e[1]) + bytes([1 if e[4] else 0]) + _lit(e[2]) + _lit(e[3]))
    elif k == "in":
        vals = list(e[2])
        body = b"".join(bytes([0]) if v is None else _lit(v) for v in vals)
        out.append(bytes([_HK[k]]) + _str(e[1]) + bytes([1 if e[3] else 0]) + struct.pack("<H", len(vals)) + body)
    else:  # isnull
        out.append(bytes([_HK[k]]) + _str(e[1]) + bytes([1 if e[2] else 0]))


JOIN_TYPES = {"inner": 0, "left": 1, "right": 2, "full": 3, "cross": 4}


def _filters(fs: list[FilterOp]) -> bytes:
    return struct.pack("<H", len(fs)) + b"".join(f.encode() for f in fs)


def _groups(gs: list[list[FilterOp]]) -> bytes:
    return struct.pack("<H", len(gs)) + b"".join(_filters(g) for g in gs)


def encode_plan(d: QueryDescriptor) -> bytes:
    """QueryDescriptor -> plan bytes read by qm_plan_compile."""
    v2 = d.distinct is not None or bool(d.orderBy) or d.having is not None
    out = bytearray(struct.pack("<IH", 0x4C504D51, 2 if v2 else 1))
    out += _filters(d.filters)
    out += _groups(d.filterGroups)
    out += struct.pack("<H", len(d.computedColumns))
    for alias, a, b in d.computedColumns:
        out += _str(alias) + bytes([1]) + _str(a) + _str(b)
    out += struct.pack("<H", len(d.groupBy)) + b"".join(_str(g) for g in d.groupBy)
    out += struct.pack("<H", len(d.aggregates))
    for a in d.aggregates:
        if a.fn not in AGG_FNS:
            raise ValueError(f"unknown aggregate {a.fn}")
        out += bytes([AGG_FNS[a.fn]]) + _str(a.column) + _str(a.out_name())
    if d.join is not None:
        j = d.join
        out += bytes([1]) + _str(j.leftKey) + _str(j.rightKey) + bytes([JOIN_TYPES[j.type]])
        out += _filters(j.right.filters) + _groups(j.right.filterGroups)
    else:
        out += bytes([0])
    if d.sortColumn:
        out += bytes([1]) + _str(d.sortColumn) + bytes([1 if d.sortDirection == "desc" else 0])
    else:
        out += bytes([0])
    out += struct.pack("<qq", -1 if d.limit is None else int(d.limit), int(d.offset or 0))
    out += struct.pack("<H", len(d.projections)) + b"".join(_str(p) for p in d.projections)
    if v2:
        if d.distinct is not None:
            out += bytes([1]) + struct.pack("<H", len(d.distinct)) + b"".join(_str(c) for c in d.distinct)
        else:
            out += bytes([0])
        out += struct.pack("<H", len(d.orderBy))
        for col, direction in d.orderBy:
            out += _str(col) + bytes([1 if direction == "desc" else 0])
        ops: list[bytes] = []
        if d.having is not None:
            _having_ops(d.having, ops)
        out += struct.pack("<H", len(ops)) + b"".join(ops)
    return bytes(out)
