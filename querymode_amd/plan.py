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


JOIN_TYPES = {"inner": 0, "left": 1, "right": 2, "full": 3, "cross": 4}


def _filters(fs: list[FilterOp]) -> bytes:
    return struct.pack("<H", len(fs)) + b"".join(f.encode() for f in fs)


def _groups(gs: list[list[FilterOp]]) -> bytes:
    return struct.pack("<H", len(gs)) + b"".join(_filters(g) for g in gs)


def encode_plan(d: QueryDescriptor) -> bytes:
    """QueryDescriptor -> plan bytes read by qm_plan_compile."""
    out = bytearray(struct.pack("<IH", 0x4C504D51, 1))
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
    return bytes(out)
