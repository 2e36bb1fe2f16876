"""QMCB columnar wire format (src/columnar.ts:22-32, 438-613, 1056-1121).

Layout (little-endian): ``u32 magic 0x42434D51, u32 rows, u16 cols``; per
column ``u16 nameLen, name, u8 dtype, u8 hasNulls``; then per column an
optional null bitmap (``ceil(rows/8)`` bytes, bit set = null) followed by the
data: F64/I64 ``8*rows``, I32/F32 ``4*rows``, BOOL bit-packed, UTF8
``u32 total + (rows+1) u32 offsets + bytes``, F32VEC ``u32 dim + 4*rows*dim``.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Any

import numpy as np

QMCB_MAGIC = 0x42434D51
DTYPE_F64, DTYPE_I64, DTYPE_UTF8, DTYPE_BOOL, DTYPE_F32VEC, DTYPE_NULL, DTYPE_I32, DTYPE_F32 = range(8)

_NP = {DTYPE_F64: np.float64, DTYPE_I64: np.int64, DTYPE_I32: np.int32, DTYPE_F32: np.float32}


@dataclass
class ColumnarColumn:
    name: str
    dtype: int
    data: Any                 # numpy array (numeric/bool), list[str] (utf8), 2-D array (f32vec), None (null)
    row_count: int
    null_bitmap: np.ndarray | None = None  # uint8, bit set = null
    vector_dim: int = 0

    def is_null(self, r: int) -> bool:
        return self.null_bitmap is not None and bool(self.null_bitmap[r >> 3] & (1 << (r & 7)))


@dataclass
class ColumnarBatch:
    columns: list[ColumnarColumn] = field(default_factory=list)
    row_count: int = 0

    def column(self, name: str) -> ColumnarColumn | None:
        for c in self.columns:
            if c.name == name:
                return c
        return None


def pack_bits(flags: np.ndarray) -> np.ndarray:
    return np.packbits(np.asarray(flags, dtype=np.uint8), bitorder="little")


def unpack_bits(bits: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(np.asarray(bits, dtype=np.uint8), bitorder="little", count=n).astype(bool)


def decode(buf: bytes | bytearray | memoryview) -> ColumnarBatch | None:
    """decodeColumnarBatch: returns None for a buffer that is not QMCB."""
    mv = memoryview(buf)
    if len(mv) < 10:
        return None
    magic, rows, ncols = struct.unpack_from("<IIH", mv, 0)
    if magic != QMCB_MAGIC:
        return None
    pos = 10
    descs = []
    for _ in range(ncols):
        (nl,) = struct.unpack_from("<H", mv, pos)
        pos += 2
        name = bytes(mv[pos:pos + nl]).decode("utf-8")
        pos += nl
        dt, hn = mv[pos], mv[pos + 1]
        pos += 2
        descs.append((name, dt, hn != 0))
    cols = []
    for name, dt, hn in descs:
        nb = None
        if hn:
            k = (rows + 7) // 8
            nb = np.frombuffer(bytes(mv[pos:pos + k]), dtype=np.uint8).copy()
            pos += k
        if dt in _NP:
            sz = rows * np.dtype(_NP[dt]).itemsize
            data = np.frombuffer(bytes(mv[pos:pos + sz]), dtype=_NP[dt]).copy()
            pos += sz
            cols.append(ColumnarColumn(name, dt, data, rows, nb))
        elif dt == DTYPE_BOOL:
            k = (rows + 7) // 8
            data = unpack_bits(np.frombuffer(bytes(mv[pos:pos + k]), dtype=np.uint8), rows)
            pos += k
            cols.append(ColumnarColumn(name, dt, data, rows, nb))
        elif dt == DTYPE_UTF8:
            (total,) = struct.unpack_from("<I", mv, pos)
            pos += 4
            offs = np.frombuffer(bytes(mv[pos:pos + 4 * (rows + 1)]), dtype=np.uint32)
            pos += 4 * (rows + 1)
            raw = bytes(mv[pos:pos + total])
            pos += total
            strs = [raw[offs[i]:offs[i + 1]].decode("utf-8") for i in range(rows)]
            cols.append(ColumnarColumn(name, dt, strs, rows, nb))
        elif dt == DTYPE_F32VEC:
            (dim,) = struct.unpack_from("<I", mv, pos)
            pos += 4
            sz = 4 * rows * dim
            data = np.frombuffer(bytes(mv[pos:pos + sz]), dtype=np.float32).reshape(rows, dim).copy()
            pos += sz
            cols.append(ColumnarColumn(name, dt, data, rows, nb, dim))
        else:
            cols.append(ColumnarColumn(name, dt, None, rows, nb))
    return ColumnarBatch(cols, rows)


def encode(batch: ColumnarBatch) -> bytes:
    """encodeColumnarBatch."""
    out = bytearray(struct.pack("<IIH", QMCB_MAGIC, batch.row_count, len(batch.columns)))
    n = batch.row_count
    for c in batch.columns:
        nb = c.name.encode("utf-8")
        out += struct.pack("<H", len(nb)) + nb + bytes([c.dtype, 1 if c.null_bitmap is not None else 0])
    for c in batch.columns:
        if c.null_bitmap is not None:
            out += bytes(np.asarray(c.null_bitmap, dtype=np.uint8)[: (n + 7) // 8])
        if c.dtype in _NP:
            out += np.ascontiguousarray(c.data, dtype=_NP[c.dtype])[:n].tobytes()
        elif c.dtype == DTYPE_BOOL:
            out += pack_bits(np.asarray(c.data, dtype=bool)[:n]).tobytes()
        elif c.dtype == DTYPE_UTF8:
            raw = [s.encode("utf-8") for s in c.data[:n]]
            offs = np.zeros(n + 1, dtype=np.uint32)
            if n:
                offs[1:] = np.cumsum([len(b) for b in raw])
            out += struct.pack("<I", int(offs[-1])) + offs.tobytes() + b"".join(raw)
        elif c.dtype == DTYPE_F32VEC:
            out += struct.pack("<I", c.vector_dim) + np.ascontiguousarray(c.data, dtype=np.float32).tobytes()
    return bytes(out)


def to_rows(batch: ColumnarBatch) -> list[dict]:
    """columnarBatchToRows: one dict per row, NULL -> None, I64 -> int."""
    rows = [dict() for _ in range(batch.row_count)]
    for c in batch.columns:
        if c.data is None:
            vals = [None] * batch.row_count
        elif c.dtype in (DTYPE_I64, DTYPE_I32):
            vals = [int(v) for v in c.data]
        elif c.dtype in (DTYPE_F64, DTYPE_F32):
            vals = [float(v) for v in c.data]
        elif c.dtype == DTYPE_BOOL:
            vals = [bool(v) for v in c.data]
        else:
            vals = list(c.data)
        if c.null_bitmap is not None:
            nulls = unpack_bits(c.null_bitmap, batch.row_count)
            vals = [None if nulls[i] else v for i, v in enumerate(vals)]
        for i, v in enumerate(vals):
            rows[i][c.name] = v
    return rows
