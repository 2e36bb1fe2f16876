"""QMCB wire format: byte-level known answers from the format's spec
(docs/src/content/docs/columnar-format.mdx:20-57) and the round trips of
src/columnar.test.ts:291-349, 586-673."""
from __future__ import annotations

import struct

import numpy as np

from querymode_amd import qmcb as Q


def test_header_bytes():
    # u32 magic 0x42434D51 ("QMCB" little-endian), u32 rows, u16 cols;
    # per column u16 nameLen, name, u8 dtype, u8 hasNulls; then data
    b = Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("x", Q.DTYPE_I32, np.array([7, -1], np.int32), 2)], 2))
    assert b[:4] == b"QMCB"
    assert b == (struct.pack("<IIH", 0x42434D51, 2, 1) + struct.pack("<H", 1) + b"x" + bytes([6, 0])
                 + struct.pack("<ii", 7, -1))


def test_null_bitmap_bytes_bit_set_is_null():
    vals = [1.0, None, 3.0]
    data = np.array([0 if v is None else v for v in vals], np.float64)
    nb = Q.pack_bits(np.array([v is None for v in vals]))
    assert nb.tolist() == [0b010]
    b = Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("x", Q.DTYPE_F64, data, 3, nb)], 3))
    hdr = 10 + 2 + 1 + 2
    assert b[hdr - 1] == 1          # hasNulls
    assert b[hdr] == 0b010          # bitmap precedes the values
    assert np.frombuffer(b[hdr + 1:], np.float64).tolist() == [1.0, 0.0, 3.0]
    assert [r["x"] for r in Q.to_rows(Q.decode(b))] == [1.0, None, 3.0]


def test_round_trip_f64():
    b = Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("val", Q.DTYPE_F64, np.array([1.1, 2.2, 3.3]), 3)], 3))
    d = Q.decode(b)
    assert d.row_count == 3 and d.columns[0].data.tolist() == [1.1, 2.2, 3.3]


def test_round_trip_strings_offsets():
    b = Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("name", Q.DTYPE_UTF8, ["alice", "bob", "charlie"], 3)], 3))
    body = b[10 + 2 + 4 + 2:]
    total, = struct.unpack_from("<I", body, 0)
    offs = np.frombuffer(body[4:4 + 16], np.uint32).tolist()
    assert total == 15 and offs == [0, 5, 8, 15] and body[20:] == b"alicebobcharlie"
    assert [r["name"] for r in Q.to_rows(Q.decode(b))] == ["alice", "bob", "charlie"]


def test_round_trip_bool_packing():
    b = Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("flag", Q.DTYPE_BOOL, np.array([True, False, True]), 3)], 3))
    assert b[-1] == 0b00000101
    assert [r["flag"] for r in Q.to_rows(Q.decode(b))] == [True, False, True]


def test_i64_materialises_as_int():
    v = np.array([2 ** 62, -3], np.int64)
    rows = Q.to_rows(Q.decode(Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("k", Q.DTYPE_I64, v, 2)], 2))))
    assert rows == [{"k": 2 ** 62}, {"k": -3}] and isinstance(rows[0]["k"], int)


def test_not_qmcb_returns_none():
    assert Q.decode(b"") is None
    assert Q.decode(b"\x00" * 9) is None
    assert Q.decode(struct.pack("<IIH", 0x12345678, 0, 0)) is None


def test_empty_batch_and_large_batch():
    d = Q.decode(Q.encode(Q.ColumnarBatch([Q.ColumnarColumn("x", Q.DTYPE_F64, np.zeros(0), 0)], 0)))
    assert d.row_count == 0 and d.columns[0].data.size == 0
    n = 100_000
    rng = np.random.default_rng(0)
    cols = [Q.ColumnarColumn("a", Q.DTYPE_I32, rng.integers(-9, 9, n).astype(np.int32), n),
            Q.ColumnarColumn("b", Q.DTYPE_F32, rng.random(n).astype(np.float32), n,
                             Q.pack_bits(rng.random(n) < 0.1))]
    d = Q.decode(Q.encode(Q.ColumnarBatch(cols, n)))
    assert np.array_equal(d.columns[0].data, cols[0].data)
    assert np.array_equal(d.columns[1].data, cols[1].data)
    assert np.array_equal(d.columns[1].null_bitmap, cols[1].null_bitmap)
