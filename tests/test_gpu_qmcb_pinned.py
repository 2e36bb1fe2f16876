"""QMCB encoded on the device into pinned host memory (qm_result_qmcb_pinned):
the whole result is byte-identical to the host encoder (qm_result_qmcb);
every row range decodes to the same rows as that slice of the whole result
(ranges starting off a byte boundary shift the NULL bitmaps and BOOL bits);
the cursor and the pull operator hand out the same rows as collect()."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from querymode_amd import _native as N
from querymode_amd import qmcb
from querymode_amd.client import pinned_qmcb
from querymode_amd.operators import drain_pipeline
from querymode_amd.plan import AggregateOp, FilterOp, QueryDescriptor

pytestmark = pytest.mark.gpu


def _host_qmcb(lib, rh) -> bytes:
    size = C.c_uint64()
    N.check(lib.qm_result_qmcb_size(rh, C.byref(size)))
    buf = np.empty(max(1, size.value), dtype=np.uint8)
    wr = C.c_uint64()
    N.check(lib.qm_result_qmcb(rh, buf.ctypes.data_as(C.c_void_p), size.value, C.byref(wr)))
    return bytes(buf[: wr.value])


def _table(n=50_003, seed=31):
    rng = np.random.default_rng(seed)
    cols = {"k": rng.integers(0, 500, n).astype(np.int64), "f": rng.normal(0, 1, n),
            "g": rng.normal(0, 1, n).astype(np.float32), "b": rng.random(n) < 0.5,
            "s": rng.integers(-100, 100, n).astype(np.int16), "t": rng.integers(-5, 5, n).astype(np.int8),
            "i": rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int32)}
    nulls = {"f": rng.random(n) < 0.1, "b": rng.random(n) < 0.2, "i": rng.random(n) < 0.05}
    return cols, nulls


QUERIES = [
    QueryDescriptor("pin_t", filters=[FilterOp("k", "lt", 300)], projections=["k", "f", "g", "b", "s", "t", "i"]),
    QueryDescriptor("pin_t", groupBy=["k"], aggregates=[AggregateOp("sum", "f"), AggregateOp("count", "*"),
                                                        AggregateOp("min", "i"), AggregateOp("max", "g")]),
    QueryDescriptor("pin_t", sortColumn="f", sortDirection="desc", limit=777, projections=["f", "b", "i"]),
    QueryDescriptor("pin_t", filters=[FilterOp("k", "lt", 0)], projections=["k", "b"]),  # empty result
]


def _rows(buf) -> list[dict]:
    return qmcb.to_rows(qmcb.decode(buf))


@pytest.mark.parametrize("qi", range(len(QUERIES)))
def test_pinned_equals_host_encoder_and_ranges(qm, qi):
    cols, nulls = _table()
    qm.register_table("pin_t", cols, nulls)
    lib = qm.device.lib
    res = qm.executor.run_device(QUERIES[qi])
    print(f"pinned qmcb: query {qi}, {res.rows} rows", flush=True)
    try:
        whole = bytes(res.fetch(0, res.rows))
        assert whole == _host_qmcb(lib, res._h)
        want = _rows(whole)
        n = res.rows
        for row0 in sorted({0, 1, 3, 7, 8, 13, 64, 65, max(0, n - 3), n}):
            for cnt in (1, 5, 8, 63, 1000, n):
                got = _rows(res.fetch(row0, cnt))
                assert got == want[row0:row0 + cnt], (row0, cnt)
    finally:
        res.close()


def test_views_outlive_the_result_and_release(qm):
    cols, nulls = _table(n=10_000, seed=32)
    qm.register_table("pin_t", cols, nulls)
    d = QUERIES[0]
    views = []
    for _ in range(50):  # buffers come back to the library and are reused
        res = qm.executor.run_device(d)
        views.append(res.fetch(0, res.rows))
        res.close()
        if len(views) > 3:
            views.pop(0)
    first = _rows(views[-1])
    assert first == qm.executor.execute(d).rows


@pytest.mark.parametrize("bs", [1, 7, 333, 100_000])
def test_cursor_equals_collect(qm, bs):
    cols, nulls = _table(n=20_000, seed=33)
    qm.register_table("pin_t", cols, nulls)
    df = qm.table("pin_t").filter("k", "lt", 250).select("k", "f", "b", "i")
    want = df.collect().rows
    got = [r for batch in df.cursor(bs) for r in batch]
    assert got == want
    got2 = [r for batch in qm.executor.cursor(df.toDescriptor(), bs) for r in batch]
    assert got2 == want
    assert drain_pipeline(qm.executor.to_operator(df.toDescriptor())) == want


def test_pinned_release_rejects_unknown(qm):
    cols, nulls = _table(n=1000, seed=34)
    qm.register_table("pin_t", cols, nulls)
    lib = qm.device.lib
    buf = (C.c_uint8 * 16)()
    assert lib.qm_qmcb_release(C.cast(buf, C.c_void_p)) != 0
    res = qm.executor.run_device(QueryDescriptor("pin_t", projections=["k"], limit=3))
    try:
        mv = pinned_qmcb(lib, res._h)
        assert qmcb.decode(mv).row_count == 3
    finally:
        res.close()
