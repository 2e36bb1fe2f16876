"""Host-side Parquet decode (querymode_amd.client.decode_parquet) on files
written here with pyarrow: every supported type, NULLs (definition levels),
several row groups and pages, and the widening of unsigned types; a string
column is refused."""
from __future__ import annotations

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from querymode_amd.client import decode_parquet
from querymode_amd.errors import QueryModeError


def _write(tmp_path, n=20_000, seed=3):
    rng = np.random.default_rng(seed)
    data = {
        "f64": rng.normal(0, 1, n), "f32": rng.normal(0, 1, n).astype(np.float32),
        "i64": rng.integers(-2 ** 62, 2 ** 62, n), "i32": rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32),
        "i16": rng.integers(-2 ** 15, 2 ** 15, n).astype(np.int16), "i8": rng.integers(-128, 128, n).astype(np.int8),
        "b": rng.random(n) < 0.5, "u32": rng.integers(0, 2 ** 32, n).astype(np.uint32),
    }
    masks = {k: rng.random(n) < 0.1 for k in ("f64", "i32", "b")}
    arrays = {k: pa.array(v, mask=masks.get(k)) for k, v in data.items()}
    path = tmp_path / "t.parquet"
    pq.write_table(pa.table(arrays), path, row_group_size=7_000, data_page_size=4096)
    return path, data, masks


def test_decode_types_nulls_row_groups(tmp_path):
    path, data, masks = _write(tmp_path)
    assert pq.ParquetFile(path).num_row_groups == 3
    cols, nulls = decode_parquet(str(path))
    assert set(cols) == set(data)
    for k, v in data.items():
        got = cols[k]
        if k == "u32":
            assert got.dtype == np.int64 and np.array_equal(got, v.astype(np.int64))
            continue
        assert got.dtype == v.dtype, k
        if k in masks:
            assert np.array_equal(nulls[k], masks[k])
            assert np.array_equal(got[~masks[k]], v[~masks[k]])
            assert not got[masks[k]].any()  # NULL slots hold 0 / False
        else:
            assert k not in nulls and np.array_equal(got, v)


def test_decode_selected_columns(tmp_path):
    path, data, _ = _write(tmp_path, n=1000)
    cols, _ = decode_parquet(str(path), ["i64", "b"])
    assert list(cols) == ["i64", "b"]


def test_strings_are_unsupported(tmp_path):
    path = tmp_path / "s.parquet"
    pq.write_table(pa.table({"s": pa.array(["a", "b"])}), path)
    with pytest.raises(QueryModeError) as e:
        decode_parquet(str(path))
    assert e.value.code == "UNSUPPORTED"
