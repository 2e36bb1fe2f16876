"""The C-ABI library (libqmgpu.so) loads on a host without a GPU, exports
every function include/qmgpu.h declares with the binding's signatures, and
fails with a status code — not a crash — when no device is present."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from querymode_amd import _native as N
from querymode_amd.errors import STATUS_CODES, QueryModeError

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "qmgpu.h"


def _declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return re.findall(r"^\s*(?:int|uint64_t|const char \*|void)\s+\**(qm_\w+)\s*\(", text, flags=re.M)


def test_header_declares_the_entry_points():
    names = _declared()
    assert len(names) == len(set(names))
    assert {"qm_init", "qm_shutdown", "qm_table_create", "qm_plan_compile", "qm_exec", "qm_batch_exec",
            "qm_result_qmcb", "qm_result_stats", "qm_result_close", "qm_last_error"} <= set(names)


def test_library_exports_every_declared_symbol():
    import os
    lib = C.CDLL(str(N.LIB_PATH), mode=os.RTLD_NOW)  # every internal symbol must resolve
    for name in _declared():
        assert hasattr(lib, name), f"{name} missing from {N.LIB_PATH.name}"


def test_binding_covers_the_header():
    assert sorted(n for n, _, _ in N.SIGNATURES) == sorted(_declared())


def test_status_codes_match_header():
    text = HEADER.read_text()
    codes = {int(v): k for k, v in re.findall(r"#define\s+QM_ERR_(\w+)\s+(-\d+)", text)}
    assert len(codes) == len(STATUS_CODES) == 11
    for v, k in codes.items():
        assert STATUS_CODES[v].startswith(k.replace("ARG", "ARGUMENT") if k == "INVALID_ARG" else k)


def test_struct_layouts_match_header():
    # qm_stats: 4 x u64, 3 x f64, 4 x u32, 4 x f64, u64, u32 (+ 4 bytes of tail padding);
    # qm_column_desc: ptr, i32, ptr, ptr
    assert C.sizeof(N.Stats) == 4 * 8 + 3 * 8 + 4 * 4 + 4 * 8 + 8 + 8 and N.Stats.pass_ms.offset == 72
    assert N.Stats.join_parts.offset == 112
    assert N.ColumnDesc.data.offset == 16 and C.sizeof(N.ColumnDesc) == 32


def test_init_without_device_reports_error():
    lib = N.load()
    ctx = C.c_void_p()
    rc = lib.qm_init(0, 0, C.byref(ctx))
    if rc == 0:  # a GPU is present: nothing to check here
        lib.qm_shutdown(ctx)
        pytest.skip("device present")
    assert rc in STATUS_CODES
    assert ctx.value is None
    assert N.last_error()
    with pytest.raises(QueryModeError):
        N.check(rc)


def test_null_handles_are_rejected():
    lib = N.load()
    out = C.c_void_p()
    assert lib.qm_exec(None, None, None, C.byref(out)) != 0
    assert lib.qm_result_close(None) == 0 or lib.qm_result_close(None) in STATUS_CODES
    n = C.c_uint64()
    assert lib.qm_result_rows(None, C.byref(n)) != 0
    assert lib.qm_plan_compile(None, b"", 0, C.byref(out)) != 0
    # asynchronous batches: argument errors come back at submission
    plans = (C.c_void_p * 1)()
    assert lib.qm_batch_submit(None, plans, 1, None, C.byref(out)) == -10 and out.value is None
    done = C.c_int()
    assert lib.qm_batch_poll(None, C.byref(done)) == -10
    assert lib.qm_batch_wait(None, plans) == -10
    assert lib.qm_batch_close(None) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", tmp_path / "libqmgpu.so")
    with pytest.raises(QueryModeError):
        N.load()
