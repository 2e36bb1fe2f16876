"""ctypes binding of the C-ABI in include/qmgpu.h (libqmgpu.so).

This is the Python counterpart of the Node-API addon a TypeScript host would
use: plain pointers and sizes across the boundary, status codes turned into
QueryModeError. There is no fallback: if the library is missing, loading
fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .errors import STATUS_CODES, QueryModeError

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "libqmgpu.so"

# dtype tags (QMCB + narrow ints)
QM_F64, QM_I64, QM_UTF8, QM_BOOL, QM_F32VEC, QM_NULL, QM_I32, QM_F32 = 0, 1, 2, 3, 4, 5, 6, 7
QM_I8, QM_I16 = 16, 17

STRATEGIES = {0: "none", 1: "direct_lds", 2: "hash_lds", 3: "direct_global", 4: "hash_global", 5: "part_hash",
              6: "part_direct", 7: "shared_interval", 8: "sorted_runs", 9: "part_runs"}


class ColumnDesc(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dtype", C.c_int32), ("data", C.c_void_p), ("null_bitmap", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [
        ("rows_scanned", C.c_uint64), ("rows_selected", C.c_uint64), ("groups", C.c_uint64),
        ("bytes_algorithmic", C.c_uint64), ("device_ms", C.c_double), ("scan_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("kernel_launches", C.c_uint32), ("retries", C.c_uint32), ("strategy", C.c_uint32), ("pages_skipped", C.c_uint32),
        ("pass_ms", C.c_double * 4), ("pass_records", C.c_uint64), ("join_parts", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("_")}
        d["strategy"] = STRATEGIES.get(self.strategy, str(self.strategy))
        d["pass_ms"] = list(self.pass_ms)
        return d


# (name, restype, argtypes) of every symbol include/qmgpu.h declares
SIGNATURES = [
    ("qm_init", C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("qm_shutdown", C.c_int, [C.c_void_p]),
    ("qm_last_error", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("qm_device_info", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("qm_table_create", C.c_int, [C.c_void_p, C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(ColumnDesc),
                                  C.POINTER(C.c_void_p)]),
    ("qm_table_drop", C.c_int, [C.c_void_p]),
    ("qm_table_rows", C.c_uint64, [C.c_void_p]),
    ("qm_plan_compile", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("qm_plan_close", C.c_int, [C.c_void_p]),
    ("qm_exec", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("qm_measure_read_bw", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    ("qm_exec_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_uint64,
                                  C.POINTER(C.c_uint64)]),
    ("qm_batch_exec", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("qm_batch_submit", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.c_void_p,
                                  C.POINTER(C.c_void_p)]),
    ("qm_batch_poll", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("qm_batch_wait", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("qm_batch_close", C.c_int, [C.c_void_p]),
    ("qm_batch_prepare", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.c_void_p,
                                   C.POINTER(C.c_void_p)]),
    ("qm_prepared_run", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("qm_prepared_info", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    ("qm_prepared_close", C.c_int, [C.c_void_p]),
    ("qm_plan_device_bytes", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    ("qm_tune_set", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("qm_tune_get", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("qm_result_rows", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("qm_result_qmcb_size", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("qm_result_qmcb", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("qm_result_qmcb_pinned", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_uint64)]),
    ("qm_qmcb_release", C.c_int, [C.c_void_p]),
    ("qm_result_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("qm_result_close", C.c_int, [C.c_void_p]),
]

_lib = None


def load() -> C.CDLL:
    """Load libqmgpu.so (built in-tree by ``querymode_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise QueryModeError("QUERY_FAILED", f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    lib = load()
    buf = C.create_string_buffer(4096)
    lib.qm_last_error(None, buf, len(buf))
    return buf.value.decode("utf-8", "replace")


def check(rc: int) -> None:
    if rc != 0:
        raise QueryModeError(STATUS_CODES.get(rc, "QUERY_FAILED"), last_error())
