"""Deterministic synthetic tables for the five benchmark workloads.

Values are pure functions of (seed, stream, row) computed by libqmdatagen.so,
so a row range can be regenerated anywhere (the oracle and the device see the
same bytes). Stream ids are ``16 * workload + column``.

* filter_groupby_small (1M rows, seed 0): status i32 U{0,1,2}, category i32
  U[0,16), amount f64 U[0,1000).
* filter_groupby_zipf (100M rows): key i64 Zipf(s=1.1) over 1M distinct ids
  (ranks mapped to sparse int64 ids by a bijective mix), a i32 U[0,1e6),
  b f64 N(100, 25), flag i8 Bernoulli(0.5).
* join_build (10M rows): id i64 = random permutation of [0, 10M), region i32
  U[0,64), weight f32 U[0,1). join_probe (200M rows): fk i64, 90% uniform over
  the id set, 10% uniform over [10M, 20M); amount f64 lognormal(3, 1).
* batch_table (500M rows): k1 i32 U[0,4), k2 i64 U[0,1e7), v1 i32 uniform
  over the full int32 range, v2 f64 U[0,1). With ``k2_order=True`` the table
  is stored in k2 order (its sort key): k2 holds the sorted sample of 500M
  uniform draws over [0, 1e7) (a function of (seed, stream, n, row)), the
  other columns are unchanged.
* batch_queries: the 256-query mix over batch_table.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libqmdatagen.so"
_lib = None

u64, u32, i64, dbl = C.c_uint64, C.c_uint32, C.c_int64, C.c_double


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing (run __graft_entry__.build())")
        lib = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        lib.qg_fill_i64.argtypes = [u64, u32, u64, u64, i64, u64, vp]
        lib.qg_fill_i32.argtypes = [u64, u32, u64, u64, i64, u64, vp]
        lib.qg_fill_i8.argtypes = [u64, u32, u64, u64, i64, u64, vp]
        lib.qg_fill_uniform_f64.argtypes = [u64, u32, u64, u64, dbl, dbl, vp]
        lib.qg_fill_uniform_f32.argtypes = [u64, u32, u64, u64, C.c_float, C.c_float, vp]
        lib.qg_fill_normal_f64.argtypes = [u64, u32, u64, u64, dbl, dbl, vp]
        lib.qg_fill_lognormal_f64.argtypes = [u64, u32, u64, u64, dbl, dbl, vp]
        lib.qg_fill_bernoulli_i8.argtypes = [u64, u32, u64, u64, dbl, vp]
        lib.qg_fill_zipf_i64.argtypes = [u64, u32, u64, u64, dbl, u64, C.c_int, vp]
        lib.qg_permutation_i64.argtypes = [u64, u32, u64, i64, vp]
        lib.qg_fill_fk_i64.argtypes = [u64, u32, u64, u64, u64, dbl, vp]
        lib.qg_sorted_uniform_i64.argtypes = [u64, u32, u64, i64, u64, vp]
        for f in ("qg_fill_i64", "qg_fill_i32", "qg_fill_i8", "qg_fill_uniform_f64", "qg_fill_uniform_f32",
                  "qg_fill_normal_f64", "qg_fill_lognormal_f64", "qg_fill_bernoulli_i8", "qg_fill_zipf_i64",
                  "qg_permutation_i64", "qg_fill_fk_i64", "qg_sorted_uniform_i64"):
            getattr(lib, f).restype = None
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def uniform_int(seed, stream, row0, n, lo, span, dtype) -> np.ndarray:
    lib = _load()
    out = np.empty(n, dtype=dtype)
    fn = {np.int64: lib.qg_fill_i64, np.int32: lib.qg_fill_i32, np.int8: lib.qg_fill_i8}[dtype]
    fn(seed, stream, row0, n, lo, span, _p(out))
    return out


def uniform_f64(seed, stream, row0, n, lo, hi) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _load().qg_fill_uniform_f64(seed, stream, row0, n, lo, hi, _p(out))
    return out


def uniform_f32(seed, stream, row0, n, lo, hi) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    _load().qg_fill_uniform_f32(seed, stream, row0, n, lo, hi, _p(out))
    return out


def normal_f64(seed, stream, row0, n, mu, sigma) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _load().qg_fill_normal_f64(seed, stream, row0, n, mu, sigma, _p(out))
    return out


def lognormal_f64(seed, stream, row0, n, mu, sigma) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _load().qg_fill_lognormal_f64(seed, stream, row0, n, mu, sigma, _p(out))
    return out


def bernoulli_i8(seed, stream, row0, n, p) -> np.ndarray:
    out = np.empty(n, dtype=np.int8)
    _load().qg_fill_bernoulli_i8(seed, stream, row0, n, p, _p(out))
    return out


def zipf_i64(seed, stream, row0, n, s, nvals, mapped=True) -> np.ndarray:
    out = np.empty(n, dtype=np.int64)
    _load().qg_fill_zipf_i64(seed, stream, row0, n, s, nvals, 1 if mapped else 0, _p(out))
    return out


def permutation_i64(seed, stream, n, base=0) -> np.ndarray:
    out = np.empty(n, dtype=np.int64)
    _load().qg_permutation_i64(seed, stream, n, base, _p(out))
    return out


def fk_i64(seed, stream, row0, n, nids, match_rate) -> np.ndarray:
    out = np.empty(n, dtype=np.int64)
    _load().qg_fill_fk_i64(seed, stream, row0, n, nids, match_rate, _p(out))
    return out


# ---- the workloads ----
def filter_groupby_small(n=1_000_000, seed=0, row0=0) -> dict:
    return {
        "status": uniform_int(seed, 17, row0, n, 0, 3, np.int32),
        "category": uniform_int(seed, 18, row0, n, 0, 16, np.int32),
        "amount": uniform_f64(seed, 19, row0, n, 0.0, 1000.0),
    }


def filter_groupby_zipf(n=100_000_000, seed=0, row0=0, nkeys=1_000_000) -> dict:
    return {
        "key": zipf_i64(seed, 33, row0, n, 1.1, nkeys, True),
        "a": uniform_int(seed, 34, row0, n, 0, 1_000_000, np.int32),
        "b": normal_f64(seed, 35, row0, n, 100.0, 25.0),
        "flag": bernoulli_i8(seed, 36, row0, n, 0.5),
    }


def join_build(n=10_000_000, seed=0) -> dict:
    return {
        "id": permutation_i64(seed, 49, n, 0),
        "region": uniform_int(seed, 50, 0, n, 0, 64, np.int32),
        "weight": uniform_f32(seed, 51, 0, n, 0.0, 1.0),
    }


def join_probe(n=200_000_000, nids=10_000_000, seed=0, row0=0) -> dict:
    return {
        "fk": fk_i64(seed, 52, row0, n, nids, 0.9),
        "amount": lognormal_f64(seed, 53, row0, n, 3.0, 1.0),
    }


def sorted_uniform_i64(seed: int, stream: int, n: int, lo: int, hi: int) -> np.ndarray:
    """Ascending sample of n uniform integers over [lo, hi)."""
    out = np.empty(n, dtype=np.int64)
    _load().qg_sorted_uniform_i64(seed, stream, n, lo, hi - lo, _p(out))
    return out


def batch_table(n=500_000_000, seed=0, row0=0, k2_order=False) -> dict:
    if k2_order and row0:
        raise ValueError("a k2-ordered table is generated whole")
    return {
        "k1": uniform_int(seed, 65, row0, n, 0, 4, np.int32),
        "k2": (sorted_uniform_i64(seed, 70, n, 0, 10_000_000) if k2_order
               else uniform_int(seed, 66, row0, n, 0, 10_000_000, np.int64)),
        "v1": uniform_int(seed, 67, row0, n, -(2 ** 31), 2 ** 32, np.int32),
        "v2": uniform_f64(seed, 68, row0, n, 0.0, 1.0),
    }


def batch_queries(nq: int = 256, seed: int = 0, min_sel: float = 1e-4, max_sel: float = 0.5):
    """The batched mix: query q filters v2 to [lo, lo + s) with s log-uniform in
    [min_sel, max_sel] and lo uniform in [0, 1 - s), groups by k1 (even q) or
    k2 (odd q), and computes SUM(v1), COUNT(*), MIN(v1), MAX(v2)."""
    from .plan import AggregateOp, FilterOp, QueryDescriptor
    rng = np.random.default_rng(np.random.SeedSequence([seed, 4]))
    out = []
    for q in range(nq):
        s = float(np.exp(rng.uniform(np.log(min_sel), np.log(max_sel))))
        lo = float(rng.uniform(0.0, 1.0 - s))
        key = "k1" if q % 2 == 0 else "k2"
        out.append(QueryDescriptor(
            "batch", filters=[FilterOp("v2", "gte", lo), FilterOp("v2", "lt", lo + s)], groupBy=[key],
            aggregates=[AggregateOp("sum", "v1"), AggregateOp("count", "*"), AggregateOp("min", "v1"),
                        AggregateOp("max", "v2")]))
    return out
