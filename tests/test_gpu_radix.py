"""The stable LSD radix sort (k_radix.hip) behind ORDER BY without LIMIT,
the multi-column ORDER BY and DISTINCT stages and the key-range sort passes,
called directly: for sizes around the 4,096-key tile, bit ranges that end
inside a digit or start above bit 0, u8 keys, many ties and all-equal keys,
the keys and values it returns equal numpy's stable argsort of the key bits
(`kind="stable"`: equal keys keep their input order). Device memory comes
from torch (plumbing only)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from querymode_amd import _native as N
    L = N.load()
    L.qmk_radix_tmp_bytes.restype = C.c_size_t
    L.qmk_radix_tmp_bytes.argtypes = [C.c_uint64, C.c_int]
    for name in ("qmk_radix_sort_u64", "qmk_radix_sort_u8"):
        f = getattr(L, name)
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_void_p,
                      C.c_size_t, C.c_void_p]
    return L


def _sort(lib, keys: np.ndarray, vals: np.ndarray, begin: int, end: int):
    import torch
    dev = torch.device("cuda:0")
    n = keys.size
    kb = 1 if keys.dtype == np.uint8 else 8
    kt = torch.from_numpy(keys.view(np.uint8 if kb == 1 else np.int64)).to(dev)
    vt = torch.from_numpy(vals.view(np.int64)).to(dev)
    ko, vo = torch.empty_like(kt), torch.empty_like(vt)
    tb = int(lib.qmk_radix_tmp_bytes(n, kb))
    tmp = torch.empty(max(tb, 1), dtype=torch.uint8, device=dev)
    fn = lib.qmk_radix_sort_u8 if kb == 1 else lib.qmk_radix_sort_u64
    torch.cuda.synchronize()
    rc = fn(kt.data_ptr(), ko.data_ptr(), vt.data_ptr(), vo.data_ptr(), n, begin, end, tmp.data_ptr(), tb, None)
    torch.cuda.synchronize()
    assert rc == 0, rc
    ko_np = ko.cpu().numpy().view(np.uint8 if kb == 1 else np.uint64)
    return ko_np, vo.cpu().numpy().view(np.uint64)


def _want(keys, vals, begin, end):
    k = keys.astype(np.uint64)
    bits = (k >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1 if end - begin < 64 else 2 ** 64 - 1)
    order = np.argsort(bits, kind="stable")
    return keys[order], vals[order]


@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 12_289, 1_000_003])
@pytest.mark.parametrize("begin,end", [(0, 64), (0, 13), (8, 24), (0, 1), (40, 64)])
def test_u64_pairs_equal_stable_argsort(lib, n, begin, end):
    rng = np.random.default_rng(n * 131 + begin * 7 + end)
    keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    keys[rng.random(n) < 0.3] = keys[0]                   # a heavy tie across tiles
    keys[: n // 7] &= np.uint64(0xFFFF)                   # small keys: ties in the high digits
    vals = np.arange(n, dtype=np.uint64) * np.uint64(3) + np.uint64(11)
    print(f"radix u64: n {n}, bits [{begin}, {end})", flush=True)
    gk, gv = _sort(lib, keys, vals, begin, end)
    wk, wv = _want(keys, vals, begin, end)
    assert np.array_equal(gk, wk) and np.array_equal(gv, wv)


@pytest.mark.parametrize("n", [5, 4096, 300_001])
@pytest.mark.parametrize("end", [1, 3, 8])
def test_u8_pairs_equal_stable_argsort(lib, n, end):
    rng = np.random.default_rng(n + end)
    keys = rng.integers(0, 256, n, dtype=np.uint8)
    vals = rng.integers(0, 2 ** 62, n, dtype=np.uint64)
    print(f"radix u8: n {n}, bits [0, {end})", flush=True)
    gk, gv = _sort(lib, keys, vals, 0, end)
    wk, wv = _want(keys, vals, 0, end)
    assert np.array_equal(gk, wk) and np.array_equal(gv, wv)


def test_all_equal_keys_keep_input_order(lib):
    n = 70_001
    keys = np.full(n, 0xDEADBEEF, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint64)[::-1].copy()
    gk, gv = _sort(lib, keys, vals, 0, 64)
    assert np.array_equal(gk, keys) and np.array_equal(gv, vals)


def test_rejects_bad_arguments(lib):
    import torch
    dev = torch.device("cuda:0")
    a = torch.zeros(16, dtype=torch.int64, device=dev)
    b = torch.zeros(16, dtype=torch.int64, device=dev)
    tb = int(lib.qmk_radix_tmp_bytes(16, 8))
    tmp = torch.empty(tb, dtype=torch.uint8, device=dev)
    # in place (kin == kout), too little scratch, a bit range past the key width: rejected before any launch
    assert lib.qmk_radix_sort_u64(a.data_ptr(), a.data_ptr(), b.data_ptr(), tmp.data_ptr(), 16, 0, 64,
                                  tmp.data_ptr(), tb, None) != 0
    c = torch.zeros(16, dtype=torch.int64, device=dev)
    d = torch.zeros(16, dtype=torch.int64, device=dev)
    assert lib.qmk_radix_sort_u64(a.data_ptr(), b.data_ptr(), c.data_ptr(), d.data_ptr(), 16, 0, 64,
                                  tmp.data_ptr(), tb - 1, None) != 0
    assert lib.qmk_radix_sort_u64(a.data_ptr(), b.data_ptr(), c.data_ptr(), d.data_ptr(), 16, 0, 65,
                                  tmp.data_ptr(), tb, None) != 0
    torch.cuda.synchronize()
