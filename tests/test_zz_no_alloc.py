"""smdRadixSortNoAlloc (the allocation-free variant of radixSort) through the
C-ABI: its output equals radixSort's on the same device arrays, bit for bit,
on two calls in a row (the second smaller, so stale counters of the first
would show), and a key count above its fixed maximum is refused before
anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p
MAX_KEYS = 1 << 24


def _lib(smd):
    L = smd.lib()
    L.allocateArray.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    L.freeArray.argtypes = [vp]
    L.copyArrayToDevice.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    L.copyArrayFromDevice.argtypes = [vp, vp, ctypes.c_uint, ctypes.c_int]
    L.threadSync.argtypes = []
    L.radixSort.argtypes = [vp] * 8 + [ctypes.c_uint, ctypes.c_uint, ctypes.c_bool]
    L.smdRadixSortNoAlloc.argtypes = [vp] * 4 + [ctypes.c_uint, ctypes.c_uint]
    return L


def _dev(L, nbytes):
    p = vp()
    assert L.allocateArray(ctypes.byref(p), nbytes) == 0
    return p


def _sorted_by(L, fn, keys, vals, bits):
    n = len(keys)
    dk, dv, tk, tv = (_dev(L, 4 * n) for _ in range(4))
    try:
        assert L.copyArrayToDevice(dk, keys.ctypes.data_as(vp), 0, 4 * n) == 0
        assert L.copyArrayToDevice(dv, vals.ctypes.data_as(vp), 0, 4 * n) == 0
        if fn == "radixSort":
            rc = L.radixSort(dk, dv, tk, tv, None, None, None, None, n, bits, False)
        else:
            rc = L.smdRadixSortNoAlloc(dk, dv, tk, tv, n, bits)
        assert rc == 0, smd_error(L)
        assert L.threadSync() == 0
        k = np.zeros(n, np.uint32)
        v = np.zeros(n, np.uint32)
        assert L.copyArrayFromDevice(k.ctypes.data_as(vp), dk, 0, 4 * n) == 0
        assert L.copyArrayFromDevice(v.ctypes.data_as(vp), dv, 0, 4 * n) == 0
        return k, v
    finally:
        for p in (dk, dv, tk, tv):
            L.freeArray(p)


def smd_error(L):
    L.smd_last_error.restype = ctypes.c_char_p
    return L.smd_last_error().decode(errors="replace")


@pytest.mark.parametrize("sizes", [((3_000_001, 22), (100_003, 17)), ((9_000_000, 24), (4097, 12))])
def test_no_alloc_sort_equals_radix_sort(smd, gpu, sizes):
    L = _lib(smd)
    for n, bits in sizes:
        print(f"\n[gpu-test] no-alloc sort n={n} bits={bits}", flush=True)
        rng = np.random.default_rng(n)
        keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 4] = keys[0]  # a long run of equal keys
        vals = rng.permutation(n).astype(np.uint32)
        want = _sorted_by(L, "radixSort", keys, vals, bits)
        got = _sorted_by(L, "noalloc", keys, vals, bits)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        # and the order is the stable one
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(got[0], keys[order]) and np.array_equal(got[1], vals[order])


def test_no_alloc_sort_rejects_above_maximum(smd, gpu):
    L = _lib(smd)
    d = _dev(L, 64)
    try:
        # refused on the host before any launch: the arrays are never touched
        assert L.smdRadixSortNoAlloc(d, d, d, d, MAX_KEYS + 1, 24) == 5  # SMD_ERR_CAPACITY
        assert "exceed the maximum" in smd_error(L)
        assert L.smdRadixSortNoAlloc(d, d, d, d, 16, 0) == 1  # SMD_ERR_INVALID
    finally:
        L.freeArray(d)
