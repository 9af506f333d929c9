"""The device radix sort under every query path (radix.hip), called directly
through its C++ symbols in libgwn.so: keys and values after a sort equal a
stable CPU sort of the same keys (numpy's stable argsort), bit for bit, for
sizes around the 4096-key tile and for 16/30/32-bit and 64-bit keys."""
import ctypes as C

import numpy as np
import pytest

from libigl_amd import _native as nat

pytestmark = pytest.mark.gpu


def _fns():
    L = nat.lib()
    tb = getattr(L, "_ZN3gwn21radix_sort_temp_bytesElb")
    tb.argtypes, tb.restype = [C.c_int64, C.c_bool], C.c_size_t
    s32 = getattr(L, "_ZN3gwn14radix_sort_u32EPjS0_liPvmP12ihipStream_t")
    s64 = getattr(L, "_ZN3gwn14radix_sort_u64EPmPjliPvmP12ihipStream_t")
    for f in (s32, s64):
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p]
        f.restype = None
    return tb, s32, s64


def _sort(keys, bits):
    from hip_mem import DeviceArray
    tb, s32, s64 = _fns()
    n = keys.shape[0]
    key64 = keys.dtype == np.uint64
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = DeviceArray.from_numpy(keys), DeviceArray.from_numpy(vals)
    nbytes = tb(n, key64)
    tmp = DeviceArray((max(nbytes, 1),), np.uint8)
    (s64 if key64 else s32)(dk.data_ptr, dv.data_ptr, n, bits, tmp.data_ptr, nbytes, None)
    return dk.numpy(), dv.numpy()


def _check(keys, bits):
    k, v = _sort(keys, bits)
    mask = (np.uint64(1) << np.uint64(bits)) - np.uint64(1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    sig = (keys.astype(np.uint64) & mask)
    order = np.argsort(sig, kind="stable").astype(np.uint32)
    assert np.array_equal(v, order)
    assert np.array_equal(k, keys[order])


@pytest.mark.parametrize("n", [2, 3, 255, 4095, 4096, 4097, 8192 + 5, 100_003, 1_000_000])
@pytest.mark.parametrize("bits", [16, 30, 32])
def test_radix_sort_u32(n, bits):
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
    _check(keys, bits)


@pytest.mark.parametrize("n", [5000, 300_001])
def test_radix_sort_u32_duplicates_and_skew(n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 7, n).astype(np.uint32) << np.uint32(20)  # few distinct keys, long equal runs
    _check(keys, 30)
    _check(np.zeros(n, np.uint32), 30)  # one digit everywhere
    _check(np.arange(n, dtype=np.uint32)[::-1].copy(), 30)  # reversed


@pytest.mark.parametrize("n,bits", [(4097, 64), (200_003, 63), (100_000, 40)])
def test_radix_sort_u64(n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64)
    if bits == 64:
        keys |= np.uint64(1) << np.uint64(63)
        keys[::3] &= ~(np.uint64(1) << np.uint64(63))
    _check(keys if bits == 64 else keys & np.uint64((1 << bits) - 1), bits)
