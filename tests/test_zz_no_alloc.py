"""The scratch-free query entry point (gwn_soup_query_device_noalloc) against
gwn_soup_query_device, through the C ABI: bitwise the same results in every
precision, on two calls in a row over the same static scratch, the second
with a smaller input; and sizes above the maximum are refused."""
import numpy as np
import pytest

from libigl_amd import _native as nat
from libigl_amd import datagen as D
from libigl_amd import gwn
from libigl_amd.devmem import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def _queries(V, F, n, seed):
    rng = np.random.default_rng(seed)
    h = n // 2
    Q = np.concatenate([rng.uniform(V.min(0) - 0.2, V.max(0) + 0.2, (h, 3)),
                        D.near_surface_queries(V, F, n - h, 1e-4, rng)])
    return np.ascontiguousarray(Q, dtype=np.float32)


@pytest.mark.parametrize("precision", [gwn.PREC_FP32, gwn.PREC_MIXED, gwn.PREC_FP64])
def test_noalloc_equals_device_query(precision):
    L = nat.lib()
    V, F = D.bumpy_sphere(4, 12, 71)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    st = Stream()
    # first a sorted batch, then a smaller one (per-lane walk): the second call
    # finds the scratch holding the first call's contents
    for n, seed in ((300_000, 1), (2_000, 2), (70_000, 3)):
        Q = _queries(V, F, n, seed)
        dQ = DeviceArray.from_numpy(Q)
        dA, dB = DeviceArray((n,), np.float64), DeviceArray((n,), np.float64)
        nat.check(L.gwn_soup_query_device(bvh.handle, dQ.data_ptr, n, 4.0, precision, dA.data_ptr, st.ptr))
        nat.check(L.gwn_soup_query_device_noalloc(bvh.handle, dQ.data_ptr, n, 4.0, precision, dB.data_ptr, st.ptr))
        st.synchronize()
        a, b = dA.numpy(), dB.numpy()
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), (n, precision, np.abs(a - b).max())
    st.close()


def test_noalloc_limits_and_empty():
    L = nat.lib()
    mx = L.gwn_soup_noalloc_max_queries()
    assert mx == 1 << 20
    V, F = D.bumpy_sphere(2, 4, 72)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    dQ = DeviceArray((4, 3), np.float32)
    dW = DeviceArray((4,), np.float64)
    # refused on the host before any device work: more queries than the scratch holds
    rc = L.gwn_soup_query_device_noalloc(bvh.handle, dQ.data_ptr, mx + 1, 2.0, gwn.PREC_FP32, dW.data_ptr, None)
    assert rc == nat.GWN_ERR_INVALID and b"1,048,576" in L.gwn_last_error()
    # nq = 0 is a no-op; an empty mesh gives 0
    assert L.gwn_soup_query_device_noalloc(bvh.handle, dQ.data_ptr, 0, 2.0, gwn.PREC_FP32, dW.data_ptr, None) == 0
    e = gwn.FastWindingNumberBVH().build(np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    dQ2 = DeviceArray.from_numpy(np.ones((4, 3), np.float32))
    dW2 = DeviceArray.from_numpy(np.full(4, 7.0))
    nat.check(L.gwn_soup_query_device_noalloc(e.handle, dQ2.data_ptr, 4, 2.0, gwn.PREC_MIXED, dW2.data_ptr, None))
    assert np.all(dW2.numpy() == 0.0)
    # the largest batch it takes, against the pooled entry point
    Q = _queries(V, F, mx, 5)
    dQ3 = DeviceArray.from_numpy(Q)
    dA, dB = DeviceArray((mx,), np.float64), DeviceArray((mx,), np.float64)
    nat.check(L.gwn_soup_query_device(bvh.handle, dQ3.data_ptr, mx, 2.0, gwn.PREC_MIXED, dA.data_ptr, None))
    nat.check(L.gwn_soup_query_device_noalloc(bvh.handle, dQ3.data_ptr, mx, 2.0, gwn.PREC_MIXED, dB.data_ptr, None))
    assert np.array_equal(dA.numpy(), dB.numpy())
