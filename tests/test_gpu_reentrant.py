"""Reentrant const queries (fast_winding_number.h:191-206: the reference's
FastWindingNumberBVH is queried through a const reference from many
igl::parallel_for threads at once, e.g. signed_distance.cpp:90-121).

Several host threads query one handle at once, through the single-point and
batch overloads, and *_device calls on two streams share one handle; every
result must equal the same query run serially, bit for bit (each query's value
depends only on the query, the tree and the precision)."""
import threading

import numpy as np
import pytest

from libigl_amd import datagen as D
from libigl_amd import gwn
from libigl_amd.devmem import DeviceArray, Stream, synchronize

pytestmark = pytest.mark.gpu


def _run_threads(fn, n):
    errs = []

    def wrap(k):
        try:
            fn(k)
        except Exception as e:  # re-raised in the main thread below
            errs.append(e)

    th = [threading.Thread(target=wrap, args=(k,)) for k in range(n)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise errs[0]


def test_soup_threads_single_point_and_batch():
    V, F = D.bumpy_sphere(4, 12, 31)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    rng = np.random.default_rng(31)
    nth, per_pt, per_batch = 8, 150, 20_000
    Qp = np.concatenate([rng.uniform(-1.4, 1.4, (nth * per_pt // 2, 3)),
                         D.near_surface_queries(V, F, nth * per_pt // 2, 1e-3, rng)])
    Qb = rng.uniform(-1.4, 1.4, (nth, per_batch, 3))
    precs = [gwn.PREC_FP32, gwn.PREC_MIXED, gwn.PREC_FP64]
    # serial answers first
    ser_pt = {p: bvh.query(Qp, 4.0, p) for p in precs}
    ser_b = {p: [bvh.query(Qb[k], 4.0, p) for k in range(nth)] for p in precs}
    got_pt = {p: np.full(Qp.shape[0], np.nan) for p in precs}
    got_b = {p: [None] * nth for p in precs}

    def work(k):
        for j, p in enumerate(precs):
            for i in range(k * per_pt, (k + 1) * per_pt):
                # the single-point overload, fast_winding_number.h:203-206
                got_pt[p][i] = gwn.fast_winding_number(bvh, 4.0, Qp[i]) if p == gwn.PREC_FP32 else \
                    bvh.query(Qp[i:i + 1], 4.0, p)[0]
            got_b[p][k] = bvh.query(Qb[k], 4.0, p)

    _run_threads(work, nth)
    for p in precs:
        assert np.array_equal(got_pt[p], ser_pt[p]), p
        for k in range(nth):
            assert np.array_equal(got_b[p][k], ser_b[p][k]), (p, k)


def test_soup_device_calls_on_two_streams():
    V, F = D.bumpy_sphere(4, 12, 32)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    rng = np.random.default_rng(32)
    ncalls, nq = 12, 60_000
    Qs = [np.ascontiguousarray(rng.uniform(-1.4, 1.4, (nq, 3)), dtype=np.float32) for _ in range(ncalls)]
    serial = [bvh.query(Q.astype(np.float64), 3.0, gwn.PREC_MIXED) for Q in Qs]
    dQ = [DeviceArray.from_numpy(Q) for Q in Qs]
    dW = [DeviceArray((nq,), np.float64) for _ in range(ncalls)]
    streams = [Stream(), Stream()]
    for c in range(ncalls):  # back to back, alternating streams, no synchronisation in between
        bvh.query_device(dQ[c].data_ptr, nq, dW[c].data_ptr, 3.0, gwn.PREC_MIXED, streams[c % 2].ptr)
    synchronize()
    for c in range(ncalls):
        assert np.array_equal(dW[c].numpy(), serial[c]), c
    for s in streams:
        s.close()


def test_cloud_and_curve_threads():
    rng = np.random.default_rng(33)
    Vs, Fs = D.noisy_sphere(4, 0.01, 33)
    P, N, _ = D.sample_surface(Vs, Fs, 200_000, rng)
    A = np.full(P.shape[0], 4 * np.pi / P.shape[0])
    cloud = gwn.PointCloudFWN(P.astype(np.float32), N.astype(np.float32), A.astype(np.float32), 2)
    Qc = [rng.uniform(-1.3, 1.3, (50_000, 3)).astype(np.float32) for _ in range(6)]
    ser_c = [cloud.query(Q, 2.0) for Q in Qc]
    Pc, Ec = D.bumpy_curve(20_000, 8, 34)
    curve = gwn.SubtendedAngle(Pc, Ec, 2)
    # batches above the curve's sort threshold use the pooled sort workspace
    Qv = [rng.uniform(-1.4, 1.4, (300_000, 2)).astype(np.float32) for _ in range(6)]
    ser_v = [curve.compute_angle(Q, 2.0) for Q in Qv]
    got_c, got_v = [None] * 6, [None] * 6

    def work(k):
        got_c[k] = cloud.query(Qc[k], 2.0)
        got_v[k] = curve.compute_angle(Qv[k], 2.0)

    _run_threads(work, 6)
    for k in range(6):
        assert np.array_equal(got_c[k], ser_c[k]), k
        assert np.array_equal(got_v[k], ser_v[k]), k


def test_exact_device_on_two_streams():
    """Few queries against many triangles split the triangles into chunks whose
    partial sums live in the call's scratch: two streams must not share it."""
    V, F = D.noisy_sphere(5, 0.01, 35)
    rng = np.random.default_rng(35)
    Os = [rng.uniform(-1.5, 1.5, (300, 3)) for _ in range(8)]
    serial = [gwn.winding_number_bruteforce(V, F, O) for O in Os]
    dV, dF = DeviceArray.from_numpy(np.ascontiguousarray(V)), DeviceArray.from_numpy(F.astype(np.int32))
    dO = [DeviceArray.from_numpy(O) for O in Os]
    dW = [DeviceArray((O.shape[0],), np.float64) for O in Os]
    streams = [Stream(), Stream()]
    from libigl_amd import _native as nat
    for c, O in enumerate(Os):
        nat.check(nat.lib().gwn_exact_winding_number_device(dV.data_ptr, V.shape[0], dF.data_ptr, F.shape[0],
                                                           dO[c].data_ptr, O.shape[0], dW[c].data_ptr,
                                                           streams[c % 2].ptr))
    synchronize()
    for c in range(len(Os)):
        assert np.array_equal(dW[c].numpy(), serial[c]), c
    for s in streams:
        s.close()


def test_exact_tree_device_on_two_streams():
    """One hierarchy queried on two streams at once: each call sorts its
    points in its own scratch."""
    V, F = D.noisy_sphere(5, 0.01, 36)
    rng = np.random.default_rng(36)
    Os = [rng.uniform(-1.5, 1.5, (5000 + 700 * c, 3)) for c in range(6)]
    T = gwn.WindingNumberAABB(V, F)
    serial = [T.winding_number(O) for O in Os]
    dO = [DeviceArray.from_numpy(O) for O in Os]
    dW = [DeviceArray((O.shape[0],), np.float64) for O in Os]
    streams = [Stream(), Stream()]
    from libigl_amd import _native as nat
    for c, O in enumerate(Os):
        nat.check(nat.lib().gwn_exact_tree_query_device(T._h, dO[c].data_ptr, O.shape[0], dW[c].data_ptr,
                                                        streams[c % 2].ptr))
    synchronize()
    for c in range(len(Os)):
        assert np.array_equal(dW[c].numpy(), serial[c]), c
    for s in streams:
        s.close()
