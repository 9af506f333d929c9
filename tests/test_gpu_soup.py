"""GPU parity of the triangle-soup fast winding number.

Against the reference (oracle/_ref, the reference's own UT_SolidAngle) and the
bit-exact C restatement:
  * tree: identical node arrays (SAH4 mode);
  * moments: BoxData computed by the GPU upward pass identical bit for bit;
  * W: max |W_gpu - W_ref| <= 1e-4 (the fast path's tolerance; observed ~1e-6,
    from the per-slot summation order and the device atan2f);
  * identical inside/outside classification at 0.5 wherever the reference is
    more than 1e-4 away from 0.5.
Precision modes against the FP64 brute-force oracle: mixed and FP64 within
the Taylor truncation bound of the chosen accuracy_scale.
"""
import numpy as np
import pytest

import oracle_py as O
from libigl_amd import datagen as D
from libigl_amd import gwn

pytestmark = pytest.mark.gpu
TOL_REF = 1e-4


def _ref_soup(V, F, order=2):
    return O.RefSoup(V, F, order) if O.ref_available() else O.Soup(V, F, order)


def _classify_same(W, Wr, band=TOL_REF):
    m = np.abs(Wr - 0.5) > band
    return np.array_equal(W[m] > 0.5, Wr[m] > 0.5)


@pytest.mark.parametrize("level,nb,seed", [(0, 0, 1), (2, 4, 2), (4, 16, 3)])
def test_tree_and_moments_bit_exact(level, nb, seed):
    V, F = D.bumpy_sphere(level, nb, seed) if nb else D.noisy_sphere(level, 0.01, seed)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    ref = _ref_soup(V, F)
    assert np.array_equal(bvh.nodes(), ref.nodes())
    bg, br = bvh.boxdata(), ref.boxdata()
    same = bg.view(np.uint32) == br.view(np.uint32)
    print(f"boxdata bit-equal fraction {same.mean():.6f}; max rel diff "
          f"{(np.abs(bg - br) / np.maximum(np.abs(br), 1e-30))[np.isfinite(br)].max():.3e}")
    assert same.all()


def test_tree_off_origin_large_mesh():
    """20,480 triangles not straddling the origin: the reference's root bound
    for 8,192+ boxes includes the origin; the GPU build's tree, moments and W
    are still the reference's."""
    V, F = D.bumpy_sphere(5, 16, 4)
    V = V + np.array([10.0, -3.0, 2.5])
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    ref = _ref_soup(V, F)
    assert np.array_equal(bvh.nodes(), ref.nodes())
    assert np.array_equal(bvh.boxdata().view(np.uint32), ref.boxdata().view(np.uint32))
    rng = np.random.default_rng(4)
    Q = np.concatenate([rng.uniform(V.min(0) - 0.3, V.max(0) + 0.3, (20000, 3)),
                        D.near_surface_queries(V, F, 20000, 1e-3, rng)])
    W, Wr = bvh.query(Q, 2.0), ref.query(Q, 2.0)
    assert np.abs(W - Wr).max() <= TOL_REF
    assert _classify_same(W, Wr)


def test_soup_81920tri_1Mq_vs_reference():
    V, F, Q = D.soup_workload(nq=1_000_000)  # 81,920 tris, 1M queries (half near the surface)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    W = bvh.query(Q, 2.0)
    ref = _ref_soup(V, F)
    Wr = ref.query(Q, 2.0)
    err = np.abs(W - Wr)
    print(f"81,920 tris, 1M queries, beta=2: max|W-W_ref| {err.max():.3e}, p99 {np.quantile(err, 0.99):.3e}")
    assert err.max() <= TOL_REF
    assert _classify_same(W, Wr)


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("beta", [1.0, 2.0, 4.5])
def test_soup_orders_and_beta_vs_reference(order, beta):
    V, F = D.bumpy_sphere(4, 16, 5)
    rng = np.random.default_rng(order * 10 + int(beta))
    Q = np.concatenate([rng.uniform(-1.5, 1.5, (20000, 3)), D.near_surface_queries(V, F, 20000, 1e-3, rng)])
    W = gwn.FastWindingNumberBVH().build(V, F, order).query(Q, beta)
    Wr = _ref_soup(V, F, order).query(Q, beta)
    assert np.abs(W - Wr).max() <= TOL_REF
    assert _classify_same(W, Wr)


def test_soup_one_shot_and_single_point_overloads():
    V, F, Q = D.soup_workload(nq=2000, level=3, nbumps=8, seed=7)
    W1 = gwn.fast_winding_number(V, F, Q)  # fast_winding_number.h:160-165 (order 2, beta 2)
    bvh = gwn.FastWindingNumberBVH()
    gwn.fast_winding_number(V, F, 2, bvh)
    W2 = gwn.fast_winding_number(bvh, 2.0, Q)
    assert np.array_equal(W1, W2)
    w0 = gwn.fast_winding_number(bvh, 2.0, Q[0])
    assert isinstance(w0, float) and w0 == W2[0]


@pytest.mark.parametrize("order", [0, 1, 2])
def test_ring_walk_and_sequential_walks_agree(order):
    """Batches under 65,536 queries take the walk with one wavefront per
    query (per-lane partial sums, then a fixed-order sum across lanes); larger
    ones the wave-coherent walk (one running sum in rope order).  Same records
    and terms, so the results differ only by FP32 summation rounding (FP64
    rounding for the FP64 form and for the mixed form's queries near 0.5)."""
    V, F = D.bumpy_sphere(4, 16, 20 + order)
    rng = np.random.default_rng(order)
    Q = np.concatenate([rng.uniform(-1.4, 1.4, (40000, 3)), D.near_surface_queries(V, F, 30000, 1e-3, rng)])
    bvh = gwn.FastWindingNumberBVH().build(V, F, order)
    ref = _ref_soup(V, F, order)
    for beta in (2.0, 4.9):
        Wr = ref.query(Q, beta)
        for prec, tol in ((gwn.PREC_FP32, 1e-5), (gwn.PREC_MIXED, 1e-5), (gwn.PREC_FP64, 1e-12)):
            Wbig = bvh.query(Q, beta, prec)  # 70,000 queries: wave-coherent walk
            Wsmall = np.concatenate([bvh.query(Q[k:k + 20000], beta, prec) for k in range(0, Q.shape[0], 20000)])
            d = np.abs(Wbig - Wsmall)
            print(f"order {order} beta {beta} prec {prec}: max |ring - wave| {d.max():.2e}")
            assert d.max() <= tol
            if prec == gwn.PREC_MIXED:
                assert np.array_equal(Wbig > 0.5, Wsmall > 0.5)
            if prec == gwn.PREC_FP32:
                for W in (Wbig, Wsmall):
                    assert np.abs(W - Wr).max() <= TOL_REF
                    assert _classify_same(W, Wr)
            # a batch's result does not depend on how it was split
            assert np.array_equal(bvh.query(Q[:20000], beta, prec), Wsmall[:20000])


def test_wave_walk_block_order_covers_every_query():
    """Batches of 65,536+ queries run their blocks of 256 costliest first (a
    per-XCD counting sort of costs read from a grid built with the tree).
    Every block runs once: an output prefilled with NaN comes back finite
    everywhere, with block counts that do not divide by the 8 XCD ranges,
    queries far outside the grid (clamped to its border cells) and a flat
    mesh whose grid has no extent along z.  Which block runs when never
    changes a result: repeated calls agree bit for bit."""
    from libigl_amd.devmem import DeviceArray, Stream
    rng = np.random.default_rng(91)
    V, F = D.bumpy_sphere(4, 16, 91)
    n = 300_001  # 1,172 blocks: XCD ranges of 146 and 147
    Q = np.concatenate([rng.uniform(-1.4, 1.4, (n // 2, 3)), D.near_surface_queries(V, F, n - n // 2 - 1000, 1e-3, rng),
                        rng.uniform(-1e4, 1e4, (1000, 3))]).astype(np.float32)
    g = np.linspace(-1, 1, 60)
    X, Y = np.meshgrid(g, g, indexing="ij")
    Vf = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    idx = np.arange(60 * 60).reshape(60, 60)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    Ff = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.int32)
    Qf = rng.uniform(-1.2, 1.2, (70_001, 3)).astype(np.float32)
    st = Stream()
    for (Vm, Fm, Qm) in ((V, F, Q), (Vf, Ff, Qf)):
        bvh = gwn.FastWindingNumberBVH().build(Vm, Fm, 2)
        ref = _ref_soup(Vm, Fm)
        dQ = DeviceArray.from_numpy(Qm)
        for prec in (gwn.PREC_FP32, gwn.PREC_MIXED, gwn.PREC_FP64):
            runs = []
            for _ in range(2):
                dW = DeviceArray.from_numpy(np.full(Qm.shape[0], np.nan))
                bvh.query_device(dQ.data_ptr, Qm.shape[0], dW.data_ptr, 4.9, prec, st.ptr)
                st.synchronize()
                runs.append(dW.numpy())
                dW.free()
            assert np.all(np.isfinite(runs[0])), (Fm.shape[0], prec)
            assert np.array_equal(runs[0], runs[1])
            sub = np.arange(0, Qm.shape[0], 97)
            assert np.abs(runs[0][sub] - ref.query(Qm[sub].astype(np.float64), 4.9)).max() <= TOL_REF
        dQ.free()


def test_soup_edge_cases():
    V, F = D.noisy_sphere(2, 0.01, 5)
    # empty mesh: W = 0 (the reference leaves it uninitialised)
    e = gwn.FastWindingNumberBVH().build(np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert np.all(e.query(np.ones((5, 3))) == 0)
    bvh = gwn.FastWindingNumberBVH().build(V, F)
    assert bvh.query(np.zeros((0, 3))).shape == (0,)
    # queries exactly on vertices and at the centre
    Q = np.concatenate([V[:50], np.zeros((1, 3))])
    W = bvh.query(Q)
    Wr = _ref_soup(V, F).query(Q)
    assert np.abs(W - Wr).max() <= TOL_REF and np.all(np.isfinite(W))
    # degenerate and non-finite triangles are handled as the reference does.
    # A NaN first vertex makes the triangle's box NaN and the build drops it
    # (FastWindingNumberForSoups.h:4535-4615); a NaN in the 2nd/3rd vertex is
    # swallowed by h_min/h_max (250-256), the box stays finite, the triangle
    # stays in the tree and its NaN moments reach every query (W = NaN).
    Vn = np.concatenate([V, [[np.nan, 0, 0]]])
    nanv = V.shape[0]
    Qr = np.random.default_rng(1).uniform(-1.4, 1.4, (3000, 3))
    F2 = np.concatenate([F, [[0, 0, 1], [2, 2, 2], [nanv, 0, 1]]]).astype(np.int32)
    b2 = gwn.FastWindingNumberBVH().build(Vn, F2)
    r2 = _ref_soup(Vn, F2)
    assert np.array_equal(b2.nodes(), r2.nodes())
    W2, R2 = b2.query(Qr), r2.query(Qr)
    assert np.all(np.isfinite(R2)) and np.abs(W2 - R2).max() <= TOL_REF
    F3 = np.concatenate([F, [[0, 1, nanv]]]).astype(np.int32)
    b3 = gwn.FastWindingNumberBVH().build(Vn, F3)
    r3 = _ref_soup(Vn, F3)
    assert np.array_equal(b3.nodes(), r3.nodes())
    W3, R3 = b3.query(Qr), r3.query(Qr)
    assert np.array_equal(np.isnan(W3), np.isnan(R3))
    ok = ~np.isnan(R3)
    assert np.abs(W3[ok] - R3[ok]).max(initial=0.0) <= TOL_REF
    # 1..5 triangles: single-node trees
    for k in range(1, 6):
        bk = gwn.FastWindingNumberBVH().build(V, F[:k].copy())
        assert np.abs(bk.query(Qr[:500]) - _ref_soup(V, F[:k].copy()).query(Qr[:500])).max() <= TOL_REF
    # bad order / bad index raise
    with pytest.raises(gwn.GWNError):
        gwn.FastWindingNumberBVH().build(V, F, 3)
    Fb = F.copy()
    Fb[1, 2] = 10 ** 6
    with pytest.raises(gwn.GWNError):
        gwn.FastWindingNumberBVH().build(V, Fb)


def test_precision_modes_vs_fp64_exact():
    V, F = D.bumpy_sphere(4, 16, 9)
    rng = np.random.default_rng(3)
    Q = np.concatenate([rng.uniform(-1.4, 1.4, (3000, 3)), D.near_surface_queries(V, F, 3000, 1e-4, rng)])
    Wx = O.winding_number_bf(V, F, Q)
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    for beta, bound in ((2.0, 1e-2), (5.0, 2e-4)):
        for prec in (gwn.PREC_FP32, gwn.PREC_MIXED, gwn.PREC_FP64):
            W = bvh.query(Q, beta, prec)
            err = np.abs(W - Wx)
            print(f"beta {beta} prec {prec}: max {err.max():.3e} p99 {np.quantile(err, .99):.3e}")
            assert err.max() < bound
    # everything evaluated exactly: FP64 form equals the FP64 oracle closely
    W64 = bvh.query(Q, 1e6, gwn.PREC_FP64)
    assert np.abs(W64 - Wx).max() < 1e-9
    # mixed: FP32 running sum over the 5,120 terms of the far triangles (its
    # rounding, ~2e-6 typical), FP64 near triangles, FP64 re-evaluation near 0.5
    Wm = bvh.query(Q, 1e6, gwn.PREC_MIXED)
    assert np.abs(Wm - Wx).max() < 2e-5
    assert np.array_equal(Wm > 0.5, Wx > 0.5)
