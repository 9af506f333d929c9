"""GPU igl::knn and igl::copyleft::cgal::point_areas through the C ABI, against
the oracle (best-first octree search / brute force for knn, numpy SVD + Qhull
Delaunay for point areas) and the closed forms of knn_cases.

Tolerances: knn indices are exact (distances are FP64 on both sides with the
same summation order; ties go to the lower index on both sides).  Point
areas: the plane fit differs from the oracle's SVD by rounding only, so
A agrees to 1e-9 relative and T to 1e-9 (same orientation); the lattice closed
forms to 1e-12."""
import numpy as np
import pytest

import oracle_py as O
from knn_cases import K_AREAS, clustered, cube_corners, lattice_patch, surface_cloud, two_points
from libigl_amd import gwn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [cube_corners, two_points])
def test_knn_golden(case):
    V, P, k, I = case()
    point_indices, CH, _, _ = gwn.octree(V)
    assert np.array_equal(gwn.knn(P, V, k, point_indices, CH), I)
    assert np.array_equal(gwn.knn(P, V, k), I)


@pytest.mark.parametrize("k", [1, 5, 8, 21, 33, 64, 80])
def test_knn_vs_oracle(k):
    rng = np.random.default_rng(k)
    V = rng.random((3000, 3))
    Q = np.concatenate([V[:1000], rng.random((1000, 3)) * 1.6 - 0.3])
    I = gwn.knn(Q, V, k)
    assert I.shape == (Q.shape[0], k)
    assert np.array_equal(I, O.knn_bf(Q, V, k))
    if k in (1, 21):
        assert np.array_equal(I, O.knn(Q, V, k))


def test_knn_clustered_with_octree():
    V = clustered(4000)
    point_indices, CH, _, _ = gwn.octree(V)
    cand = gwn.octree_candidates(point_indices, CH)
    _, first = np.unique(V, axis=0, return_index=True)
    assert np.array_equal(cand, np.sort(first))  # coincident points: the first copy only
    I = gwn.knn(V, V, K_AREAS, point_indices, CH)
    Io = O.knn_bf(V, V, K_AREAS, cand)
    assert np.array_equal(I, Io)
    assert np.array_equal(I, O.knn(V, V, K_AREAS))  # the oracle's best-first search over its FP64 octree


def test_knn_edge_cases():
    rng = np.random.default_rng(3)
    V = rng.random((10, 3))
    # nv <= k: every point of V, nearest first
    I = gwn.knn(V[:4], V, 21)
    assert I.shape == (4, 10) and np.array_equal(I, O.knn_bf(V[:4], V, 21))
    assert gwn.knn(np.zeros((0, 3)), V, 3).shape == (0, 3)
    assert gwn.knn(V, np.zeros((0, 3)), 3).shape == (10, 0)
    with pytest.raises(gwn.GWNError):  # fewer candidates than k
        gwn.lib()  # loaded
        from libigl_amd import _native as nat
        I = np.zeros((10, 5), np.int32)
        c = np.array([0, 1, 2], np.int32)
        nat.check(gwn.lib().gwn_knn(V, 10, np.ascontiguousarray(V), 10, 5, c.ctypes.data, 3, I))
    with pytest.raises(gwn.GWNError):  # non-finite point
        Vb = V.copy()
        Vb[3, 1] = np.nan
        gwn.knn(V, Vb, 3)


def test_knn_large_sample():
    """1M surface points, k = 21 (the point-area pre-step), checked on a sample."""
    P, _ = surface_cloud(1_000_000, seed=9)
    I = gwn.knn(P, P, K_AREAS)
    sub = np.arange(0, P.shape[0], 4999)
    Io = O.knn_bf(P[sub], P, K_AREAS)
    assert np.array_equal(I[sub], Io)
    assert np.all(I[:, 0] == np.arange(P.shape[0]))  # each point is its own nearest


@pytest.mark.parametrize("stretch", [1.0, 1.15])
def test_point_areas_lattice(stretch):
    P, N, interior, area = lattice_patch(stretch=stretch)
    I = gwn.knn(P, P, K_AREAS)
    A, T = gwn.point_areas(P, I, N)
    assert np.allclose(A[interior], area, rtol=1e-12, atol=0)
    assert np.allclose(T[interior], N[interior], rtol=0, atol=1e-12)
    # the patch's border: hull edges through exactly collinear lattice points,
    # where Qhull may keep flat slivers (area ~1e-8 of the point's) that the
    # star leaves out as collinear, or a flat face whose NaN term zeroes the
    # oracle's sum (Ao = 0); the star gives such a corner the value of its
    # mirror image
    Ao, To = O.point_areas(P, I, N)
    ok = Ao > 0
    assert ok.sum() > 0.9 * ok.size
    assert np.allclose(A[ok], Ao[ok], rtol=1e-7, atol=0)
    assert np.all(A > 0)


@pytest.mark.parametrize("k", [K_AREAS, 12, 40])
def test_point_areas_vs_oracle(k):
    P, N = surface_cloud(4000, seed=k)
    I = O.knn(P, P, k)
    A, T = gwn.point_areas(P, I, N)
    Ao, To = O.point_areas(P, I, N)
    assert np.allclose(A, Ao, rtol=1e-9, atol=1e-15), np.max(np.abs(A - Ao) / np.maximum(Ao, 1e-300))
    assert np.allclose(T, To, rtol=0, atol=1e-9)
    # without normals: no filter, T unoriented
    A2, T2 = gwn.point_areas(P, I)
    Ao2, To2 = O.point_areas(P, I, None)
    assert np.allclose(A2, Ao2, rtol=1e-9, atol=1e-15)
    assert np.allclose(np.abs(np.einsum("ij,ij->i", T2, To2)), 1.0, rtol=0, atol=1e-9)


def test_point_areas_filters_and_degenerate():
    P, N, interior, area = lattice_patch()
    I = gwn.knn(P, P, K_AREAS)
    c = int(np.nonzero(interior)[0][0])
    N2 = N.copy()
    N2[I[c, 1::2]] *= -1
    A2, _ = gwn.point_areas(P, I, N2)
    Ao2, _ = O.point_areas(P, I, N2)
    assert np.isclose(A2[c], Ao2[c], rtol=1e-12) and A2[c] != area
    A4, T4 = gwn.point_areas(P[:2], np.array([[0, 1], [1, 0]], np.int32), N[:2])
    assert np.all(A4 == 0) and np.all(T4 == 0)
    L = np.stack([np.linspace(0, 1, 8), np.zeros(8), np.zeros(8)], 1)
    A5, _ = gwn.point_areas(L, np.tile(np.arange(8, dtype=np.int32), (8, 1)), np.tile([0, 0, 1.0], (8, 1)))
    assert np.all(A5 == 0)
    with pytest.raises(gwn.GWNError):
        gwn.point_areas(P[:3], np.array([[0, 1, 5]] * 3, np.int32), N[:3])


def test_cloud_with_estimated_areas():
    """igl::copyleft::cgal::fast_winding_number(P,N,Q,...): areas from knn +
    point areas, then the cloud FWN; against the oracle's pipeline."""
    P, N = surface_cloud(20000, seed=4)
    rng = np.random.default_rng(4)
    Q = np.concatenate([rng.random((3000, 3)) * 3 - 1.5, P[:500] * 0.9, P[500:1000] * 1.1])
    c = gwn.PointCloudFWN.with_estimated_areas(P, N, 2)
    W = c.query(Q, 2.0)
    I = O.knn(P, P, K_AREAS)
    Ao, _ = O.point_areas(P, I, N)
    assert np.allclose(c.A, Ao, rtol=1e-9, atol=1e-15)
    Wo = O.Cloud(P, N, Ao, 2, np.float64).eval(Q, 2.0)
    assert np.max(np.abs(W - Wo)) < 1e-4
    assert np.all(np.abs(W[3000:3500] - 1) < 0.05) and np.all(np.abs(W[3500:]) < 0.05)
    assert np.allclose(gwn.fast_winding_number_estimated_areas(P, N, Q), W, rtol=0, atol=0)


def test_cloud_estimated_areas_coincident_points():
    """Exact copies: knn's candidates are the first copies only, on the GPU as
    in the oracle's octree search."""
    P, N = surface_cloud(6000, seed=8)
    rng = np.random.default_rng(8)
    dup, src = rng.choice(6000, 10, replace=False), rng.choice(6000, 10, replace=False)
    P[dup], N[dup] = P[src], N[src]
    c = gwn.PointCloudFWN.with_estimated_areas(P, N, 2)
    I = O.knn(P, P, K_AREAS)
    Ao, _ = O.point_areas(P, I, N)
    assert np.allclose(c.A, Ao, rtol=1e-9, atol=1e-15)
