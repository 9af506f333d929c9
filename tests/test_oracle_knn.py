"""The oracle's knn and point-area restatements (CPU), pinned against the
reference's golden vectors (tests/include/igl/knn.cpp), against each other
(best-first octree search vs brute force) and against closed forms."""
import numpy as np
import pytest

import oracle_py as O
from knn_cases import K_AREAS, clustered, cube_corners, lattice_patch, surface_cloud, two_points


@pytest.mark.parametrize("case", [cube_corners, two_points])
def test_knn_golden(case):
    V, P, k, I = case()
    assert np.array_equal(O.knn(P, V, k), I)
    assert np.array_equal(O.knn_bf(P, V, k), I)


def _dist(P, V, I):
    return np.sqrt(((V[I] - P[:, None, :]) ** 2).sum(-1))


@pytest.mark.parametrize("seed", [0, 1])
def test_best_first_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    V = rng.random((700, 3))
    Q = np.concatenate([V[:200], rng.random((200, 3)) * 1.4 - 0.2])
    tree = O.Octree(V, np.float64)
    cand = O.octree_candidates(tree)
    assert cand.shape[0] == V.shape[0]  # distinct points: every point is a leaf's
    for k in (1, 5, 21):
        Ib = O.knn(Q, V, k, tree)
        Ibf = O.knn_bf(Q, V, k, cand)
        assert np.array_equal(Ib, Ibf)
        d = _dist(Q, V, Ib)
        assert np.all(np.diff(d, axis=1) >= 0)
        # brute force over every point agrees on the distances
        dall = np.sort(np.sqrt(((V[None] - Q[:, None]) ** 2).sum(-1)), axis=1)[:, :k]
        assert np.allclose(d, dall, rtol=0, atol=1e-15)


def test_clustered_and_coincident():
    V = clustered(1500)
    tree = O.Octree(V, np.float64)
    cand = O.octree_candidates(tree)
    # a coincident point is reachable only through the first of its copies
    _, first = np.unique(V, axis=0, return_index=True)
    assert np.array_equal(np.sort(cand), np.sort(first))
    Ib = O.knn(V, V, K_AREAS, tree)
    Ibf = O.knn_bf(V, V, K_AREAS, cand)
    d1, d2 = _dist(V, V, Ib), _dist(V, V, Ibf)
    assert np.array_equal(d1, d2)  # same distances; indices may differ only on ties
    ties = np.nonzero((Ib != Ibf).any(1))[0]
    for i in ties:
        assert np.array_equal(np.sort(Ib[i]), np.sort(Ibf[i])) or np.any(np.diff(d1[i]) == 0)
    assert np.isin(Ib, cand).all()


def test_point_areas_lattice_closed_forms():
    for stretch in (1.0, 1.15):
        P, N, interior, area = lattice_patch(stretch=stretch)
        I = O.knn(P, P, K_AREAS)
        A, T = O.point_areas(P, I, N)
        assert interior.sum() > 20
        assert np.allclose(A[interior], area, rtol=1e-12, atol=0), (stretch, A[interior][:5], area)
        assert np.allclose(T[interior], N[interior], rtol=0, atol=1e-12)


def test_point_areas_filters_and_degenerate():
    P, N, interior, area = lattice_patch()
    I = O.knn(P, P, K_AREAS)
    c = int(np.nonzero(interior)[0][0])
    # neighbours whose normal opposes the first neighbour's are dropped:
    # same as triangulating the kept ones only
    N2 = N.copy()
    N2[I[c, 1::2]] *= -1
    A2, _ = O.point_areas(P, I, N2)
    keep = I[c, ::2]
    A3, _ = O.point_areas(P[keep], np.tile(np.arange(keep.size), (keep.size, 1)), N[keep])
    assert A2[c] == A3[0] and 0 < A2[c] != area
    # two kept neighbours or fewer: 0
    A4, T4 = O.point_areas(P[:2], np.array([[0, 1], [1, 0]]), N[:2])
    assert np.all(A4 == 0) and np.all(T4 == 0)
    # collinear neighbours: no triangle
    L = np.stack([np.linspace(0, 1, 8), np.zeros(8), np.zeros(8)], 1)
    A5, _ = O.point_areas(L, np.tile(np.arange(8), (8, 1)), np.tile([0, 0, 1.0], (8, 1)))
    assert np.all(A5 == 0)


def test_point_areas_surface_partition():
    """On a closed surface the areas add up to about the surface's area."""
    P, N = surface_cloud(3000)
    I = O.knn(P, P, K_AREAS)
    A, T = O.point_areas(P, I, N)
    from libigl_amd import datagen as D
    V, F = D.bumpy_sphere(3, 6, 5)
    total = D.face_normals(V, F)[1].sum()
    assert abs(A.sum() / total - 1) < 0.1
    assert np.all(np.einsum("ij,ij->i", T, N) >= 0)
