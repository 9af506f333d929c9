"""GPU parity of the exact FP64 winding number (igl::winding_number) against
the oracle: the reference's hierarchy (caps for boxes that do not contain the
query) and the FP64 brute force over every face.  Tolerance: max |dW| <= 1e-6
(stated by the workload definition); observed differences are ~1e-13
(summation order and the device atan2)."""
import numpy as np
import pytest

import oracle_py as O
from libigl_amd import datagen as D
from libigl_amd import gwn

pytestmark = pytest.mark.gpu
TOL = 1e-6


def test_exact_full_workload_vs_oracle_subset():
    V, F, Q = D.exact_workload()  # 20,480 tris, 100k queries
    W = gwn.winding_number(V, F, Q)
    assert W.shape == (Q.shape[0],) and np.all(np.isfinite(W))
    sub = np.arange(0, Q.shape[0], 20)  # 5,000 queries through the oracle
    Wo = O.winding_number_bf(V, F, Q[sub])
    err = np.abs(W[sub] - Wo)
    print(f"exact: max|dW| {err.max():.3e} over {sub.size} queries")
    assert err.max() <= TOL
    assert np.array_equal(W[sub] > 0.5, Wo > 0.5)
    # whole batch: a closed, outward surface gives w in {0, 1} away from it
    r = np.linalg.norm(Q, axis=1)
    assert np.all(np.abs(W[r < 0.9] - 1.0) < 1e-9)
    assert np.all(np.abs(W[r > 1.1]) < 1e-9)


def test_exact_hierarchy_matches_oracle_tree():
    """The GPU hierarchy is the oracle's node for node (same rand() state),
    and its sums equal the oracle's hierarchical sums and the brute force."""
    V, F, Q = D.exact_workload()
    Q = Q[:20000]
    O.libc_srand(4242)
    tree = gwn.WindingNumberAABB(V, F)
    O.libc_srand(4242)
    otree = O.WNTree(V, F)
    assert np.array_equal(tree.nodes(), otree.nodes())
    W = tree.winding_number(Q)
    assert np.abs(W - otree.query(Q)).max() < 1e-12
    assert np.abs(W - gwn.winding_number_bruteforce(V, F, Q)).max() < 1e-12


def test_exact_hierarchy_result_independent_of_batch():
    """One wavefront per query: batches of 4,096+ queries take them in
    Hilbert order, smaller ones in input order, and a query's lanes add the
    same terms in the same order either way, so its result does not depend on
    the batch it is in: bit for bit."""
    V, F, Q = D.exact_workload()
    rng = np.random.default_rng(7)
    Q = np.concatenate([Q[:12000], D.near_surface_queries(V, F, 4000, 1e-4, rng), V[:50]])
    tree = gwn.WindingNumberAABB(V, F)
    W = tree.winding_number(Q)  # 16,050 queries, sorted
    Wl = np.concatenate([tree.winding_number(Q[k:k + 4000]) for k in range(0, Q.shape[0], 4000)])
    assert Q.shape[0] % 4000 != 0 and Q[16000:].shape[0] < 4096
    assert np.array_equal(W, Wl)
    assert np.array_equal(tree.winding_number(Q[::-1][:3000]), W[::-1][:3000])


def test_exact_hierarchy_1310720_tris_vs_brute_force():
    """The 1,310,720-triangle mesh: exact queries through the hierarchy within
    1e-6 of the sum over every face (observed ~1e-12), 80% uniform, 20% within
    1e-4 of the surface."""
    V, F = D.bumpy_sphere(8, 64, 44)
    assert F.shape[0] == 1_310_720
    rng = np.random.default_rng(45)
    Q = np.concatenate([D.uniform_queries(V.min(0), V.max(0), 16000, rng), D.near_surface_queries(V, F, 4000, 1e-4, rng)])
    tree = gwn.WindingNumberAABB(V, F)
    W = tree.winding_number(Q)
    Wb = gwn.winding_number_bruteforce(V, F, Q)
    err = np.abs(W - Wb)
    print(f"1.31M tris, {tree.nodes().shape[0]} nodes: max|W_tree - W_brute| {err.max():.3e}")
    assert err.max() <= TOL and err.max() < 1e-9
    assert np.array_equal(W > 0.5, Wb > 0.5)
    chk = np.arange(0, Q.shape[0], 400)
    assert np.abs(W[chk] - O.winding_number_bf(V, F, Q[chk])).max() < 1e-9


def test_exact_small_cases():
    # single tet (tests/include/igl/signed_distance.cpp:4-44)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    F = np.array([[0, 1, 3], [0, 2, 1], [0, 3, 2], [1, 2, 3]], np.int32)
    Q = np.array([[0.5, 0.5, 0.5], [0.1, 0.1, 0.1], [0, 0, 0], [0.2, 0.2, 0.0]])
    W = gwn.winding_number(V, F, Q)
    Wo = O.winding_number_bf(V, F, Q)
    assert np.abs(W - Wo).max() <= 1e-14
    assert abs(W[0]) < 1e-15 and abs(W[1] - 1) < 1e-15
    # empty mesh and empty query set
    assert np.all(gwn.winding_number(np.zeros((0, 3)), np.zeros((0, 3), np.int32), Q) == 0)
    assert gwn.winding_number(V, F, np.zeros((0, 3))).shape == (0,)


def test_exact_ragged_sizes():
    # query counts and face counts that are not multiples of the tile sizes
    V, F = D.noisy_sphere(2, 0.01, 3)
    for nf in (1, 7, 255, 257, 320):
        for nq in (1, 3, 511, 513):
            Q = np.random.default_rng(nf * 1000 + nq).uniform(-1.3, 1.3, (nq, 3))
            Wo = O.winding_number_bf(V, F[:nf], Q)
            for W in (gwn.winding_number(V, F[:nf], Q), gwn.winding_number_bruteforce(V, F[:nf], Q)):
                assert np.abs(W - Wo).max() <= 1e-12, (nf, nq)


def test_exact_split_over_triangle_chunks():
    # few queries against thousands of triangles: the kernel splits the
    # triangles into chunks (1,025 -> 2, 3,001 -> 3 of 1,001, 5,120 -> 5) and
    # adds the partial sums afterwards
    V, F = D.noisy_sphere(4, 0.01, 5)
    for nf in (1025, 3001, 5120):
        for nq in (1, 300, 4097):
            Q = np.random.default_rng(nf + 7 * nq).uniform(-1.3, 1.3, (nq, 3))
            W = gwn.winding_number_bruteforce(V, F[:nf], Q)
            Wo = O.winding_number_bf(V, F[:nf], Q)
            assert np.abs(W - Wo).max() <= 1e-12, (nf, nq)
            assert np.abs(gwn.winding_number(V, F[:nf], Q) - Wo).max() <= 1e-12, (nf, nq)


def test_exact_rejects_bad_index():
    V, F = D.noisy_sphere(1, 0.01, 3)
    F = F.copy()
    F[0, 0] = -1
    with pytest.raises(gwn.GWNError):
        gwn.winding_number(V, F, np.zeros((1, 3)))


def test_exact_2d_curves_vs_oracle():
    """2D curves (igl::winding_number with F.cols() == 2): GPU FP64 against the
    oracle's restatement of signed_angle summed over edges (pinned by closed
    forms), on a noisy closed star polygon and an open arc, plus the square's
    closed forms through the C ABI."""
    import oracle_py as Op
    rng = np.random.default_rng(3)
    n = 3000
    t = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = 1 + 0.3 * np.sin(7 * t) + 0.01 * rng.normal(size=n)
    V = np.stack([r * np.cos(t), r * np.sin(t)], 1)
    E = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    O = rng.uniform(-1.6, 1.6, (20000, 2))
    W = gwn.winding_number(V, E, O)
    Wo = Op.winding_number_2d(V, E, O)
    assert np.abs(W - Wo).max() < 1e-12
    inside = np.abs(Wo - 1) < 1e-9
    outside = np.abs(Wo) < 1e-9
    assert inside.sum() > 1000 and outside.sum() > 1000 and np.all(np.abs(W[outside]) < 1e-9)
    Wa = gwn.winding_number(V, E[: n // 3], O[:500])
    assert np.abs(Wa - Op.winding_number_2d(V, E[: n // 3], O[:500])).max() < 1e-12
    Vs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    Es = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    assert np.allclose(gwn.winding_number(Vs, Es, np.array([[0.5, 0.5], [3.0, 3.0]])), [1, 0], atol=1e-15)


def test_exact_2d_few_queries_many_edges():
    """A few queries against thousands of edges: the edges are split into
    chunks whose partial sums are added in chunk order."""
    import oracle_py as Op
    n = 12_000
    t = np.linspace(0, 2 * np.pi, n, endpoint=False)
    V = np.stack([np.cos(t) * (1 + 0.2 * np.cos(5 * t)), np.sin(t) * (1 + 0.2 * np.cos(5 * t))], 1)
    E = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    for nq in (1, 300):
        Q = np.random.default_rng(nq).uniform(-1.4, 1.4, (nq, 2))
        W = gwn.winding_number(V, E, Q)
        assert np.abs(W - Op.winding_number_2d(V, E, Q)).max() < 1e-12, nq

