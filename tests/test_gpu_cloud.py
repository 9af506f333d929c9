"""GPU parity of the point-cloud path (igl::octree + fast_winding_number(P,N,A,...)).

Against the FP32 instantiation of the oracle restatement (itself pinned to the
reference's golden values and brute force):
  * octree: CH, cell centres, widths and the point set of every cell identical;
  * CM and R bit-identical (FP32 sums in ascending point order, as the
    reference), EC within 2e-5 relative (accumulated in FP64 on the GPU);
  * W within 1e-4 at beta = 2 and identical 0.5 classification outside a 1e-4
    band;
  * beta <= 0 brute force within 1e-5.
"""
import numpy as np
import pytest

import oracle_py as O
from libigl_amd import datagen as D
from libigl_amd import gwn

pytestmark = pytest.mark.gpu


def _cloud(n, seed=1, level=3):
    V, F = D.noisy_sphere(level, 0.01, 11)
    rng = np.random.default_rng(seed)
    P, N, _ = D.sample_surface(V, F, n, rng)
    _, area = D.face_normals(V, F)
    A = np.full(n, area.sum() / n)
    return P.astype(np.float32), N.astype(np.float32), A.astype(np.float32)


def test_one_point_golden_fp32():
    # tests/include/igl/fast_winding_number.cpp:17-53 (double there; FP32 here)
    P = np.array([[0.1, 0.2, 0.3]])
    N = np.array([[4.0, 5.0, 6.0]])
    N /= np.linalg.norm(N)
    A = np.array([0.7])
    Q = np.array([[0, 0, 0], [4, -3, 2], [-1, 1, 3], [0, 5, -2]], float)
    gold = np.array([0.38779369004261133, -0.00041235296362485, -0.00362978253577090, -0.00041235296362485])
    W = gwn.fast_winding_number(P, N, A, Q)
    assert np.abs(W - gold).max() < 1e-6 * np.abs(gold).max()


@pytest.mark.parametrize("n", [2, 3, 17, 500, 20000])
def test_octree_identical_to_reference_rules(n):
    P, N, A = _cloud(n, seed=n)
    c = gwn.PointCloudFWN(P, N, A, 2)
    CH, CN, W, cs, cc, pl = c.octree()
    o = O.Octree(P, np.float32)
    assert np.array_equal(CH, o.CH)
    assert np.array_equal(CN, o.CN) and np.array_equal(W, o.W)
    assert np.array_equal(cc, o.cell_count)
    for cell in range(CH.shape[0]):
        assert np.array_equal(np.sort(pl[cs[cell]:cs[cell] + cc[cell]]), o.point_indices(cell))


@pytest.mark.parametrize("order", [0, 1, 2])
def test_expansion_coefficients(order):
    P, N, A = _cloud(5000, seed=7)
    c = gwn.PointCloudFWN(P, N, A, order)
    CM, R, EC = c.expansion()
    oc = O.Cloud(P, N, A, order, np.float32)
    ne = oc.oct.cell_count > 0
    assert np.array_equal(np.isnan(CM[:, 0]), ~ne)
    scale = np.abs(oc.EC[ne]).max()
    # CM is summed as the reference sums it (FP32, ascending point index), so
    # CM and R match bit for bit and every near/far decision is the same
    assert np.array_equal(CM[ne], oc.CM[ne])
    assert np.array_equal(R[ne], oc.R[ne])
    assert np.abs(EC[ne] - oc.EC[ne]).max() < 2e-5 * scale
    assert np.all(EC[~ne] == 0) and np.all(R[~ne] == 0)


@pytest.mark.parametrize("beta", [1.0, 2.0, 4.0])
def test_eval_vs_oracle(beta):
    P, N, A = _cloud(20000, seed=3)
    rng = np.random.default_rng(4)
    V, F = D.noisy_sphere(3, 0.01, 11)
    Q = np.concatenate([rng.uniform(-1.3, 1.3, (4000, 3)), D.near_surface_queries(V, F, 2000, 1e-2, rng)])
    Q = Q.astype(np.float32)
    W = gwn.PointCloudFWN(P, N, A, 2).query(Q, beta)
    Wo = O.Cloud(P, N, A, 2, np.float32).eval(Q, beta)
    err = np.abs(W.astype(np.float64) - Wo)
    print(f"cloud beta {beta}: max {err.max():.3e}")
    assert err.max() <= 1e-4
    m = np.abs(Wo - 0.5) > 1e-4
    assert np.array_equal(W[m] > 0.5, Wo[m] > 0.5)


def test_brute_force_branch_and_large_beta():
    P, N, A = _cloud(3000, seed=5)
    Q = np.random.default_rng(6).uniform(-1.3, 1.3, (1000, 3)).astype(np.float32)
    c = gwn.PointCloudFWN(P, N, A, 2)
    Wb = c.query(Q, 0.0)
    Wo = O.Cloud(P, N, A, 2, np.float32).eval(Q, 0.0)
    assert np.abs(Wb - Wo).max() < 1e-5
    # very large beta: every cell is near, the walk ends at leaves -> brute force
    assert np.abs(c.query(Q, 1e9) - Wb).max() < 1e-4


def test_cloud_edge_cases():
    Q = np.random.default_rng(1).uniform(-1, 1, (100, 3)).astype(np.float32)
    # empty cloud
    c0 = gwn.PointCloudFWN(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), 2)
    assert np.all(c0.query(Q) == 0)
    # one point (root leaf) and a query on the point (0.5, fast_winding_number.cpp:163-166)
    P = np.array([[0.1, 0.2, 0.3]], np.float32)
    N = np.array([[0.0, 0.0, 1.0]], np.float32)
    A = np.array([0.5], np.float32)
    c1 = gwn.PointCloudFWN(P, N, A, 2)
    assert c1.query(P)[0] == 0.5
    # duplicate points: their cell cannot be split; values stay finite
    P2, N2, A2 = _cloud(200, seed=9)
    P2 = np.concatenate([P2, P2[:5]])
    N2 = np.concatenate([N2, N2[:5]])
    A2 = np.concatenate([A2, A2[:5]])
    c2 = gwn.PointCloudFWN(P2, N2, A2, 2)
    W2 = c2.query(Q, 2.0)
    assert np.all(np.isfinite(W2))
    assert np.abs(W2 - c2.query(Q, 0.0)).max() < 0.05
    with pytest.raises(ValueError):
        gwn.PointCloudFWN(P, N, A, 3)


def test_cloud_workload_full_size():
    P, N, A, Q = D.cloud_workload()  # 2M points, 160^3 grid
    c = gwn.PointCloudFWN(P, N, A, 2)
    W = c.query(Q, 2.0)
    assert W.shape == (Q.shape[0],) and np.all(np.isfinite(W))
    sub = np.arange(0, Q.shape[0], 4096)
    oc = O.Cloud(P.astype(np.float32), N.astype(np.float32), A.astype(np.float32), 2, np.float32)
    Wo = oc.eval(Q[sub].astype(np.float32), 2.0)
    err = np.abs(W[sub] - Wo)
    print(f"cloud 2M: cells {c.ncells()}, max|dW| on {sub.size} grid queries {err.max():.3e}")
    assert err.max() <= 1e-4
    r = np.linalg.norm(Q, axis=1)
    assert np.all(np.abs(W[r < 0.8] - 1) < 0.05) and np.all(np.abs(W[r > 1.2]) < 0.05)
