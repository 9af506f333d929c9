"""GPU antipodal-method winding numbers (WindingNumberAntipodalScene,
winding_number_antipodal) through the C ABI, against the oracle's
restatement (brute-force signed crossings, oracle/oracle_antipodal.c) and the
exact FP64 winding number.

Tolerances: the boundary term is the same FP64 sum in the same segment order
on both sides (device and host atan2 may differ by an ulp): 1e-12; the
crossing counts are integers and must agree exactly, so W agrees to 1e-12;
against the exact winding number 1e-9 (the reference's own test asserts 1e-5
for random soups, tests/include/igl/embree/winding_number_antipodal.cpp:79-125)."""
import numpy as np
import pytest

import oracle_py as O
from libigl_amd import datagen as D
from libigl_amd import gwn

pytestmark = pytest.mark.gpu


def _soup(seed, nv=10, nf=30):
    rng = np.random.default_rng(seed)
    V = rng.uniform(-1, 1, (nv, 3))
    F = []
    while len(F) < nf:
        a, b, c = rng.integers(0, nv, 3)
        if a != b and b != c and a != c:
            F.append([a, b, c])
    return V, np.array(F, np.int32), rng.uniform(-1.5, 1.5, (100, 3))


def test_closed_mesh_integer_valued():
    V, F = D.bumpy_sphere(4, 8, 5)  # 5,120 triangles, closed
    rng = np.random.default_rng(0)
    Q = np.concatenate([D.uniform_queries(V.min(0) - 0.3, V.max(0) + 0.3, 20000, rng),
                        D.near_surface_queries(V, F, 2000, 1e-4, rng)])
    s = gwn.WindingNumberAntipodalScene(V, F)
    assert s.info()["num_boundary_segments"] == 0 and s.info()["num_faces"] == F.shape[0]
    W = s.winding_number(Q)
    assert np.all(W == np.round(W))  # no boundary: the signed crossing count
    Wo, ns = O.winding_number_antipodal(V, F, Q)
    assert ns == 0 and np.array_equal(W, Wo)
    We = O.winding_number_bf(V, F, Q[::10])
    assert np.max(np.abs(W[::10] - We)) < 1e-9


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_soup_matches_exact(seed):
    V, F, Q = _soup(seed)
    W = gwn.winding_number_antipodal(V, F, Q)
    Wo, ns = O.winding_number_antipodal(V, F, Q)
    assert ns > 0
    assert np.max(np.abs(W - Wo)) < 1e-12
    assert np.max(np.abs(W - O.winding_number_bf(V, F, Q))) < 1e-9


def test_single_triangle_and_open_mesh():
    # tests/include/igl/embree/winding_number_antipodal.cpp:48-77
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    F = np.array([[0, 1, 2]], np.int32)
    c = np.array([1 / 3, 1 / 3, 0])
    P = np.array([c + [0, 0, 0.5], c - [0, 0, 0.5]])
    W = gwn.winding_number_antipodal(V, F, P)
    assert abs(W[0]) > 1e-3 and abs(W[1]) > 1e-3 and W[0] * W[1] < 0
    assert np.max(np.abs(W - O.winding_number_bf(V, F, P))) < 1e-12
    # an open hemisphere of the bumpy sphere
    V2, F2 = D.bumpy_sphere(3, 6, 5)
    F2 = F2[V2[F2].mean(1)[:, 2] > 0]
    rng = np.random.default_rng(3)
    Q = D.uniform_queries(V2.min(0) - 0.3, V2.max(0) + 0.3, 5000, rng)
    W2 = gwn.winding_number_antipodal(V2, F2, Q)
    Wo2, ns = O.winding_number_antipodal(V2, F2, Q)
    assert ns > 0 and np.max(np.abs(W2 - Wo2)) < 1e-12
    assert np.max(np.abs(W2 - O.winding_number_bf(V2, F2, Q))) < 1e-9


@pytest.mark.parametrize("x0", [[0, 0, 1.0], [-1.0, 0, 0], [0.3, -0.5, 0.8]])
def test_other_directions(x0):
    V, F, Q = _soup(7)
    x = np.asarray(x0) / np.linalg.norm(x0)
    s = gwn.WindingNumberAntipodalScene(V, F, x)
    assert np.allclose(s.info()["x0"], x, rtol=0, atol=0)
    W = s.winding_number(Q)
    Wo, _ = O.winding_number_antipodal(V, F, Q, x)
    assert np.max(np.abs(W - Wo)) < 1e-12
    assert np.max(np.abs(W - O.winding_number_bf(V, F, Q))) < 1e-9


def test_large_mesh_and_edge_cases():
    """81,920 triangles, 1M queries: integer-valued, and the exact sum on a sample."""
    V, F, Q = D.soup_workload()
    W = gwn.WindingNumberAntipodalScene(V, F).winding_number(Q)
    assert np.all(W == np.round(W)) and set(np.unique(W)) <= {0.0, 1.0}
    sub = np.arange(0, Q.shape[0], 997)
    assert np.array_equal(W[sub], O.winding_number_antipodal(V, F, Q[sub])[0])
    We = O.winding_number_bf(V, F, Q[sub])
    assert np.max(np.abs(W[sub] - We)) < 1e-6  # near-surface queries: the exact FP64 sum is 1e-9-ish off integers
    # empty mesh and no queries
    assert np.all(gwn.winding_number_antipodal(V, F[:0], Q[:5]) == 0)
    assert gwn.WindingNumberAntipodalScene(V, F).winding_number(np.zeros((0, 3))).shape == (0,)
