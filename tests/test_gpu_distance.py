"""GPU igl::signed_distance (all five sign types) through the C ABI, against
the oracle's brute-force restatement (itself pinned by closed forms, the
single-tet golden value and, for the pseudonormals, by
tests/test_oracle_pseudonormal.py).

Tolerances: squared distances and closest points are FP64 on both sides and
agree to 1e-12 (relative); the closest face agrees except where several faces
are equally close (then the GPU's face must be at the same distance); the
fast-winding-number sign differs only through the FWN value (FP32, order 2,
accuracy_scale 2), within 1e-4 in w, so |dS| <= 2e-4 |S|; the winding-number
sign comes from the exact hierarchy on the GPU and from the brute-force sum in
the oracle, which agree to ~1e-12 in w, so |dS| <= 1e-11 max(1, |S|); the
pseudonormal sign is +-1 exactly and its normal agrees to 1e-12."""
import numpy as np
import pytest

import oracle_py as O
from libigl_amd import datagen as D
from libigl_amd import gwn

pytestmark = pytest.mark.gpu

SIGNS = (gwn.SIGN_UNSIGNED, gwn.SIGN_WINDING_NUMBER, gwn.SIGN_DEFAULT, gwn.SIGN_FAST_WINDING_NUMBER,
         gwn.SIGN_PSEUDONORMAL)


def _check(V, F, P, S, I, C, So, Io, Co, sign):
    nan = np.isnan(So)
    assert np.array_equal(np.isnan(S), nan)
    assert np.array_equal(I[nan], Io[nan]) and np.all(C[nan] == Co[nan])
    m = ~nan
    if sign == gwn.SIGN_FAST_WINDING_NUMBER:
        assert np.all(np.abs(S[m] - So[m]) <= 2e-4 * np.abs(So[m]) + 1e-12)
    elif sign in (gwn.SIGN_WINDING_NUMBER, gwn.SIGN_DEFAULT):
        assert np.all(np.abs(S[m] - So[m]) <= 1e-11 * np.maximum(1.0, np.abs(So[m])))
    else:
        assert np.all(np.abs(S[m] - So[m]) <= 1e-12 * np.maximum(1.0, np.abs(So[m])))
    same = I[m] == Io[m]
    assert np.allclose(C[m][same], Co[m][same], rtol=0, atol=1e-12)
    # ties: the GPU's face is as close as the oracle's
    for p in np.nonzero(m)[0][~same]:
        t = F[I[p]]
        d2, _ = O.point_tri_sqrdist(P[p], V[t[0]], V[t[1]], V[t[2]])
        to = F[Io[p]]
        do, _ = O.point_tri_sqrdist(P[p], V[to[0]], V[to[1]], V[to[2]])
        assert abs(d2 - do) <= 1e-12 * max(1.0, do), (p, d2, do)


def test_single_tet_golden():
    # tests/include/igl/signed_distance.cpp:4-44
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    F = np.array([[0, 1, 3], [0, 2, 1], [0, 3, 2], [1, 2, 3]], np.int32)
    for sign, tol in ((gwn.SIGN_PSEUDONORMAL, 1e-15), (gwn.SIGN_WINDING_NUMBER, 1e-15), (gwn.SIGN_DEFAULT, 1e-15),
                      (gwn.SIGN_UNSIGNED, 1e-15), (gwn.SIGN_FAST_WINDING_NUMBER, 1e-7)):
        S, I, C = gwn.signed_distance(np.array([[0.5, 0.5, 0.5]]), V, F, sign)[:3]
        assert abs(S[0] - np.sqrt(1 / 12)) < tol, (sign, S)
        assert I[0] == 3 and np.allclose(C[0], [1 / 3, 1 / 3, 1 / 3], atol=1e-15)
    N = gwn.signed_distance(np.array([[0.5, 0.5, 0.5]]), V, F, gwn.SIGN_PSEUDONORMAL)[3]
    assert np.allclose(N[0], np.full(3, 1 / np.sqrt(3)), atol=1e-15)
    with pytest.raises(gwn.GWNError):
        gwn.signed_distance(np.zeros((1, 3)), V, F, 7)


@pytest.mark.parametrize("sign", SIGNS)
def test_signed_distance_vs_oracle(sign):
    V, F = D.bumpy_sphere(4, 8, 5)  # 5,120 triangles
    rng = np.random.default_rng(sign)
    lo, hi = V.min(0) - 0.3, V.max(0) + 0.3
    P = np.concatenate([D.uniform_queries(lo, hi, 3000, rng), D.near_surface_queries(V, F, 1000, 1e-3, rng),
                        V[:50] + 0.0])  # points on vertices: ties between faces
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    S, I, C, N = bvh.signed_distance(P, sign, with_normals=True)
    So, Io, Co, No = O.signed_distance(V, F, P, sign, with_normals=True)
    _check(V, F, P, S, I, C, So, Io, Co, sign)
    if sign == gwn.SIGN_PSEUDONORMAL:
        # the same normal (at ties between faces the test lands on the shared
        # vertex's or edge's normal from either face); on this closed mesh the
        # sign is the winding number's
        assert np.allclose(N, No, rtol=0, atol=1e-12)
        Sw, _, _ = O.signed_distance(V, F, P, gwn.SIGN_WINDING_NUMBER)
        assert np.array_equal(np.sign(S), np.sign(Sw))
    else:
        assert np.all(N == 0)  # not written for the other sign types


def test_pseudonormal_cube_and_open_mesh():
    """Corner, edge and face normals of a triangulated cube (closed forms, as
    the oracle's test pins them), and an open mesh against the oracle."""
    import test_oracle_pseudonormal as T
    V, F = T.cube()
    r3, r2 = 1 / np.sqrt(3), 1 / np.sqrt(2)
    P = np.array([[1.5, -0.6, -0.7], [-0.4, 1.5, 1.6], [1.5, 1.5, 0.5], [0.5, -0.3, -0.3], [0.3, 0.6, 1.5],
                  [0.9, 0.4, 0.6]])
    S, I, C, N = gwn.signed_distance(P, V, F, gwn.SIGN_PSEUDONORMAL)
    want = np.array([[r3, -r3, -r3], [-r3, r3, r3], [r2, r2, 0], [0, -r2, -r2], [0, 0, 1], [1, 0, 0]])
    assert np.allclose(N, want, atol=1e-14)
    assert np.all(S[:5] > 0) and abs(S[5] + 0.1) < 1e-15
    Vo, Fo = D.noisy_sphere(3, 0.01, 8)
    Fo = Fo[Vo[Fo].mean(1)[:, 2] < 0.4].copy()  # open: a cap removed
    rng = np.random.default_rng(12)
    Po = rng.uniform(-1.3, 1.3, (3000, 3))
    S, I, C, N = gwn.FastWindingNumberBVH().build(Vo, Fo, 2).signed_distance(Po, gwn.SIGN_PSEUDONORMAL,
                                                                             with_normals=True)
    So, Io, Co, No = O.signed_distance(Vo, Fo, Po, gwn.SIGN_PSEUDONORMAL, with_normals=True)
    _check(Vo, Fo, Po, S, I, C, So, Io, Co, gwn.SIGN_PSEUDONORMAL)
    assert np.allclose(N, No, rtol=0, atol=1e-12)


@pytest.mark.parametrize("bounds", [(-0.05, 0.05), (0.02, 0.1), (-0.2, -0.01)])
def test_signed_distance_bounds(bounds):
    V, F = D.bumpy_sphere(3, 6, 9)
    rng = np.random.default_rng(11)
    P = D.uniform_queries(V.min(0) - 0.2, V.max(0) + 0.2, 2000, rng)
    S, I, C = gwn.FastWindingNumberBVH().build(V, F, 2).signed_distance(P, gwn.SIGN_WINDING_NUMBER, *bounds)
    So, Io, Co = O.signed_distance(V, F, P, gwn.SIGN_WINDING_NUMBER, *bounds)
    assert np.isnan(So).any() and (~np.isnan(So)).any()
    _check(V, F, P, S, I, C, So, Io, Co, gwn.SIGN_WINDING_NUMBER)


def test_signed_distance_edge_cases():
    V, F = D.icosphere(2)
    # empty mesh: signed_distance.cpp:397-403
    S, I, C = gwn.signed_distance(np.zeros((3, 3)), V, F[:0], gwn.SIGN_FAST_WINDING_NUMBER)
    assert np.all(np.isnan(S)) and np.all(I == -1) and np.all(np.isnan(C))
    # no points
    S, I, C = gwn.signed_distance(np.zeros((0, 3)), V, F)
    assert S.shape == (0,) and I.shape == (0,) and C.shape == (0, 3)
    # a single triangle, points in every region
    V1 = np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0]])
    F1 = np.array([[0, 1, 2]], np.int32)
    P = np.array([[0.2, 0.3, 2.0], [-1, -1, 0.5], [2, -1, 0], [-0.5, 3, 1], [0.5, -1, 0], [-1, 0.5, -2],
                  [0.8, 0.8, 0]])
    S, I, C = gwn.signed_distance(P, V1, F1, gwn.SIGN_UNSIGNED)
    So, Io, Co = O.signed_distance(V1, F1, P, gwn.SIGN_UNSIGNED)
    assert np.allclose(S, So, rtol=1e-15, atol=0) and np.array_equal(I, Io) and np.allclose(C, Co, atol=1e-15)


def test_closest_point_deep_interior_queries():
    """Points deep inside the closed 81,920-triangle surface: their wavefronts
    outgrow the closest-point walk's record budget and hand their queries to
    the walk with one wavefront per query (lanes sharing the branch and bound).
    Face, closest point and distance match the oracle's brute force, with
    points near the surface and a NaN point in the same batch."""
    V, F, _ = D.soup_workload(nq=10)
    rng = np.random.default_rng(17)
    d = rng.normal(size=(3000, 3))
    inner = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0, 0.3, (3000, 1))
    near = D.near_surface_queries(V, F, 60000, 1e-3, rng)
    P = np.concatenate([inner, near[:30000], [[np.nan, 0.0, 0.0]], near[30000:]])
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    for sign in (gwn.SIGN_UNSIGNED, gwn.SIGN_FAST_WINDING_NUMBER):
        S, I, C = bvh.signed_distance(P, sign)
        sub = np.concatenate([np.arange(3000), np.arange(3000, P.shape[0], 61), [33000]])
        So, Io, Co = O.signed_distance(V, F, P[sub], sign)
        _check(V, F, P[sub], S[sub], I[sub], C[sub], So, Io, Co, sign)
        assert np.all(S[:3000] < -0.3) if sign == gwn.SIGN_FAST_WINDING_NUMBER else np.all(S[:3000] > 0.3)


def test_signed_distance_large_batch():
    """1M points on the 81,920-triangle workload mesh, checked on a sample."""
    V, F, Q = D.soup_workload()
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    S, I, C = bvh.signed_distance(Q, gwn.SIGN_FAST_WINDING_NUMBER)
    sub = np.arange(0, Q.shape[0], 997)
    So, Io, Co = O.signed_distance(V, F, Q[sub], gwn.SIGN_FAST_WINDING_NUMBER)
    _check(V, F, Q[sub], S[sub], I[sub], C[sub], So, Io, Co, gwn.SIGN_FAST_WINDING_NUMBER)
    # a batch of one gives the same closest point as the point in the large
    # batch; S = (1 - 2|w|) d (signed_distance.cpp:121-123) with w from the
    # fast winding number, whose batches of one and of 1M sum the same terms
    # in different orders (they agree to FP32 rounding, 1e-5 in w)
    S1, I1, C1 = bvh.signed_distance(Q[12345:12346], gwn.SIGN_FAST_WINDING_NUMBER)
    assert I1[0] == I[12345] and np.array_equal(C1[0], C[12345])
    assert abs(S1[0] - S[12345]) <= 2e-5 * np.linalg.norm(C1[0] - Q[12345])
