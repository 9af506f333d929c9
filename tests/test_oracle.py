"""CPU tests pinning the oracle (no GPU).

The oracle is the checker for the HIP path, so it is itself checked against
things that do not depend on it: the compiled reference (oracle/_ref), the
reference's own golden values, closed forms and invariants.  Each check is
chosen so that a plausible mistake (dropped term, wrong sign, transposed
operand, wrong tree split) fails it.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from libigl_amd import datagen as D

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


# ---------------- exact path ----------------
def test_solid_angle_octant_closed_form():
    # the triangle (e_x, e_y, e_z) subtends 1/8 of the sphere from the origin:
    # w = 1/8 (sign: counter-clockwise seen from the origin's side is negative)
    a, b, c = np.eye(3)
    w = O.solid_angle(a, b, c, np.zeros(3))
    assert abs(abs(w) - 0.125) < 1e-15
    # orientation flip negates; a point beyond the plane sees the opposite sign
    assert abs(O.solid_angle(a, c, b, np.zeros(3)) + w) < 1e-15
    assert np.sign(O.solid_angle(a, b, c, np.ones(3))) == -np.sign(w)


def test_solid_angle_square_pyramid():
    # a unit square at height h seen from the origin: Omega = 4 asin(1/(1+4h^2)) ... use
    # the exact formula for a rectangle a x b at distance d: 4 atan(ab / (2d sqrt(4d^2+a^2+b^2)))
    d = 0.37
    V = np.array([[-0.5, -0.5, d], [0.5, -0.5, d], [0.5, 0.5, d], [-0.5, 0.5, d]])
    F = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    w = O.winding_number_bf(V, F, np.zeros((1, 3)))[0]
    omega = 4 * math.atan(1.0 / (2 * d * math.sqrt(4 * d * d + 2)))
    assert abs(abs(w) - omega / (4 * math.pi)) < 1e-14


def test_exact_closed_mesh_inside_outside():
    V, F = D.noisy_sphere(3, 0.01, 7)
    O_in = np.array([[0, 0, 0], [0.3, -0.2, 0.1], [-0.5, 0.4, 0.2]])
    O_out = np.array([[1.5, 0, 0], [0, -2, 0.3], [3, 3, 3]])
    w = O.winding_number_bf(V, F, np.concatenate([O_in, O_out]))
    assert np.allclose(w[:3], 1.0, atol=1e-12)
    assert np.allclose(w[3:], 0.0, atol=1e-12)
    w_flip = O.winding_number_bf(V, F[:, ::-1].copy(), O_in)
    assert np.allclose(w_flip, -1.0, atol=1e-12)


def test_single_tet_outside_point():
    # tests/include/igl/signed_distance.cpp:4-44: the tet, P=(0.5,0.5,0.5) is outside (w = 0)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    F = np.array([[0, 1, 3], [0, 2, 1], [0, 3, 2], [1, 2, 3]], np.int32)
    assert abs(O.winding_number_bf(V, F, np.array([[0.5, 0.5, 0.5]]))[0]) < 1e-15
    assert abs(O.winding_number_bf(V, F, np.array([[0.1, 0.1, 0.1]]))[0] - 1.0) < 1e-15


@needs_ref
def test_tri_angle_f64_matches_reference_formula():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, c, q = rng.normal(size=(4, 3))
        # the VOS det/atan2 form and the reference's normalised form agree
        w1 = O.solid_angle(a, b, c, q)
        w2 = O.ref().ref_solid_angle_tri_f64(a, b, c, q) / (4 * math.pi)
        assert abs(w1 - w2) < 1e-14
        assert abs(O.lib().oracle_solid_angle_tri_f64(a, b, c, q) / (4 * math.pi) - w1) < 1e-14


# ---------------- soup path ----------------
@needs_ref
@pytest.mark.parametrize("level,nb,seed", [(0, 0, 1), (1, 2, 2), (3, 8, 3), (5, 16, 4)])
def test_soup_oracle_bit_exact_vs_reference(level, nb, seed):
    V, F = D.bumpy_sphere(level, nb, seed) if nb else D.noisy_sphere(level, 0.01, seed)
    o, r = O.Soup(V, F), O.RefSoup(V, F)
    assert np.array_equal(o.nodes(), r.nodes())
    assert np.array_equal(o.boxdata().view(np.uint32), r.boxdata().view(np.uint32))
    rng = np.random.default_rng(seed)
    Q = np.concatenate([rng.uniform(-1.6, 1.6, (300, 3)), D.near_surface_queries(V, F, 300, 1e-3, rng)])
    for beta in (1.0, 2.0, 4.0):
        assert np.array_equal(o.omega(Q, beta).view(np.uint32), r.omega(Q, beta).view(np.uint32))


@needs_ref
@pytest.mark.parametrize("order", [0, 1])
def test_soup_oracle_lower_orders_vs_reference(order):
    V, F = D.bumpy_sphere(3, 8, 9)
    o, r = O.Soup(V, F, order), O.RefSoup(V, F, order)
    rng = np.random.default_rng(1)
    Q = rng.uniform(-1.6, 1.6, (400, 3))
    assert np.array_equal(o.omega(Q, 2.0).view(np.uint32), r.omega(Q, 2.0).view(np.uint32))


@needs_ref
def test_soup_oracle_degenerate_and_nan_inputs():
    V, F = D.noisy_sphere(2, 0.01, 5)
    V = V.copy()
    F = np.concatenate([F, [[0, 0, 1], [2, 2, 2]]]).astype(np.int32)  # zero-area triangles
    Vn = np.concatenate([V, [[np.nan, 0, 0]]])
    F2 = np.concatenate([F, [[0, 1, V.shape[0]]]]).astype(np.int32)  # a triangle with a NaN vertex
    for VV, FF in ((V, F), (Vn, F2)):
        o, r = O.Soup(VV, FF), O.RefSoup(VV, FF)
        assert np.array_equal(o.nodes(), r.nodes())
        Q = np.random.default_rng(2).uniform(-1.5, 1.5, (200, 3))
        oo, rr = o.omega(Q), r.omega(Q)
        assert np.array_equal(oo.view(np.uint32), rr.view(np.uint32))


def test_soup_oracle_closed_mesh_and_large_beta_is_exact():
    V, F = D.bumpy_sphere(3, 8, 12)
    s = O.Soup(V, F)
    rng = np.random.default_rng(4)
    Q = rng.uniform(-1.2, 1.2, (200, 3))
    w_exact = O.winding_number_bf(V, F, Q)
    # accuracy_scale -> infinity: every triangle is evaluated exactly (FP32)
    assert np.abs(s.query(Q, 1e6) - w_exact).max() < 2e-5
    # default accuracy: the Taylor far field is within a few 1e-3
    assert np.abs(s.query(Q, 2.0) - w_exact).max() < 1e-2


def test_soup_oracle_near_count():
    """The near-triangle count of the instrumented walk: every exact triangle
    at a huge distance, none at a point far outside, monotone in the distance,
    and by brute force (the sphere test over the walk's exact triangles equals
    the sphere test over all triangles once accuracy_scale makes every
    triangle exact)."""
    V, F = D.bumpy_sphere(3, 8, 12)
    s = O.Soup(V, F)
    rng = np.random.default_rng(5)
    Q = np.concatenate([rng.uniform(-1.2, 1.2, (60, 3)), D.near_surface_queries(V, F, 60, 1e-3, rng)])
    T = V[F].astype(np.float32)
    P = T.mean(1)
    half = np.maximum(np.abs(T.min(1) - P), np.abs(T.max(1) - P))
    r = np.sqrt((half.astype(np.float64) ** 2).sum(1))
    for q in Q.astype(np.float32):
        _, _, exact = s.count(q, 2.0)
        assert s.count_near(q, 2.0, 1e30) == exact
        n0, n1 = s.count_near(q, 2.0, 0.0), s.count_near(q, 2.0, 0.05)
        assert n0 <= n1 <= exact
        d = np.sqrt(((q.astype(np.float64) - P) ** 2).sum(1))
        bf = int((d < r + 0.05).sum())
        assert abs(s.count_near(q, 1e6, 0.05) - bf) <= 1  # one float rounding at the sphere's edge at most
    assert s.count_near(np.array([50.0, 0, 0], np.float32), 2.0, 0.05) == 0


# ---------------- cloud path ----------------
def test_cloud_one_point_golden():
    # tests/include/igl/fast_winding_number.cpp:17-53
    P = np.array([[0.1, 0.2, 0.3]])
    N = np.array([[4.0, 5.0, 6.0]])
    N /= np.linalg.norm(N)
    A = np.array([0.7])
    c = O.Cloud(P, N, A, 2)
    Q = np.array([[0, 0, 0], [4, -3, 2], [-1, 1, 3], [0, 5, -2]], float)
    gold = np.array([0.38779369004261133, -0.00041235296362485, -0.00362978253577090, -0.00041235296362485])
    assert np.abs(c.eval(Q, 2) - gold).max() < 1e-15


def test_octree_tiny_by_hand():
    # two points on a diagonal: root splits once; octants by (x>=c)+2(y>=c)+4(z>=c)
    P = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    t = O.Octree(P)
    assert t.t.ncells == 9
    assert list(t.CH[0]) == list(range(1, 9))
    assert np.allclose(t.CN[0], [0.5, 0.5, 0.5]) and t.W[0] == 1.0
    assert list(t.point_indices(1)) == [0] and list(t.point_indices(8)) == [1]
    assert np.allclose(t.CN[8], [0.75, 0.75, 0.75]) and t.W[8] == 0.5
    # three points, two close: centres along the x axis are 0.5, 0.25, 0.125
    # (0.1 stays in octant 0 each time) and 0.0625, where x=0.1 >= c moves
    # point 1 to octant 1: three nested splits below the root -> 1 + 4*8 cells
    P = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [1.0, 1.0, 1.0]])
    t = O.Octree(P)
    assert t.t.ncells == 33
    assert list(t.CH[1]) == list(range(9, 17))
    assert list(t.CH[9]) == list(range(17, 25))
    assert list(t.CH[17]) == list(range(25, 33))
    assert list(t.point_indices(25)) == [0] and list(t.point_indices(26)) == [1]
    assert list(t.point_indices(8)) == [2]
    assert np.allclose(t.CN[17], [0.0625] * 3) and t.W[17] == 0.125


def test_cloud_taylor_terms_closed_form():
    # a cluster of dipoles far away: expansion (order 2) ~ brute force, order 0 worse
    rng = np.random.default_rng(8)
    P = 0.05 * rng.normal(size=(64, 3))
    N = rng.normal(size=(64, 3))
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    A = rng.uniform(0.1, 1.0, 64)
    Q = np.array([[2.0, 0.3, -0.1], [-1.0, 1.5, 0.7]])
    brute = O.Cloud(P, N, A, 2).eval(Q, 0)
    e = {}
    for order in (0, 1, 2):
        w = O.Cloud(P, N, A, order).eval(Q, 2.0)
        e[order] = np.abs(w - brute).max()
    # each order removes one power of (cell radius / distance) ~ 1/20 here
    assert e[2] < 0.1 * e[1] and e[1] < 0.5 * e[0]


def test_cloud_count_closed_forms():
    """The instrumented cloud walk: at an enormous beta every point is a
    direct dipole term and no expansion is used; far outside the cloud the
    root's non-empty children are tested and expanded, nothing else."""
    rng = np.random.default_rng(9)
    P = rng.standard_normal((3000, 3))
    N = P / np.linalg.norm(P, axis=1)[:, None]
    A = np.full(3000, 4 * np.pi / 3000)
    c = O.Cloud(P, N, A, 2, np.float64)
    t, d, e = c.count(np.array([0.1, 0.2, 0.3]), 1e12)
    assert d == 3000 and e == 0
    t, d, e = c.count(np.array([1e4, 0.0, 0.0]), 2.0)
    assert d == 0 and e == t and 1 <= t <= 8


def test_cloud_large_beta_equals_brute_force():
    V, F = D.noisy_sphere(2, 0.01, 3)
    rng = np.random.default_rng(5)
    P, N, _ = D.sample_surface(V, F, 500, rng)
    A = np.full(500, 4 * np.pi / 500)
    c = O.Cloud(P, N, A, 2)
    Q = rng.uniform(-1.3, 1.3, (50, 3))
    assert np.abs(c.eval(Q, 1e9) - c.eval(Q, 0)).max() < 1e-12
    w = c.eval(Q, 2.0)
    inside = np.linalg.norm(Q, axis=1) < 0.8
    outside = np.linalg.norm(Q, axis=1) > 1.2
    assert np.all(np.abs(w[inside] - 1) < 0.2) and np.all(np.abs(w[outside]) < 0.2)


def test_parallel_for_slicing_matches_serial():
    V, F = D.noisy_sphere(2, 0.01, 3)
    s = O.Soup(V, F)
    Q = np.random.default_rng(1).uniform(-1.5, 1.5, (3000, 3))
    assert np.array_equal(s.query(Q, 2.0, nthreads=1), s.query(Q, 2.0, nthreads=7))


def test_oracle_octree_from_lists_matches_own_octree():
    """Plumbing of caller-made octrees: the oracle's own octree passed back as
    plain lists gives the same precompute and evaluation bit for bit."""
    rng = np.random.default_rng(3)
    P = rng.normal(size=(3000, 3)).astype(np.float32)
    N = (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32)
    A = np.full(3000, 1e-3, np.float32)
    c = O.Cloud(P, N, A, 2, np.float32)
    o = c.oct
    lists = [o.point_indices(k).tolist() for k in range(o.CH.shape[0])]
    CM, R, EC, ev = O.cloud_over(P, N, A, O.Octree.from_lists(o.CH, lists, np.float32), 2)
    ne = o.cell_count > 0
    assert np.array_equal(CM[ne], c.CM[ne]) and np.array_equal(R[ne], c.R[ne]) and np.array_equal(EC[ne], c.EC[ne])
    Q = rng.uniform(-2, 2, (200, 3)).astype(np.float32)
    assert np.array_equal(ev(Q, 2.0), c.eval(Q, 2.0))


# ---------------- signed distance ----------------

def test_point_triangle_closed_forms():
    """Every Voronoi region of the right triangle (0,0,0),(1,0,0),(0,1,0)
    (point_simplex_squared_distance.cpp:43-114) against hand-derived answers."""
    a, b, c = np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0., 1, 0])
    cases = [  # point, closest point
        ((0.2, 0.3, 2.0), (0.2, 0.3, 0.0)),    # face interior
        ((-1.0, -1.0, 0.5), (0.0, 0.0, 0.0)),  # vertex a
        ((2.0, -1.0, 0.0), (1.0, 0.0, 0.0)),   # vertex b
        ((-0.5, 3.0, 1.0), (0.0, 1.0, 0.0)),   # vertex c
        ((0.5, -1.0, 0.0), (0.5, 0.0, 0.0)),   # edge ab
        ((-1.0, 0.5, -2.0), (0.0, 0.5, 0.0)),  # edge ac
        ((0.8, 0.8, 0.0), (0.5, 0.5, 0.0)),    # edge bc
        ((0.25, 0.25, 0.0), (0.25, 0.25, 0.0)),  # on the face
    ]
    for p, q in cases:
        d2, cp = O.point_tri_sqrdist(p, a, b, c)
        assert np.allclose(cp, q, atol=1e-15), (p, cp)
        assert abs(d2 - np.sum((np.array(p) - q) ** 2)) < 1e-14


def test_point_triangle_vs_dense_sampling():
    """Random triangles: the closest point lies on the triangle, and no point
    of a fine barycentric grid is closer (up to the grid's resolution)."""
    rng = np.random.default_rng(3)
    m = 80
    u, v = np.meshgrid(np.linspace(0, 1, m), np.linspace(0, 1, m))
    keep = (u + v) <= 1.0
    u, v = u[keep], v[keep]
    for _ in range(50):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        d2, cp = O.point_tri_sqrdist(p, a, b, c)
        grid = a + np.outer(u, b - a) + np.outer(v, c - a)
        g2 = np.min(np.sum((grid - p) ** 2, axis=1))
        assert d2 <= g2 + 1e-12
        h = max(np.linalg.norm(b - a), np.linalg.norm(c - a)) / (m - 1)
        assert np.sqrt(g2) - np.sqrt(d2) <= 2 * h
        # cp is on the triangle's plane and inside it (barycentric >= 0)
        M = np.stack([b - a, c - a], 1)
        lam, *_ = np.linalg.lstsq(M, cp - a, rcond=None)
        assert np.allclose(M @ lam, cp - a, atol=1e-10) and lam.min() > -1e-10 and lam.sum() < 1 + 1e-10


def test_signed_distance_single_tet_golden():
    # tests/include/igl/signed_distance.cpp:4-44
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    F = np.array([[0, 1, 3], [0, 2, 1], [0, 3, 2], [1, 2, 3]], np.int32)
    P = np.array([[0.5, 0.5, 0.5]])
    for t, tol in ((1, 1e-15), (2, 1e-15), (3, 1e-15), (4, 1e-7)):
        S, I, C = O.signed_distance(V, F, P, t)
        assert abs(S[0] - np.sqrt(1 / 12)) < tol, (t, S)
        assert I[0] == 3 and np.allclose(C[0], [1 / 3, 1 / 3, 1 / 3], atol=1e-15)


def test_signed_distance_signs_bounds_and_empty():
    V, F = D.icosphere(3)
    rng = np.random.default_rng(4)
    P = rng.uniform(-1.4, 1.4, (300, 3))
    r = np.linalg.norm(P, axis=1)
    ok = np.abs(r - 1) > 0.05
    Su, _, _ = O.signed_distance(V, F, P, 3)
    assert np.all(Su >= 0)
    # exact winding number: s = 1 - 2w = -1 / +1 to rounding; the fast winding
    # number (order 2, accuracy_scale 2) is within ~1e-3 of 0 or 1
    for t, rtol in ((1, 1e-12), (4, 5e-3)):
        S, _, _ = O.signed_distance(V, F, P, t)
        assert np.all(np.sign(S[ok]) == np.where(r[ok] < 1, -1, 1))
        assert np.allclose(np.abs(S[ok]), Su[ok], rtol=rtol)
    # bounds: only points with distance in [0.1, 0.3) keep a value
    S, I, C = O.signed_distance(V, F, P, 3, lower=-0.3, upper=0.3)
    # signed_distance.cpp:79-84: up = 0.3^2, low = max(0.3 - 0.6, 0)^2 = 0
    inside = Su < 0.3
    assert np.all(np.isnan(S[~inside])) and np.all(I[~inside] == F.shape[0] + 1) and np.all(C[~inside] == 0)
    assert np.allclose(S[inside], Su[inside])
    S, I, C = O.signed_distance(V, F[:0], P[:5], 4)
    assert np.all(np.isnan(S)) and np.all(I == -1) and np.all(np.isnan(C))


def test_winding_number_2d_closed_forms():
    """2D curves (winding_number.cpp with F.cols() == 2): a counter-clockwise
    square is 1 inside and 0 outside, clockwise is -1; a point on the curve's
    vertex contributes 0 for that edge end; half of a closed curve seen from its
    centre is 1/2; a circle polygon matches the angle it subtends."""
    V = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    E = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    Q = np.array([[0.5, 0.5], [0.1, 0.9], [2, 0.5], [-1, -1], [0.5, 3]])
    W = O.winding_number_2d(V, E, Q)
    assert np.allclose(W, [1, 1, 0, 0, 0], atol=1e-15)
    assert np.allclose(O.winding_number_2d(V, E[:, ::-1], Q), [-1, -1, 0, 0, 0], atol=1e-15)
    # half the square's boundary seen from its centre: the two edges cover 180 degrees
    assert abs(O.winding_number_2d(V, E[:2], Q[:1])[0] - 0.5) < 1e-15
    # an open arc of a circle from angle 0 to t seen from the centre: t / (2 pi)
    t = 2.0
    ang = np.linspace(0, t, 200)
    C = np.stack([np.cos(ang), np.sin(ang)], 1)
    Ea = np.stack([np.arange(199), np.arange(1, 200)], 1).astype(np.int32)
    assert abs(O.winding_number_2d(C, Ea, np.zeros((1, 2)))[0] - t / (2 * np.pi)) < 1e-14
