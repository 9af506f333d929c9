"""CPU tests pinning the 2D curve oracle (UT_SubtendedAngle restated in
oracle/oracle_curve.c).

Pins, independent of the restatement itself:
  - the compiled reference UT_SubtendedAngle (oracle/_ref), bit for bit: node
    arrays, BoxData (the fields each order defines) and the angle per query;
  - closed curves: 2 pi inside and 0 outside (exact evaluation, accuracy_scale
    large), orientation flip negates;
  - the exact segment angle against an FP64 atan2 sum;
  - the far field: each order's error falls off as r^-(order+2) with the
    distance r (a dropped or wrong term in the moments breaks the rate).
"""
import math

import numpy as np
import pytest

import oracle_py as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")

# BoxData columns each order defines: maxP2, P, N (20); + NijDiag, Nxy+Nyx (32); + third order (48)
ORDER_COLS = {0: 20, 1: 32, 2: 48}


def polygon(n, noise=0.05, seed=0, centre=(0.0, 0.0)):
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 2 * np.pi, n, endpoint=False)
    r = 1 + noise * rng.standard_normal(n)
    P = np.stack([r * np.cos(t), r * np.sin(t)], 1) + np.asarray(centre)
    E = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    return P, E


def exact_angles(P, E, Q):
    """FP64 sum of the signed segment angles (UTsignedAngleSegment, 6279-6307)."""
    a = P[E[:, 0]][None, :, :] - Q[:, None, :]
    b = P[E[:, 1]][None, :, :] - Q[:, None, :]
    num = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    den = a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]
    return np.arctan2(num, den).sum(1)


def curve_cases():
    yield polygon(1, seed=1)
    yield polygon(3, seed=2)
    yield polygon(7, seed=3)
    yield polygon(100, seed=4, centre=(0.5, -2.0))
    yield polygon(3000, seed=5)
    rng = np.random.default_rng(6)
    yield rng.random((400, 2)), rng.integers(0, 400, (1500, 2)).astype(np.int32)  # a random segment soup
    yield polygon(9000, seed=7, centre=(3.0, 1.0))  # >= 8192 boxes away from the origin


@needs_ref
@pytest.mark.parametrize("order", [0, 1, 2])
def test_curve_oracle_bit_exact_vs_reference(order):
    rng = np.random.default_rng(order)
    for P, E in curve_cases():
        o, r = O.Curve(P, E, order), O.RefCurve(P, E, order)
        assert np.array_equal(o.nodes(), r.nodes())
        k = ORDER_COLS[order]
        assert np.array_equal(o.boxdata()[:, :k].view(np.uint32), r.boxdata()[:, :k].view(np.uint32))
        lo, hi = P.min(0) - 0.5, P.max(0) + 0.5
        Q = np.concatenate([rng.uniform(lo, hi, (400, 2)), P[E[:, 0]][:50] + 1e-4])
        for beta in (1.0, 2.0, 4.0):
            assert np.array_equal(o.angle(Q, beta).view(np.uint32), r.angle(Q, beta).view(np.uint32))


@needs_ref
def test_curve_oracle_degenerate_inputs_vs_reference():
    P, E = polygon(300, seed=9)
    P = np.concatenate([P, [[np.nan, 0.0], [0.2, np.inf]]])
    n = 300
    E2 = np.concatenate([E, [[0, 0], [5, 5], [n, 1], [2, n + 1], [1, n]], E[:20]]).astype(np.int32)  # zero-length, NaN/inf, duplicates
    o, r = O.Curve(P, E2), O.RefCurve(P, E2)
    assert np.array_equal(o.nodes(), r.nodes())
    Q = np.concatenate([np.random.default_rng(1).uniform(-1.5, 1.5, (300, 2)), P[:5], [[0.0, 0.0]]])
    assert np.array_equal(o.angle(Q).view(np.uint32), r.angle(Q).view(np.uint32))
    # all segments at one point (max axis length 0)
    Pp = np.zeros((2, 2))
    Ep = np.array([[0, 1]] * 40, np.int32)
    o, r = O.Curve(Pp, Ep), O.RefCurve(Pp, Ep)
    assert np.array_equal(o.nodes(), r.nodes())
    assert np.array_equal(o.angle(Q), r.angle(Q))


def test_closed_curve_inside_outside():
    P, E = polygon(500, noise=0.02, seed=11)
    c = O.Curve(P, E)
    Qin = np.array([[0.0, 0.0], [0.3, -0.2], [-0.5, 0.4]])
    Qout = np.array([[1.6, 0.0], [0.0, -2.0], [3.0, 3.0]])
    big = 1e6  # accuracy_scale so large that every child is evaluated exactly
    a = c.angle(np.concatenate([Qin, Qout]), big)
    assert np.allclose(a[:3], 2 * np.pi, atol=2e-5) and np.allclose(a[3:], 0.0, atol=2e-5)
    flip = O.Curve(P, E[:, ::-1].copy())
    assert np.allclose(flip.angle(Qin, big), -2 * np.pi, atol=2e-5)
    # at accuracy_scale 2 the far field keeps inside/outside apart
    w = c.angle(np.concatenate([Qin, Qout]), 2.0) / (2 * np.pi)
    assert np.allclose(w[:3], 1.0, atol=2e-3) and np.allclose(w[3:], 0.0, atol=2e-3)


def test_exact_evaluation_matches_fp64_sum():
    rng = np.random.default_rng(12)
    P = rng.random((300, 2))
    E = rng.integers(0, 300, (800, 2)).astype(np.int32)
    Q = rng.uniform(-0.5, 1.5, (200, 2))
    a = O.Curve(P, E).angle(Q, 1e6)
    ref = exact_angles(P.astype(np.float32).astype(np.float64), E, Q.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(a - ref)) < 2e-4  # float segment angles summed in float


@pytest.mark.parametrize("order", [0, 1, 2])
def test_far_field_convergence_order(order):
    """A cluster of segments seen from distance r: the order-k expansion's
    error decays as r^-(k+2), so doubling r divides it by about 2^(k+2)."""
    rng = np.random.default_rng(20 + order)
    P = rng.random((64, 2)) - 0.5
    E = rng.integers(0, 64, (48, 2)).astype(np.int32)
    E = E[E[:, 0] != E[:, 1]]
    c = O.Curve(P, E, order)
    dirs = np.stack([np.cos(np.linspace(0.1, 6.2, 12)), np.sin(np.linspace(0.1, 6.2, 12))], 1)
    errs = []
    for r in (4.0, 8.0):
        Q = dirs * r
        approx = c.angle(Q, 1e-3).astype(np.float64)  # tiny accuracy_scale: the root's children are far
        errs.append(np.max(np.abs(approx - exact_angles(P, E, Q))))
    rate = errs[0] / errs[1]
    assert rate > 0.6 * 2 ** (order + 2), (order, errs)
    assert rate < 1.6 * 2 ** (order + 2) or errs[1] < 1e-6, (order, errs)


def test_visit_counts():
    P, E = polygon(2000, seed=13)
    c = O.Curve(P, E)
    v, a, e = c.count(np.array([5.0, 5.0], np.float32), 2.0)
    assert v >= 1 and e == 0 and a >= 1  # a distant query is answered by far fields
    v2, a2, e2 = c.count(np.array([1.0, 0.0], np.float32), 2.0)
    assert e2 > 0 and v2 > v
