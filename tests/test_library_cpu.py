"""CPU tests of the product library that need no GPU: it loads, exports every
symbol include/gwn.h declares, and its host-side tree build reproduces the
reference topology."""
import os
import re

import numpy as np
import pytest

import oracle_py as O
from libigl_amd import _native as nat
from libigl_amd import datagen as D
from libigl_amd import gwn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gwn.h")).read()
    return sorted(set(re.findall(r"\b(gwn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = nat.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(nat.SIGNATURES), set(names) ^ set(nat.SIGNATURES)


def test_version_and_error_string():
    L = nat.lib()
    assert b"gfx950" in L.gwn_version()
    assert isinstance(L.gwn_last_error(), bytes)


def test_sah4_tree_matches_reference_nodes():
    for level, nb, seed in ((0, 0, 1), (2, 4, 2), (4, 16, 3), (6, 16, 22)):
        V, F = D.bumpy_sphere(level, nb, seed) if nb else D.noisy_sphere(level, 0.01, seed)
        ours = gwn.sah4_tree(V, F)
        theirs = (O.RefSoup(V, F) if O.ref_available() else O.Soup(V, F)).nodes()
        assert np.array_equal(ours, theirs)


@pytest.mark.parametrize("nf", [8191, 8192, 20000])
def test_sah4_tree_root_bound_off_origin(nf):
    """A mesh away from the origin: from 8192 boxes on, the reference's root
    bound (computeFullBoundingBox's parallel branch) also contains the origin,
    which moves the root split (FastWindingNumberForSoups.h:4856-4913)."""
    rng = np.random.default_rng(nf)
    V = rng.random((4000, 3)) + np.array([2.0, 0.5, -3.0])
    F = rng.integers(0, 4000, (nf, 3)).astype(np.int32)
    ref = O.RefSoup(V, F) if O.ref_available() else O.Soup(V, F)
    assert np.array_equal(gwn.sah4_tree(V, F), ref.nodes())
    Vs, Fs = D.bumpy_sphere(5, 16, 4)
    Vs = Vs + 10.0
    ref = O.RefSoup(Vs, Fs) if O.ref_available() else O.Soup(Vs, Fs)
    assert np.array_equal(gwn.sah4_tree(Vs, Fs), ref.nodes())


def test_sah4_tree_2d_matches_reference_nodes():
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 7, 33, 1000, 20000):
        t = np.linspace(0, 2 * np.pi, n, endpoint=False)
        P = np.stack([np.cos(t), np.sin(t)], 1) * (1 + 0.05 * rng.standard_normal((n, 1))) + 0.3
        E = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
        ref = O.RefCurve(P, E) if O.ref_available() else O.Curve(P, E)
        assert np.array_equal(gwn.sah4_tree_2d(P, E), ref.nodes())
    P = rng.random((3000, 2)) + np.array([4.0, 1.0])
    E = rng.integers(0, 3000, (12000, 2))
    ref = O.RefCurve(P, E) if O.ref_available() else O.Curve(P, E)
    assert np.array_equal(gwn.sah4_tree_2d(P, E), ref.nodes())
    assert gwn.sah4_tree_2d(np.zeros((0, 2)), np.zeros((0, 2), np.int32)).shape == (0, 4)
    with pytest.raises(nat.GWNError):
        gwn.sah4_tree_2d(P, np.array([[0, 3000]]))


def test_sah4_tree_degenerate_inputs():
    V, F = D.noisy_sphere(2, 0.01, 5)
    Vn = np.concatenate([V, [[np.nan, 0, 0]], [[np.inf, 1, 1]]])
    F2 = np.concatenate([F, [[0, 0, 1], [2, 2, 2], [0, 1, V.shape[0]], [3, 4, V.shape[0] + 1]]]).astype(np.int32)
    ref = O.RefSoup(Vn, F2) if O.ref_available() else O.Soup(Vn, F2)
    assert np.array_equal(gwn.sah4_tree(Vn, F2), ref.nodes())
    # tiny meshes: 1..5 triangles (single root node, exhaustive small split)
    for k in range(1, 6):
        FF = F[:k].copy()
        ref = O.RefSoup(V, FF) if O.ref_available() else O.Soup(V, FF)
        assert np.array_equal(gwn.sah4_tree(V, FF), ref.nodes())
    # all boxes at one point (max axis length 0)
    Vp = np.zeros((3, 3))
    Fp = np.array([[0, 1, 2]] * 40, np.int32)
    ref = O.RefSoup(Vp, Fp) if O.ref_available() else O.Soup(Vp, Fp)
    assert np.array_equal(gwn.sah4_tree(Vp, Fp), ref.nodes())


def test_sah4_tree_rejects_bad_faces():
    V, F = D.noisy_sphere(1, 0.01, 5)
    F = F.copy()
    F[3, 1] = V.shape[0] + 5
    with pytest.raises(nat.GWNError):
        gwn.sah4_tree(V, F)


def test_empty_mesh_tree():
    assert gwn.sah4_tree(np.zeros((0, 3)), np.zeros((0, 3), np.int32)).shape == (0, 4)


def test_datagen_sizes_and_determinism():
    V, F = D.icosphere(6)
    assert F.shape[0] == 81920 and V.shape[0] == 40962
    V1, F1, Q1 = D.soup_workload(nq=1000)
    V2, F2, Q2 = D.soup_workload(nq=1000)
    assert np.array_equal(V1, V2) and np.array_equal(Q1, Q2)
    # outward orientation: the exact winding number at the centre is +1
    w = O.winding_number_bf(*D.noisy_sphere(2, 0.01, 1), np.zeros((1, 3)))
    assert abs(w[0] - 1.0) < 1e-12
