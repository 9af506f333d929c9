"""The exact winding number through the reference's hierarchy
(igl::WindingNumberAABB, winding_number.cpp:39-53), on the CPU.

The oracle's restatement (oracle/oracle_wntree.c) is pinned by:
  * the brute-force sum over every face (winding_number.cpp:62-89), which
    the hierarchy equals up to rounding (caps are exact for points outside a
    node's box), on closed, open, non-manifold and duplicated-vertex meshes;
  * closed forms: a closed mesh has no root cap, and W = 0 outside its box;
  * each node's cap closing its boundary: faces and cap give the same sum
    at points outside the node's box;
  * the library's host build of the same tree (std::nth_element and rand()
    called as the reference calls them), node for node.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from libigl_amd import _native as nat
from libigl_amd import datagen as D


def host_nodes(V, F):
    L = nat.lib()
    V = np.ascontiguousarray(V, np.float64)
    F = np.ascontiguousarray(F, np.int32)
    n = C.c_int64()
    nat.check(L.gwn_exact_tree_host_nodes(V, V.shape[0], F, F.shape[0], None, 0, C.byref(n)))
    out = np.empty((n.value, 5), np.int64)
    return out, V, F, n


def build_both(V, F, seed):
    out, V, F, n = host_nodes(V, F)
    O.libc_srand(seed)
    nat.check(nat.lib().gwn_exact_tree_host_nodes(V, V.shape[0], F, F.shape[0], out.ctypes.data, n.value,
                                                  C.byref(n)))
    O.libc_srand(seed)
    T = O.WNTree(V, F)
    return out, T


def open_hemisphere(level, seed):
    V, F = D.noisy_sphere(level, 0.01, seed)
    c = V[F].mean(1)
    return V, F[c[:, 2] > -0.3].copy()


def meshes():
    rng = np.random.default_rng(5)
    out = [("closed", *D.bumpy_sphere(4, 12, 3)), ("open", *open_hemisphere(4, 4))]
    Vr = rng.random((300, 3))
    out.append(("soup", Vr, rng.integers(0, 300, (2500, 3)).astype(np.int32)))
    Vs, Fs = D.noisy_sphere(3, 0.01, 6)
    Vd = Vs[Fs].reshape(-1, 3)  # every face with its own copies of its corners
    out.append(("duplicated", Vd, np.arange(Vd.shape[0], dtype=np.int32).reshape(-1, 3)))
    Vg, Fg = D.noisy_sphere(2, 0.0, 1)  # exact icosphere: many equal barycenter coordinates (median ties)
    out.append(("ties", Vg, Fg))
    return out


@pytest.mark.parametrize("name,V,F", meshes(), ids=[m[0] for m in meshes()])
def test_hierarchy_equals_brute_force_and_host_build(name, V, F):
    rng = np.random.default_rng(11)
    lo, hi = V.min(0), V.max(0)
    Q = np.concatenate([rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), (3000, 3)), V[:200] + 1e-9])
    for seed in (1, 12345):
        nodes, T = build_both(V, F, seed)
        assert np.array_equal(nodes, T.nodes()), name
        W = T.query(Q)
        Wb = O.winding_number_bf(V, F, Q)
        assert np.abs(W - Wb).max() < 1e-11, (name, np.abs(W - Wb).max())
    if name == "closed":
        assert nodes[0, 1] == 0 and nodes[0, 2] == -1  # no root cap
        far = np.array([[5.0, 0, 0], [0, -7.0, 1], [3, 3, 3]])
        assert np.all(T.query(far) == 0.0)
    if name in ("closed", "open", "duplicated"):
        assert nodes.shape[0] > 1  # split below the root
    if name == "soup":
        assert nodes[0, 1] - 2 >= F.shape[0]  # random triangles: the cap is not smaller, no split


def test_small_and_empty_meshes():
    V, F = D.noisy_sphere(1, 0.01, 2)  # 80 faces: a single leaf
    nodes, T = build_both(V, F, 3)
    assert nodes.shape == (1, 5) and nodes[0, 3] == -1
    Q = np.random.default_rng(1).uniform(-1.5, 1.5, (500, 3))
    assert np.abs(T.query(Q) - O.winding_number_bf(V, F, Q)).max() < 1e-13
    Te = O.WNTree(np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert np.all(Te.query(Q[:10]) == 0.0)
    one = O.WNTree(V, F[:1])
    assert np.abs(one.query(Q) - O.winding_number_bf(V, F[:1], Q)).max() < 1e-15


def test_caps_close_each_node():
    """For a node of the tree, sum over its faces = sum over its cap at
    points outside its box: checked through the hierarchy, which uses the cap
    there, against the brute force restricted to the faces of that node's
    subtree (a sub-mesh made of the node's faces is its own root)."""
    V, F = D.bumpy_sphere(4, 12, 8)
    F = F[V[F].mean(1)[:, 0] > 0.2].copy()  # an open piece: the root has a cap
    nodes, T = build_both(V, F, 9)
    assert nodes[0, 1] > 0
    lo, hi = V[np.unique(F)].min(0), V[np.unique(F)].max(0)
    rng = np.random.default_rng(2)
    P = rng.uniform(lo - 1.0, hi + 1.0, (4000, 3))
    out = np.any((P < lo) | (P > hi), axis=1)
    W = T.query(P[out])
    Wb = O.winding_number_bf(V, F, P[out])
    assert out.sum() > 1000 and np.abs(W - Wb).max() < 1e-12
