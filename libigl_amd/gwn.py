"""Reference-shaped Python interface of the MI355X winding-number library.

Mirrors libigl's public functions on this path (names, argument meaning,
defaults and output shapes), each call going through the C ABI of libgwn.so:

  winding_number(V, F, O)                       winding_number.h:41-48
  fast_winding_number(V, F, Q)                  fast_winding_number.h:160-165  (order 2, accuracy_scale 2)
  fast_winding_number(V, F, order, bvh)         fast_winding_number.h:176-180
  fast_winding_number(bvh, accuracy_scale, Q)   fast_winding_number.h:191-196 / 203-206
  fast_winding_number(P, N, A, Q)               fast_winding_number.h:126-131  (order 2, beta 2)
  fast_winding_number(P, N, A, Q, order, beta)  fast_winding_number.h:111-118
  fast_winding_number(P, N, A, point_indices, CH, order) -> (CM, R, EC)   .h:44-53
  fast_winding_number(P, N, A, point_indices, CH, CM, R, EC, Q, beta)     .h:88-99
  octree(P) -> (point_indices, CH, CN, W)       octree.h
  SubtendedAngle(P, E, order).compute_angle(Q, accuracy_scale)
                                                UT_SubtendedAngle (FastWindingNumberForSoups.h:6315-6370)

Errors: the reference asserts; here invalid input raises GWNError (or
ValueError for shape errors caught before the C ABI).
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from ._native import (PREC_FP32, PREC_FP64, PREC_MIXED, SIGN_DEFAULT, SIGN_FAST_WINDING_NUMBER, SIGN_PSEUDONORMAL,
                      SIGN_UNSIGNED, SIGN_WINDING_NUMBER, TREE_LBVH, TREE_SAH4, GWNError, check, lib)

DBL_MIN, DBL_MAX = 2.2250738585072014e-308, 1.7976931348623157e308

__all__ = ["FastWindingNumberBVH", "PointCloudFWN", "SubtendedAngle", "WindingNumberAABB", "fast_winding_number",
           "winding_number", "winding_number_bruteforce", "octree",
           "PREC_FP32", "PREC_MIXED", "PREC_FP64", "TREE_SAH4", "TREE_LBVH", "GWNError"]


def _rows3(X, name, dtype):
    X = np.asarray(X)
    if X.ndim == 1 and X.shape[0] == 3:
        X = X.reshape(1, 3)
    if X.ndim != 2 or X.shape[1] != 3:
        raise ValueError(f"{name} should be n x 3")
    return np.ascontiguousarray(X, dtype=dtype)


def winding_number(V, F, O):
    """Exact generalized winding number, FP64 (igl::winding_number(V,F,O,W)):
    triangles in 3D (F n x 3) or edges of a 2D curve (F n x 2, V and O 2D)."""
    F = np.asarray(F)
    if F.ndim == 2 and F.shape[1] == 2:
        V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).reshape(-1, 2))
        E = np.ascontiguousarray(F, dtype=np.int32)
        O = np.ascontiguousarray(np.asarray(O, dtype=np.float64).reshape(-1, 2))
        W = np.empty(O.shape[0], np.float64)
        check(lib().gwn_exact_winding_number_2d(V, V.shape[0], E, E.shape[0], O, O.shape[0], W))
        return W
    V = _rows3(V, "V", np.float64)
    F = _rows3(F, "F", np.int32)
    O = _rows3(O, "O", np.float64)
    W = np.empty(O.shape[0], np.float64)
    check(lib().gwn_exact_winding_number(V, V.shape[0], F, F.shape[0], O, O.shape[0], W))
    return W


def winding_number_bruteforce(V, F, O):
    """igl::winding_number(V, F, p) for each point (winding_number.cpp:62-89):
    the exact FP64 solid angles summed over every face on the GPU (no
    hierarchy, no rand()); equal to winding_number(V, F, O) up to rounding."""
    V = _rows3(V, "V", np.float64)
    F = _rows3(F, "F", np.int32)
    O = _rows3(O, "O", np.float64)
    W = np.empty(O.shape[0], np.float64)
    if O.shape[0]:
        check(lib().gwn_exact_winding_number_bruteforce(V, V.shape[0], F, F.shape[0], O, O.shape[0], W))
    return W


class WindingNumberAABB:
    """igl::WindingNumberAABB<double,int> (WindingNumberAABB.h) with grow():
    the hierarchy winding_number(V,F,O,W) builds (the build draws rand() as
    the reference's does), kept on the GPU for repeated exact queries."""

    def __init__(self, V, F):
        V = _rows3(V, "V", np.float64)
        F = _rows3(F, "F", np.int32)
        h = nat.vp()
        check(lib().gwn_exact_tree_build(V, V.shape[0], F, F.shape[0], C_byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib().gwn_exact_tree_free(self._h)
            self._h = None

    def winding_number(self, O):
        O = _rows3(O, "O", np.float64)
        W = np.empty(O.shape[0], np.float64)
        if O.shape[0]:
            check(lib().gwn_exact_tree_query(self._h, O, O.shape[0], W))
        return W

    def nodes(self):
        """Per node in construction order: #faces, #cap triangles, fan apex
        (into the de-duplicated vertices), left child, right child."""
        import ctypes
        n, t = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gwn_exact_tree_info(self._h, ctypes.byref(n), ctypes.byref(t)))
        out = np.empty((n.value, 5), np.int64)
        if n.value:
            check(lib().gwn_exact_tree_nodes(self._h, out))
        return out


class FastWindingNumberBVH:
    """igl::FastWindingNumberBVH (fast_winding_number.h:140-152): owns the
    tree and moments on the GPU.  Not copyable; freed with the object."""

    def __init__(self):
        self._h = None
        self.order = 2

    def _free(self):
        if self._h is not None:
            lib().gwn_soup_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None:
            raise GWNError(nat.GWN_ERR_INVALID, "FastWindingNumberBVH not built")
        return self._h

    def build(self, V, F, order=2, tree=TREE_SAH4):
        V = np.asarray(V)
        F = _rows3(F, "F", np.int32)
        if V.ndim != 2 or V.shape[1] != 3:
            raise ValueError("V should be 3D")
        self._free()
        h = nat.vp()
        if V.dtype == np.float32:
            check(lib().gwn_soup_build_f32(np.ascontiguousarray(V), V.shape[0], F, F.shape[0], order, tree,
                                           C_byref(h)))
        else:
            check(lib().gwn_soup_build(np.ascontiguousarray(V, dtype=np.float64), V.shape[0], F, F.shape[0],
                                       order, tree, C_byref(h)))
        self._h = h.value
        self.order = order
        return self

    def info(self):
        i = nat.SoupInfo()
        check(lib().gwn_soup_get_info(self.handle, C_byref(i)))
        return i

    def nodes(self):
        n = self.info().num_nodes
        out = np.empty((n, 4), np.uint32)
        if n:
            check(lib().gwn_soup_copy_nodes(self.handle, out))
        return out

    def boxdata(self):
        n = self.info().num_nodes
        out = np.empty((n, 92), np.float32)
        if n:
            check(lib().gwn_soup_copy_boxdata(self.handle, out))
        return out

    def set_mixed(self, near_dist=1e-3, ambiguous_band=0.05):
        p = nat.MixedParams(near_dist, ambiguous_band)
        check(lib().gwn_soup_set_mixed_params(self.handle, C_byref(p)))

    def query(self, Q, accuracy_scale=2.0, precision=PREC_FP32):
        Q = _rows3(Q, "Q", np.float64)
        W = np.empty(Q.shape[0], np.float64)
        if Q.shape[0]:
            check(lib().gwn_soup_query(self.handle, Q, Q.shape[0], float(accuracy_scale), precision, W))
        return W

    def query_device(self, Q_ptr, nq, W_ptr, accuracy_scale=2.0, precision=PREC_FP32, stream=None):
        """Device pointers (e.g. libigl_amd.devmem.DeviceArray.data_ptr): Q float nq x 3, W double nq."""
        check(lib().gwn_soup_query_device(self.handle, Q_ptr, nq, float(accuracy_scale), precision, W_ptr, stream))

    def signed_distance(self, P, sign_type=SIGN_FAST_WINDING_NUMBER, lower_bound=DBL_MIN, upper_bound=DBL_MAX,
                        with_normals=False):
        """igl::signed_distance over this mesh: (S, I, C), and N (np x 3, the
        pseudonormal sign's normals) with with_normals."""
        P = _rows3(P, "P", np.float64)
        S = np.empty(P.shape[0])
        I = np.empty(P.shape[0], np.int32)
        C = np.empty((P.shape[0], 3))
        N = np.zeros((P.shape[0], 3)) if with_normals else None
        check(lib().gwn_signed_distance_with_normals(self.handle, P, P.shape[0], int(sign_type), float(lower_bound),
                                                     float(upper_bound), S, I, C,
                                                     N.ctypes.data if with_normals else None))
        return (S, I, C, N) if with_normals else (S, I, C)


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


class PointCloudFWN:
    """Device octree + expansion coefficients of the point-cloud path.

    PointCloudFWN(P, N, A, order) builds the octree (igl::octree) and the
    precompute; with N and A None it holds the octree alone.
    PointCloudFWN.from_octree(...) works over a caller's octree."""

    def __init__(self, P, N=None, A=None, order=2):
        P = _rows3(P, "P", np.float32)
        if order not in (0, 1, 2):
            raise ValueError("expansion_order must be 0, 1 or 2")
        h = nat.vp()
        if N is None and A is None:
            check(lib().gwn_cloud_build(P, None, None, P.shape[0], order, C_byref(h)))
        else:
            N = _rows3(N, "N", np.float32)
            A = np.ascontiguousarray(np.asarray(A).reshape(-1), dtype=np.float32)
            if not (P.shape[0] == N.shape[0] == A.shape[0]):
                raise ValueError("P, N, A sizes differ")
            check(lib().gwn_cloud_build(P, N.ctypes.data, A.ctypes.data, P.shape[0], order, C_byref(h)))
        self._h = h.value
        self.order = order
        self.n = P.shape[0]
        self.user_octree = False

    @classmethod
    def from_octree(cls, P, N, A, point_indices, CH, order=2, CM=None, R=None, EC=None):
        """Precompute (CM, R, EC None) or evaluation state (CM, R, EC given)
        over a caller's octree: point_indices (one index list per cell) and
        CH (ncells x 8, -1 for leaves), as igl::octree returns them."""
        P = _rows3(P, "P", np.float32)
        N = _rows3(N, "N", np.float32)
        A = np.ascontiguousarray(np.asarray(A).reshape(-1), dtype=np.float32)
        if not (P.shape[0] == N.shape[0] == A.shape[0]):
            raise ValueError("P, N, A sizes differ")
        if order not in (0, 1, 2):
            raise ValueError("expansion_order must be 0, 1 or 2")
        CH = np.ascontiguousarray(np.asarray(CH).reshape(-1, 8), dtype=np.int32)
        m = CH.shape[0]
        if len(point_indices) != m:
            raise ValueError("point_indices and CH disagree on the number of cells")
        lens = np.fromiter((len(x) for x in point_indices), dtype=np.int64, count=m)
        start = np.zeros(m + 1, np.int64)
        np.cumsum(lens, out=start[1:])
        flat = np.ascontiguousarray(
            np.concatenate([np.asarray(x, dtype=np.int64).reshape(-1) for x in point_indices])
            if m and start[-1] else np.zeros(1, np.int64), dtype=np.int32)
        given = [x is not None for x in (CM, R, EC)]
        if any(given) and not all(given):
            raise ValueError("give all of CM, R, EC or none")
        ptrs = [None, None, None]
        keep = []
        if all(given):
            nt = {0: 3, 1: 12, 2: 39}[order]
            CMa = np.ascontiguousarray(np.asarray(CM).reshape(m, 3), dtype=np.float32)
            Ra = np.ascontiguousarray(np.asarray(R).reshape(m), dtype=np.float32)
            ECa = np.ascontiguousarray(np.asarray(EC).reshape(m, nt), dtype=np.float32)
            keep = [CMa, Ra, ECa]
            ptrs = [x.ctypes.data for x in keep]
        self = cls.__new__(cls)
        h = nat.vp()
        check(lib().gwn_cloud_build_from_octree(P, N, A, P.shape[0], CH, m, start, flat, order, ptrs[0], ptrs[1],
                                                ptrs[2], C_byref(h)))
        self._h = h.value
        self.order = order
        self.n = P.shape[0]
        self.user_octree = True
        return self

    @classmethod
    def with_estimated_areas(cls, P, N, order=2):
        """The cloud of igl::copyleft::cgal::fast_winding_number(P,N,Q,...)
        (copyleft/cgal/fast_winding_number.cpp:16-43): areas from knn (k = 21)
        and point_areas, computed on the GPU; `self.A` holds them."""
        P = _rows3(P, "P", np.float64)
        N = _rows3(N, "N", np.float64)
        if P.shape[0] != N.shape[0]:
            raise ValueError("P and N sizes differ")
        if order not in (0, 1, 2):
            raise ValueError("expansion_order must be 0, 1 or 2")
        A = np.zeros(P.shape[0], np.float64)
        self = cls.__new__(cls)
        h = nat.vp()
        check(lib().gwn_cloud_build_estimated_areas(P, N, P.shape[0], order, A.ctypes.data, C_byref(h)))
        self._h = h.value
        self.order = order
        self.n = P.shape[0]
        self.user_octree = False
        self.A = A
        return self

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                lib().gwn_cloud_free(self._h)
                self._h = None
        except Exception:
            pass

    def ncells(self):
        import ctypes
        n = ctypes.c_int64()
        check(lib().gwn_cloud_num_cells(self._h, ctypes.byref(n)))
        return n.value

    def octree(self):
        m = self.ncells()
        CH = np.empty((m, 8), np.int32)
        CN = np.empty((m, 3), np.float32)
        W = np.empty(m, np.float32)
        cs = np.empty(m, np.int64)
        cc = np.empty(m, np.int32)
        pl = np.empty(max(self.n, 1), np.int32)
        check(lib().gwn_cloud_copy_octree(self._h, CH.ctypes.data, CN.ctypes.data, W.ctypes.data, cs.ctypes.data,
                                          cc.ctypes.data, pl.ctypes.data))
        return CH, CN, W, cs, cc, pl[: self.n]

    def expansion(self):
        m = self.ncells()
        nt = {0: 3, 1: 12, 2: 39}[self.order]
        CM = np.empty((m, 3), np.float32)
        R = np.empty(m, np.float32)
        EC = np.empty((m, nt), np.float32)
        check(lib().gwn_cloud_copy_expansion(self._h, CM.ctypes.data, R.ctypes.data, EC.ctypes.data))
        return CM, R, EC

    def query(self, Q, beta=2.0):
        Q = _rows3(Q, "Q", np.float32)
        W = np.empty(Q.shape[0], np.float32)
        if Q.shape[0]:
            check(lib().gwn_cloud_query(self._h, Q, Q.shape[0], float(beta), W))
        return W


def octree(P):
    """igl::octree(P, point_indices, CH, CN, W) computed on the GPU (FP32)."""
    c = PointCloudFWN(P, None, None, 0)
    CH, CN, W, cs, cc, pl = c.octree()
    point_indices = [np.sort(pl[cs[i]:cs[i] + cc[i]]).tolist() for i in range(CH.shape[0])]
    return point_indices, CH, CN, W


class WindingNumberAntipodalScene:
    """igl::WindingNumberAntipodalScene (WindingNumberAntipodalScene.h:48-256)
    on the GPU: the mesh's weighted open boundary, and its own tree for the
    signed ray crossings the reference asks of an intersector."""

    def __init__(self, V, F, x0=None):
        V = _rows3(V, "V", np.float64)
        F = np.ascontiguousarray(np.asarray(F).reshape(-1, 3), dtype=np.int32)
        xd = None if x0 is None else np.ascontiguousarray(np.asarray(x0, np.float64).reshape(3))
        h = nat.vp()
        check(lib().gwn_antipodal_build(V, V.shape[0], F, F.shape[0], None if xd is None else xd.ctypes.data,
                                        C_byref(h)))
        self._h = h.value

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                lib().gwn_antipodal_free(self._h)
                self._h = None
        except Exception:
            pass

    def info(self):
        import ctypes
        nb, nf = ctypes.c_int64(), ctypes.c_int64()
        x0 = np.zeros(3)
        check(lib().gwn_antipodal_info(self._h, ctypes.byref(nb), ctypes.byref(nf), x0.ctypes.data))
        return {"num_boundary_segments": nb.value, "num_faces": nf.value, "x0": x0}

    def winding_number(self, O):
        O = _rows3(O, "O", np.float64)
        W = np.zeros(O.shape[0])
        if O.shape[0]:
            check(lib().gwn_antipodal_query(self._h, O, O.shape[0], W))
        return W


def _rows2(X, name, dtype):
    X = np.asarray(X)
    if X.ndim == 1 and X.shape[0] == 2:
        X = X.reshape(1, 2)
    if X.ndim != 2 or X.shape[1] != 2:
        raise ValueError(f"{name} should be n x 2")
    return np.ascontiguousarray(X, dtype=dtype)


class SubtendedAngle:
    """UT_SubtendedAngle<float,float> (FastWindingNumberForSoups.h:6315-6370):
    the 2D fast winding number of a curve given as segments.  The constructor
    is init(nsegments, segment_points, npoints, positions, order) (7342-7604)
    with P np x 2 and E ne x 2; compute_angle(Q, accuracy_scale) is
    computeAngle (7700-7828) for each row of Q (float angles in radians; the
    winding number is angle / 2 pi)."""

    def __init__(self, P, E, order=2, tree=TREE_SAH4):
        P = _rows2(P, "P", np.float32)
        E = _rows2(np.asarray(E).reshape(-1, 2), "E", np.int32)
        h = nat.vp()
        check(lib().gwn_curve_build(P, P.shape[0], E, E.shape[0], int(order), tree, C_byref(h)))
        self._h = h.value
        self.order = order

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                lib().gwn_curve_free(self._h)
                self._h = None
        except Exception:
            pass

    def info(self):
        i = nat.CurveInfo()
        check(lib().gwn_curve_get_info(self._h, C_byref(i)))
        return i

    def nodes(self):
        n = self.info().num_nodes
        out = np.empty((n, 4), np.uint32)
        if n:
            check(lib().gwn_curve_copy_nodes(self._h, out))
        return out

    def boxdata(self):
        n = self.info().num_nodes
        out = np.empty((n, 48), np.float32)
        if n:
            check(lib().gwn_curve_copy_boxdata(self._h, out))
        return out

    def compute_angle(self, Q, accuracy_scale=2.0):
        Q = _rows2(Q, "Q", np.float32)
        out = np.empty(Q.shape[0], np.float32)
        if Q.shape[0]:
            check(lib().gwn_curve_query(self._h, Q, Q.shape[0], float(accuracy_scale), out))
        return out

    def query_device(self, Q_ptr, nq, angle_ptr, accuracy_scale=2.0, stream=None):
        """Device pointers (e.g. libigl_amd.devmem.DeviceArray.data_ptr): Q float nq x 2, angle float nq."""
        check(lib().gwn_curve_query_device(self._h, Q_ptr, nq, float(accuracy_scale), angle_ptr, stream))


def sah4_tree_2d(P, E):
    """Host-only reference-topology tree of a curve (BVH<4>::init<BOX_AREA,S,2>)."""
    import ctypes
    P = _rows2(P, "P", np.float32)
    E = _rows2(np.asarray(E).reshape(-1, 2), "E", np.int32)
    n = ctypes.c_int64()
    check(lib().gwn_sah4_tree_2d(P, P.shape[0], E, E.shape[0], None, 0, ctypes.byref(n)))
    out = np.empty((n.value, 4), np.uint32)
    check(lib().gwn_sah4_tree_2d(P, P.shape[0], E, E.shape[0], out.ctypes.data, n.value, ctypes.byref(n)))
    return out


def winding_number_antipodal(V, F, O):
    """igl::winding_number_antipodal(V,F,intersector,O,W) (the GPU counts the
    ray crossings itself; no intersector is needed)."""
    return WindingNumberAntipodalScene(V, F).winding_number(O)


def octree_candidates(point_indices, CH):
    """The points igl::knn can return over an octree (knn.cpp:101-123): the
    first point of every non-empty leaf reachable from cell 0."""
    CH = np.asarray(CH).reshape(-1, 8)
    out = []
    if CH.shape[0] == 0:
        return np.zeros(0, np.int32)
    stack = [0]
    while stack:
        c = stack.pop()
        if CH[c, 0] == -1:
            if len(point_indices[c]) > 0:
                out.append(int(point_indices[c][0]))
        else:
            stack.extend(int(x) for x in CH[c])
    return np.asarray(sorted(out), np.int32)


def knn(P, V, k, point_indices=None, CH=None):
    """igl::knn(P,V,k,point_indices,CH,CN,W,I) (knn.h:51-67): I, np x min(nv,k),
    the k points of V nearest to each point of P, nearest first (equal
    distances: lower index first).  With an octree (point_indices, CH, as
    igl::octree returns them over V) the candidates are the points its search
    can reach; without one, every point of V.  knn(P,k,...) is knn(P,P,k,...)."""
    P = _rows3(P, "P", np.float64)
    V = _rows3(V, "V", np.float64)
    k = int(k)
    if k < 1:
        raise ValueError("k must be positive")
    cols = lib().gwn_knn_columns(V.shape[0], k)
    I = np.zeros((P.shape[0], cols), np.int32)
    cand, nc = None, 0
    if point_indices is not None:
        c = octree_candidates(point_indices, CH)
        cand, nc = c, c.shape[0]
    keep = cand
    check(lib().gwn_knn(P, P.shape[0], V, V.shape[0], k, None if cand is None else cand.ctypes.data, nc, I))
    del keep
    return I


def point_areas(P, I, N=None):
    """igl::copyleft::cgal::point_areas(P,I,N,A,T) (point_areas.h:44-62) on the
    GPU: returns (A, T)."""
    P = _rows3(P, "P", np.float64)
    I = np.ascontiguousarray(np.asarray(I), dtype=np.int32)
    if I.ndim != 2 or I.shape[0] != P.shape[0]:
        raise ValueError("I must have one row per point")
    Nn = None if N is None else _rows3(N, "N", np.float64)
    if Nn is not None and Nn.shape[0] != P.shape[0]:
        raise ValueError("P and N sizes differ")
    A = np.zeros(P.shape[0], np.float64)
    T = np.zeros((P.shape[0], 3), np.float64)
    if P.shape[0] and I.shape[1]:
        check(lib().gwn_point_areas(P, P.shape[0], I, I.shape[1], None if Nn is None else Nn.ctypes.data, A,
                                    T.ctypes.data))
    return A, T


def fast_winding_number_estimated_areas(P, N, Q, expansion_order=2, beta=2.0):
    """igl::copyleft::cgal::fast_winding_number(P,N,Q,expansion_order,beta,WN)
    (copyleft/cgal/fast_winding_number.h): a cloud with normals but no areas;
    the areas are estimated on the GPU (knn + point_areas)."""
    return PointCloudFWN.with_estimated_areas(P, N, expansion_order).query(Q, beta)


def fast_winding_number(*args, **kw):
    """Overloads of igl::fast_winding_number, dispatched on the arguments."""
    n = len(args)
    if n >= 1 and isinstance(args[0], FastWindingNumberBVH):
        bvh, accuracy_scale, Q = args
        Qa = np.asarray(Q)
        single = Qa.ndim == 1
        W = bvh.query(Qa, accuracy_scale, kw.get("precision", PREC_FP32))
        return float(W[0]) if single else W
    if n == 4 and isinstance(args[3], FastWindingNumberBVH):
        V, F, order, bvh = args
        bvh.build(V, F, order, kw.get("tree", TREE_SAH4))
        return bvh
    if n == 3:
        V, F, Q = args
        bvh = FastWindingNumberBVH().build(V, F, 2, kw.get("tree", TREE_SAH4))
        return bvh.query(Q, 2.0, kw.get("precision", PREC_FP32))
    if n == 4:
        P, N, A, Q = args
        return PointCloudFWN(P, N, A, 2).query(Q, 2.0)
    if n == 6 and isinstance(args[3], (list, tuple)):
        P, N, A, point_indices, CH, order = args
        return PointCloudFWN.from_octree(P, N, A, point_indices, CH, order).expansion()
    if n == 6:
        P, N, A, Q, order, beta = args
        return PointCloudFWN(P, N, A, order).query(Q, beta)
    if n == 10:
        P, N, A, point_indices, CH, CM, R, EC, Q, beta = args
        order = {3: 0, 12: 1, 39: 2}[np.asarray(EC).shape[1]]
        return PointCloudFWN.from_octree(P, N, A, point_indices, CH, order, CM, R, EC).query(Q, beta)
    raise TypeError("no fast_winding_number overload takes these arguments")


def signed_distance(P, V, F, sign_type=SIGN_FAST_WINDING_NUMBER, lower_bound=DBL_MIN, upper_bound=DBL_MAX):
    """igl::signed_distance(P,V,F,sign_type[,lower,upper],S,I,C,N) for every
    sign type (the FWN sign uses an order-2 tree, as signed_distance.cpp:65-67
    builds it).  Returns (S, I, C), and N as well for SIGN_PSEUDONORMAL."""
    bvh = FastWindingNumberBVH().build(V, F, 2)
    return bvh.signed_distance(P, sign_type, lower_bound, upper_bound, with_normals=sign_type == SIGN_PSEUDONORMAL)


def read_triangle_mesh(path):
    """igl::read_triangle_mesh for .off / .obj: (V float64 nv x 3, F int32 nf x 3),
    polygons split into fans."""
    import ctypes
    pv, pf = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_int32)()
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    check(lib().gwn_read_triangle_mesh(str(path).encode(), ctypes.byref(pv), ctypes.byref(nv), ctypes.byref(pf),
                                       ctypes.byref(nf)))
    try:
        V = np.ctypeslib.as_array(pv, shape=(max(nv.value, 1) * 3,))[: 3 * nv.value].copy().reshape(-1, 3)
        F = np.ctypeslib.as_array(pf, shape=(max(nf.value, 1) * 3,))[: 3 * nf.value].copy().reshape(-1, 3)
    finally:
        lib().gwn_free(ctypes.cast(pv, ctypes.c_void_p))
        lib().gwn_free(ctypes.cast(pf, ctypes.c_void_p))
    return V, F


def sah4_tree(V, F):
    """Host-only reference-topology tree (4 uint32 per node, reference encoding)."""
    import ctypes
    V = _rows3(V, "V", np.float32)
    F = _rows3(F, "F", np.int32)
    n = ctypes.c_int64()
    check(lib().gwn_sah4_tree(V, V.shape[0], F, F.shape[0], None, 0, ctypes.byref(n)))
    out = np.empty((n.value, 4), np.uint32)
    check(lib().gwn_sah4_tree(V, V.shape[0], F, F.shape[0], out.ctypes.data, n.value, ctypes.byref(n)))
    return out
