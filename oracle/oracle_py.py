"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the reference
build (oracle/_ref/libref_soup.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
(libigl_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_soup.so")
# the same reference sources built with its Release flags (-O3 -DNDEBUG): the
# CPU baseline that bench.py times
REF_RELEASE_SO = os.path.join(HERE, "_ref", "libref_soup_release.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")

_lib = None
_ref = {}


class OctreeT(C.Structure):
    _fields_ = [("ncells", C.c_int32), ("CH", C.POINTER(C.c_int32)), ("CN", C.c_void_p), ("W", C.c_void_p),
                ("cell_start", C.POINTER(C.c_int64)), ("cell_count", C.POINTER(C.c_int32)),
                ("point_list", C.POINTER(C.c_int32)), ("npoint_list", C.c_int64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.oracle_default_num_threads.restype = C.c_int
        L.oracle_solid_angle.restype = C.c_double
        L.oracle_solid_angle.argtypes = [_f64p, _f64p, _f64p, _f64p]
        L.oracle_solid_angle_tri_f64.restype = C.c_double
        L.oracle_solid_angle_tri_f64.argtypes = [_f64p, _f64p, _f64p, _f64p]
        L.oracle_winding_number_bf.argtypes = [_f64p, C.c_int64, _i32p, C.c_int64, _f64p, C.c_int64, _f64p, C.c_int]
        L.oracle_soup_build.restype = C.c_void_p
        L.oracle_soup_build.argtypes = [_f32p, C.c_int32, _i32p, C.c_int32, C.c_int]
        L.oracle_soup_build_nodes.restype = C.c_void_p
        L.oracle_soup_build_nodes.argtypes = [_f32p, C.c_int32, _i32p, C.c_int32, C.c_int, _u32p, C.c_int32]
        L.oracle_soup_free.argtypes = [C.c_void_p]
        L.oracle_soup_num_nodes.restype = C.c_int32
        L.oracle_soup_num_nodes.argtypes = [C.c_void_p]
        L.oracle_soup_nodes.argtypes = [C.c_void_p, _u32p]
        L.oracle_soup_boxdata.argtypes = [C.c_void_p, _f32p]
        L.oracle_soup_solid_angle.restype = C.c_float
        L.oracle_soup_solid_angle.argtypes = [C.c_void_p, _f32p, C.c_float]
        L.oracle_soup_query.argtypes = [C.c_void_p, _f64p, C.c_int64, C.c_float, _f64p, C.c_int]
        L.oracle_soup_count.argtypes = [C.c_void_p, _f32p, C.c_float, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.oracle_soup_count_near.argtypes = [C.c_void_p, _f32p, C.c_float, C.c_float] + [C.POINTER(C.c_int64)] * 4
        for sfx, fp in (("f64", _f64p), ("f32", _f32p)):
            f = getattr(L, f"oracle_octree_{sfx}")
            f.restype = C.c_int
            f.argtypes = [fp, C.c_int32, C.POINTER(OctreeT)]
            f = getattr(L, f"oracle_cloud_precompute_{sfx}")
            f.restype = C.c_int
            f.argtypes = [fp, fp, fp, C.POINTER(OctreeT), C.c_int, fp, fp, fp]
            f = getattr(L, f"oracle_cloud_eval_{sfx}")
            f.argtypes = [fp, fp, fp, C.POINTER(OctreeT), C.c_int, fp, fp, fp, fp, C.c_int64, C.c_double, fp, C.c_int]
            f = getattr(L, f"oracle_cloud_count_{sfx}")
            f.argtypes = [C.POINTER(OctreeT), fp, fp, fp, C.c_double] + [C.POINTER(C.c_int64)] * 3
        L.oracle_octree_release.argtypes = [C.POINTER(OctreeT)]
        L.oracle_curve_build.restype = C.c_void_p
        L.oracle_curve_build.argtypes = [_f32p, C.c_int32, _i32p, C.c_int32, C.c_int]
        L.oracle_curve_free.argtypes = [C.c_void_p]
        L.oracle_curve_num_nodes.restype = C.c_int32
        L.oracle_curve_num_nodes.argtypes = [C.c_void_p]
        L.oracle_curve_nodes.argtypes = [C.c_void_p, _u32p]
        L.oracle_curve_boxdata.argtypes = [C.c_void_p, _f32p]
        L.oracle_curve_angle.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_float, _f32p, C.c_int]
        L.oracle_curve_count.argtypes = [C.c_void_p, _f32p, C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64)]
        L.oracle_winding_number_antipodal.argtypes = [_f64p, C.c_int64, _i32p, C.c_int64, C.c_void_p, _f64p, C.c_int64,
                                                      _f64p, C.POINTER(C.c_int64), C.c_int]
        L.oracle_knn.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, C.c_int64, C.POINTER(OctreeT), _i32p, C.c_int]
        L.oracle_knn_bf.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, _i32p,
                                    C.c_int]
        L.oracle_winding_number_2d.argtypes = [_f64p, _i32p, C.c_int64, _f64p, C.c_int64, _f64p]
        L.oracle_point_tri_sqrdist.restype = C.c_double
        L.oracle_point_tri_sqrdist.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p]
        L.oracle_signed_distance.restype = C.c_int
        L.oracle_signed_distance.argtypes = [_f64p, C.c_int64, _i32p, C.c_int64, _f64p, C.c_int64, C.c_int,
                                             C.c_double, C.c_double, _f64p, _i32p, _f64p, C.c_int]
        L.oracle_signed_distance_n.restype = C.c_int
        L.oracle_signed_distance_n.argtypes = [_f64p, C.c_int64, _i32p, C.c_int64, _f64p, C.c_int64, C.c_int,
                                               C.c_double, C.c_double, _f64p, _i32p, _f64p, _f64p, C.c_int]
        L.oracle_pseudonormals.restype = C.c_int
        L.oracle_pseudonormals.argtypes = [_f64p, C.c_int64, _i32p, C.c_int64, _f64p, _f64p, _f64p]
        L.oracle_wntree_build.restype = C.c_void_p
        L.oracle_wntree_build.argtypes = [_f64p, C.c_int64, _i32p, C.c_int64]
        L.oracle_wntree_free.argtypes = [C.c_void_p]
        L.oracle_wntree_num_nodes.restype = C.c_int64
        L.oracle_wntree_num_nodes.argtypes = [C.c_void_p]
        L.oracle_wntree_nodes.argtypes = [C.c_void_p, _i64p]
        L.oracle_wntree_query.argtypes = [C.c_void_p, _f64p, C.c_int64, _f64p, C.c_int]
        L.oracle_wntree_count.restype = C.c_int64
        L.oracle_wntree_count.argtypes = [C.c_void_p, _f64p]
        _lib = L
    return _lib


def libc_srand(seed):
    """Seeds the C library's rand(), which the hierarchy build draws its fan
    apexes from (triangle_fan.cpp:31)."""
    C.CDLL(None).srand(C.c_uint(seed))


class WNTree:
    """The reference's exact winding-number hierarchy (WindingNumberAABB),
    restated; building it calls rand() as the reference does."""

    def __init__(self, V, F):
        self.V = _c(V, np.float64).reshape(-1, 3)
        self.F = _c(F, np.int32).reshape(-1, 3)
        self.h = lib().oracle_wntree_build(self.V, self.V.shape[0], self.F, self.F.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_wntree_free(self.h)
            self.h = None

    def nodes(self):
        """Per node in construction order: #faces, #cap triangles, fan apex,
        left, right."""
        n = lib().oracle_wntree_num_nodes(self.h)
        out = np.empty((n, 5), np.int64)
        lib().oracle_wntree_nodes(self.h, out)
        return out

    def query(self, O, nthreads=0):
        O = _c(O, np.float64).reshape(-1, 3)
        W = np.empty(O.shape[0], np.float64)
        lib().oracle_wntree_query(self.h, O, O.shape[0], W, nthreads)
        return W

    def count(self, p):
        """Solid angles the query of p evaluates."""
        return lib().oracle_wntree_count(self.h, _c(p, np.float64).reshape(3))


def ref_available(release=False) -> bool:
    return os.path.exists(REF_RELEASE_SO if release else REF_SO)


def ref(release=False):
    """The compiled reference: the parity build (-O2 -ffp-contract=off) or,
    with release=True, the Release-flag build."""
    if release not in _ref:
        path = REF_RELEASE_SO if release else REF_SO
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        R = C.CDLL(path)
        R.ref_soup_build.restype = C.c_void_p
        R.ref_soup_build.argtypes = [_f32p, C.c_int, _i32p, C.c_int, C.c_int]
        R.ref_soup_free.argtypes = [C.c_void_p]
        R.ref_soup_num_nodes.restype = C.c_int
        R.ref_soup_num_nodes.argtypes = [C.c_void_p]
        R.ref_soup_nodes.argtypes = [C.c_void_p, _u32p]
        R.ref_soup_boxdata.argtypes = [C.c_void_p, _f32p]
        R.ref_soup_solid_angle.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_float, _f32p]
        R.ref_soup_query.argtypes = [C.c_void_p, _f64p, C.c_int64, C.c_float, _f64p]
        R.ref_solid_angle_tri_f64.restype = C.c_double
        R.ref_solid_angle_tri_f64.argtypes = [_f64p, _f64p, _f64p, _f64p]
        R.ref_num_threads.restype = C.c_int
        R.ref_curve_build.restype = C.c_void_p
        R.ref_curve_build.argtypes = [_f32p, C.c_int, _i32p, C.c_int, C.c_int]
        R.ref_curve_free.argtypes = [C.c_void_p]
        R.ref_curve_num_nodes.restype = C.c_int
        R.ref_curve_num_nodes.argtypes = [C.c_void_p]
        R.ref_curve_nodes.argtypes = [C.c_void_p, _u32p]
        R.ref_curve_boxdata.argtypes = [C.c_void_p, _f32p]
        R.ref_curve_angle.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_float, _f32p]
        R.ref_curve_angle_par.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_float, _f32p]
        _ref[release] = R
    return _ref[release]


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------- 2D curves ----------------
def winding_number_2d(V, E, O):
    V = _c(V, np.float64).reshape(-1, 2)
    E = _c(E, np.int32).reshape(-1, 2)
    O = _c(O, np.float64).reshape(-1, 2)
    W = np.empty(O.shape[0])
    lib().oracle_winding_number_2d(V, E, E.shape[0], O, O.shape[0], W)
    return W


# ---------------- signed distance ----------------
def point_tri_sqrdist(p, a, b, c):
    """(squared distance, closest point) of triangle (a, b, c) to p."""
    out = np.zeros(3)
    d2 = lib().oracle_point_tri_sqrdist(_c(p, np.float64), _c(a, np.float64), _c(b, np.float64), _c(c, np.float64),
                                        out)
    return d2, out


def signed_distance(V, F, P, sign_type, lower=2.2250738585072014e-308, upper=1.7976931348623157e308, nthreads=0,
                    with_normals=False):
    """igl::signed_distance restated (brute force): returns S, I, C (and N
    with with_normals, written by the pseudonormal sign type)."""
    V = _c(V, np.float64).reshape(-1, 3)
    F = _c(F, np.int32).reshape(-1, 3)
    P = _c(P, np.float64).reshape(-1, 3)
    S = np.empty(P.shape[0])
    I = np.empty(P.shape[0], np.int32)
    C_ = np.empty((P.shape[0], 3))
    N = np.zeros((P.shape[0], 3))
    rc = lib().oracle_signed_distance_n(V, V.shape[0], F, F.shape[0], P, P.shape[0], int(sign_type), float(lower),
                                        float(upper), S, I, C_, N, nthreads)
    if rc != 0:
        raise ValueError("oracle_signed_distance failed")
    return (S, I, C_, N) if with_normals else (S, I, C_)


def pseudonormals(V, F):
    """(FN nf x 3, VN nv x 3, EN nf x 3 x 3) as signed_distance.cpp:68-76 builds them."""
    V = _c(V, np.float64).reshape(-1, 3)
    F = _c(F, np.int32).reshape(-1, 3)
    FN = np.empty((F.shape[0], 3))
    VN = np.empty((V.shape[0], 3))
    EN = np.empty((F.shape[0], 3, 3))
    if lib().oracle_pseudonormals(V, V.shape[0], F, F.shape[0], FN, VN, EN) != 0:
        raise MemoryError("oracle_pseudonormals")
    return FN, VN, EN


# ---------------- exact ----------------
def winding_number_bf(V, F, O, nthreads=0):
    V = _c(V, np.float64)
    F = _c(F, np.int32)
    O = _c(O, np.float64).reshape(-1, 3)
    W = np.empty(O.shape[0], np.float64)
    lib().oracle_winding_number_bf(V, V.shape[0], F, F.shape[0], O, O.shape[0], W, nthreads)
    return W


def solid_angle(a, b, c, p):
    return lib().oracle_solid_angle(_c(a, np.float64), _c(b, np.float64), _c(c, np.float64), _c(p, np.float64))


# ---------------- soup ----------------
class Soup:
    """Oracle UT_SolidAngle<float,float> (restated)."""

    def __init__(self, V, F, order=2, nodes=None):
        """nodes (n x 4 uint32, the reference's encoding): build the moments
        over that tree instead of the reference's own build."""
        self.U = _c(V, np.float32)
        self.F = _c(F, np.int32)
        if nodes is None:
            self.h = lib().oracle_soup_build(self.U, self.U.shape[0], self.F, self.F.shape[0], order)
        else:
            self._nodes = _c(nodes, np.uint32).reshape(-1, 4)
            self.h = lib().oracle_soup_build_nodes(self.U, self.U.shape[0], self.F, self.F.shape[0], order,
                                                   self._nodes, self._nodes.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_soup_free(self.h)
            self.h = None

    def nodes(self):
        n = lib().oracle_soup_num_nodes(self.h)
        out = np.empty((n, 4), np.uint32)
        lib().oracle_soup_nodes(self.h, out)
        return out

    def boxdata(self):
        n = lib().oracle_soup_num_nodes(self.h)
        out = np.empty((n, 92), np.float32)
        lib().oracle_soup_boxdata(self.h, out)
        return out

    def omega(self, Q, accuracy_scale=2.0):
        Q = _c(Q, np.float32).reshape(-1, 3)
        out = np.empty(Q.shape[0], np.float32)
        for i in range(Q.shape[0]):
            out[i] = lib().oracle_soup_solid_angle(self.h, Q[i].copy(), accuracy_scale)
        return out

    def query(self, Q, accuracy_scale=2.0, nthreads=0):
        Q = _c(Q, np.float64).reshape(-1, 3)
        W = np.empty(Q.shape[0], np.float64)
        lib().oracle_soup_query(self.h, Q, Q.shape[0], accuracy_scale, W, nthreads)
        return W

    def count(self, q, accuracy_scale=2.0):
        v, a, e = C.c_int64(), C.c_int64(), C.c_int64()
        lib().oracle_soup_count(self.h, _c(q, np.float32), accuracy_scale, C.byref(v), C.byref(a), C.byref(e))
        return v.value, a.value, e.value

    def count_near(self, q, accuracy_scale, near_dist):
        """Exact triangles of the query's walk within near_dist of their
        slot's bounding sphere (the mixed form's FP64 near terms)."""
        v, a, e, n = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        lib().oracle_soup_count_near(self.h, _c(q, np.float32), accuracy_scale, near_dist, C.byref(v),
                                     C.byref(a), C.byref(e), C.byref(n))
        return n.value


class RefSoup:
    """The reference's own UT_SolidAngle (compiled from the reference header);
    release=True uses the Release-flag build (the timed CPU baseline)."""

    def __init__(self, V, F, order=2, release=False):
        self.R = ref(release)
        self.U = _c(V, np.float32)
        self.F = _c(F, np.int32)
        self.h = self.R.ref_soup_build(self.U, self.U.shape[0], self.F, self.F.shape[0], order)

    def __del__(self):
        if getattr(self, "h", None):
            self.R.ref_soup_free(self.h)
            self.h = None

    def nodes(self):
        n = self.R.ref_soup_num_nodes(self.h)
        out = np.empty((n, 4), np.uint32)
        self.R.ref_soup_nodes(self.h, out)
        return out

    def boxdata(self):
        n = self.R.ref_soup_num_nodes(self.h)
        out = np.empty((n, 92), np.float32)
        self.R.ref_soup_boxdata(self.h, out)
        return out

    def omega(self, Q, accuracy_scale=2.0):
        Q = _c(Q, np.float32).reshape(-1, 3)
        out = np.empty(Q.shape[0], np.float32)
        self.R.ref_soup_solid_angle(self.h, Q, Q.shape[0], accuracy_scale, out)
        return out

    def query(self, Q, accuracy_scale=2.0):
        Q = _c(Q, np.float64).reshape(-1, 3)
        W = np.empty(Q.shape[0], np.float64)
        self.R.ref_soup_query(self.h, Q, Q.shape[0], accuracy_scale, W)
        return W


# ---------------- 2D curve ----------------
class _CurveBase:
    _prefix = None

    def _api(self):
        raise NotImplementedError

    def __init__(self, P, E, order=2):
        self.P = _c(P, np.float32).reshape(-1, 2)
        self.E = _c(E, np.int32).reshape(-1, 2)
        self.h = getattr(self._api(), self._prefix + "build")(self.P, self.P.shape[0], self.E, self.E.shape[0], order)

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self._api(), self._prefix + "free")(self.h)
            self.h = None

    def nodes(self):
        n = getattr(self._api(), self._prefix + "num_nodes")(self.h)
        out = np.empty((n, 4), np.uint32)
        getattr(self._api(), self._prefix + "nodes")(self.h, out)
        return out

    def boxdata(self):
        n = getattr(self._api(), self._prefix + "num_nodes")(self.h)
        out = np.empty((n, 48), np.float32)
        getattr(self._api(), self._prefix + "boxdata")(self.h, out)
        return out


class Curve(_CurveBase):
    """Oracle UT_SubtendedAngle<float,float> (restated)."""
    _prefix = "oracle_curve_"

    def _api(self):
        return lib()

    def angle(self, Q, accuracy_scale=2.0, nthreads=0):
        Q = _c(Q, np.float32).reshape(-1, 2)
        out = np.empty(Q.shape[0], np.float32)
        lib().oracle_curve_angle(self.h, Q, Q.shape[0], accuracy_scale, out, nthreads)
        return out

    def count(self, q, accuracy_scale=2.0):
        v, a, e = C.c_int64(), C.c_int64(), C.c_int64()
        lib().oracle_curve_count(self.h, _c(q, np.float32), accuracy_scale, C.byref(v), C.byref(a), C.byref(e))
        return v.value, a.value, e.value


class RefCurve(_CurveBase):
    """The reference's own UT_SubtendedAngle (compiled from the reference header)."""
    _prefix = "ref_curve_"
    _release = False

    def _api(self):
        return ref(self._release)

    def angle(self, Q, accuracy_scale=2.0):
        Q = _c(Q, np.float32).reshape(-1, 2)
        out = np.empty(Q.shape[0], np.float32)
        self._api().ref_curve_angle(self.h, Q, Q.shape[0], accuracy_scale, out)
        return out

    def angle_parallel(self, Q, accuracy_scale=2.0):
        """computeAngle over the host cores with the reference's igl::parallel_for."""
        Q = _c(Q, np.float32).reshape(-1, 2)
        out = np.empty(Q.shape[0], np.float32)
        self._api().ref_curve_angle_par(self.h, Q, Q.shape[0], accuracy_scale, out)
        return out


class RefCurveRelease(RefCurve):
    """RefCurve from the Release-flag build (the timed CPU baseline)."""
    _release = True


# ---------------- cloud ----------------
class Octree:
    def __init__(self, P, dtype=np.float64):
        self.dt = dtype
        self.sfx = "f64" if dtype == np.float64 else "f32"
        self.P = _c(P, dtype)
        self.t = OctreeT()
        rc = getattr(lib(), f"oracle_octree_{self.sfx}")(self.P, self.P.shape[0], C.byref(self.t))
        if rc != 0:
            raise MemoryError("oracle octree failed")
        n = self.t.ncells
        ctp = C.c_double if dtype == np.float64 else C.c_float
        self.CH = np.ctypeslib.as_array(self.t.CH, shape=(n, 8)).copy()
        self.CN = np.ctypeslib.as_array(C.cast(self.t.CN, C.POINTER(ctp)), shape=(n, 3)).copy()
        self.W = np.ctypeslib.as_array(C.cast(self.t.W, C.POINTER(ctp)), shape=(n,)).copy()
        self.cell_start = np.ctypeslib.as_array(self.t.cell_start, shape=(n,)).copy()
        self.cell_count = np.ctypeslib.as_array(self.t.cell_count, shape=(n,)).copy()
        m = max(self.t.npoint_list, 1)
        self.point_list = np.ctypeslib.as_array(self.t.point_list, shape=(m,))[: self.t.npoint_list].copy()

    @classmethod
    def from_lists(cls, CH, point_indices, dtype=np.float64):
        """An octree given by the caller (CH ncells x 8 and one index list per
        cell), for precompute/eval over arbitrary trees and list orders."""
        self = cls.__new__(cls)
        self.dt = dtype
        self.sfx = "f64" if dtype == np.float64 else "f32"
        self.CH = np.ascontiguousarray(CH, dtype=np.int32).reshape(-1, 8)
        m = self.CH.shape[0]
        self.cell_count = np.array([len(x) for x in point_indices], np.int32)
        self.cell_start = np.zeros(m, np.int64)
        self.cell_start[1:] = np.cumsum(self.cell_count, dtype=np.int64)[:-1]
        flat = [np.asarray(x, np.int64).reshape(-1) for x in point_indices]
        self.point_list = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros(0), dtype=np.int32)
        self._pl = self.point_list if self.point_list.size else np.zeros(1, np.int32)
        t = OctreeT()
        t.ncells = m
        t.CH = self.CH.ctypes.data_as(C.POINTER(C.c_int32))
        t.CN = None
        t.W = None
        t.cell_start = self.cell_start.ctypes.data_as(C.POINTER(C.c_int64))
        t.cell_count = self.cell_count.ctypes.data_as(C.POINTER(C.c_int32))
        t.point_list = self._pl.ctypes.data_as(C.POINTER(C.c_int32))
        t.npoint_list = int(self.point_list.size)
        self.t_borrowed = t  # arrays owned by numpy: never released by the oracle
        return self

    def point_indices(self, c):
        s = self.cell_start[c]
        return self.point_list[s: s + self.cell_count[c]]

    def __del__(self):
        if getattr(self, "t", None) is not None:
            lib().oracle_octree_release(C.byref(self.t))
            self.t = None


def cloud_over(P, N, A, octree, order=2, dtype=np.float32, CM=None, R=None, EC=None):
    """Precompute (or take CM, R, EC as given) over an Octree.from_lists tree;
    returns (CM, R, EC, eval) where eval(Q, beta) runs the restated
    evaluation over that tree."""
    sfx = "f64" if dtype == np.float64 else "f32"
    P, N = _c(P, dtype), _c(N, dtype)
    A = _c(A, dtype).reshape(-1)
    t = octree.t_borrowed
    m = t.ncells
    nt = {0: 3, 1: 12, 2: 39}[order]
    if CM is None:
        CM = np.empty((m, 3), dtype)
        R = np.empty(m, dtype)
        EC = np.empty((m, nt), dtype)
        getattr(lib(), f"oracle_cloud_precompute_{sfx}")(P, N, A, C.byref(t), order, CM, R, EC)
    else:
        CM, R, EC = _c(CM, dtype), _c(R, dtype).reshape(-1), _c(EC, dtype)

    def ev(Q, beta=2.0, _keep=octree):  # _keep: the tree's arrays must outlive t
        Q = _c(Q, dtype).reshape(-1, 3)
        W = np.empty(Q.shape[0], dtype)
        getattr(lib(), f"oracle_cloud_eval_{sfx}")(P, N, A, C.byref(t), order, CM, R, EC, Q, Q.shape[0],
                                                   float(beta), W, 0)
        return W
    return CM, R, EC, ev


class Cloud:
    """Oracle octree + precompute + evaluation (restated)."""

    def __init__(self, P, N, A, order=2, dtype=np.float64):
        self.dt = dtype
        self.sfx = "f64" if dtype == np.float64 else "f32"
        self.P = _c(P, dtype)
        self.N = _c(N, dtype)
        self.A = _c(A, dtype).reshape(-1)
        self.order = order
        self.oct = Octree(self.P, dtype)
        n = self.oct.t.ncells
        nt = {0: 3, 1: 12, 2: 39}[order]
        self.CM = np.empty((n, 3), dtype)
        self.R = np.empty(n, dtype)
        self.EC = np.empty((n, nt), dtype)
        getattr(lib(), f"oracle_cloud_precompute_{self.sfx}")(self.P, self.N, self.A, C.byref(self.oct.t), order,
                                                              self.CM, self.R, self.EC)

    def eval(self, Q, beta=2.0, nthreads=0):
        Q = _c(Q, self.dt).reshape(-1, 3)
        W = np.empty(Q.shape[0], self.dt)
        getattr(lib(), f"oracle_cloud_eval_{self.sfx}")(self.P, self.N, self.A, C.byref(self.oct.t), self.order,
                                                        self.CM, self.R, self.EC, Q, Q.shape[0], float(beta), W,
                                                        nthreads)
        return W

    def count(self, q, beta=2.0):
        """(distance tests, dipole terms, expansions) of one query's walk."""
        q = _c(q, self.dt).reshape(3)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        getattr(lib(), f"oracle_cloud_count_{self.sfx}")(C.byref(self.oct.t), self.CM, self.R, q, float(beta),
                                                         C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value


# ---------------- knn, point areas ----------------
def knn(P, V, k, octree=None, nthreads=0):
    """knn.cpp:24-127 restated: best-first search over `octree` (an Octree of V
    in FP64; built here when None).  I is np x min(nv, k)."""
    P = _c(P, np.float64).reshape(-1, 3)
    V = _c(V, np.float64).reshape(-1, 3)
    cols = min(V.shape[0], k)
    I = np.zeros((P.shape[0], cols), np.int32)
    if P.shape[0] == 0 or V.shape[0] == 0:
        return I
    if octree is None:
        octree = Octree(V, np.float64)
    rc = lib().oracle_knn(P, P.shape[0], V, V.shape[0], k, C.byref(octree.t), I, nthreads)
    if rc != 0:
        raise ValueError("oracle knn: fewer reachable points than k")
    return I


def knn_bf(P, V, k, cand=None, nthreads=0):
    """Brute-force k nearest over cand (indices into V; None: all), ordered by
    (squared distance, index)."""
    P = _c(P, np.float64).reshape(-1, 3)
    V = _c(V, np.float64).reshape(-1, 3)
    cols = min(V.shape[0], k)
    I = np.zeros((P.shape[0], cols), np.int32)
    if P.shape[0] == 0 or V.shape[0] == 0:
        return I
    c = None if cand is None else _c(cand, np.int32)
    rc = lib().oracle_knn_bf(P, P.shape[0], V, V.shape[0], k, None if c is None else c.ctypes.data,
                             0 if c is None else c.shape[0], I, nthreads)
    if rc != 0:
        raise ValueError("oracle knn_bf: fewer candidates than k")
    return I


def octree_candidates(octree):
    """The first point of every non-empty leaf reachable from cell 0
    (knn.cpp:113-117): the points igl::knn can return."""
    CH = octree.CH
    out, stack = [], [0]
    while stack:
        c = stack.pop()
        if CH[c, 0] == -1:
            if octree.cell_count[c] > 0:
                out.append(int(octree.point_list[octree.cell_start[c]]))
        else:
            stack.extend(int(x) for x in CH[c])
    return np.asarray(sorted(out), np.int32)


def _star_term(X, Y, Z):
    """point_areas.cpp:121-155 for the triangle (X, Y, Z) at X."""
    x = np.sqrt(np.sum((Y - Z) ** 2))
    y = np.sqrt(np.sum((X - Z) ** 2))
    z = np.sqrt(np.sum((Y - X) ** 2))
    cosX = (z * z + y * y - x * x) / (2 * y * z)
    cosY = (z * z + x * x - y * y) / (2 * x * z)
    cosZ = (x * x + y * y - z * z) / (2 * y * x)
    bary = np.array([x * cosX, y * cosY, z * cosZ])
    bary /= bary.sum()
    full = 0.25 * np.sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)))
    part = bary * full
    if cosX < 0:
        return 0.5 * full
    if cosY < 0 or cosZ < 0:
        return 0.25 * full
    return (part[1] + part[2]) / 2


def point_areas(P, I, N=None, rows=None):
    """point_areas.cpp:37-170 restated per point: the normal filter, the PCA
    plane (numpy's SVD in place of Eigen's JacobiSVD), a Delaunay triangulation
    of the projected neighbours (scipy / Qhull in place of CGAL; faces taken
    counter-clockwise as CGAL gives them) and the area sum over the faces at
    vertex 0.  Returns (A, T); with `rows`, only those points' entries are
    computed (the others stay 0)."""
    import warnings
    from scipy.spatial import Delaunay, QhullError
    with np.errstate(all="ignore"):  # degenerate faces give NaN sums, which count as 0
        return _point_areas(P, I, N, rows, Delaunay, QhullError, warnings)


def _point_areas(P, I, N, rows, Delaunay, QhullError, warnings):
    P = np.asarray(P, np.float64).reshape(-1, 3)
    I = np.asarray(I)
    n = P.shape[0]
    A = np.zeros(n)
    T = np.zeros((n, 3))
    for i in (range(n) if rows is None else rows):
        nb = P[I[i]]
        if N is not None and nb.shape[0] > 1:
            nn = np.asarray(N, np.float64)[I[i]]
            n0 = nn[0]
            dot = n0[0] * nn[:, 0] + n0[1] * nn[:, 1] + n0[2] * nn[:, 2]
            nb = nb[dot > 0]
        if nb.shape[0] <= 2:
            continue
        X = nb - nb.mean(axis=0)
        U, S, Vt = np.linalg.svd(X, full_matrices=False)
        t = Vt[2].copy()
        if N is not None and t @ np.asarray(N, np.float64)[i] < 0:
            t = -t
        T[i] = t
        plane = (U * S)[:, :2]
        try:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                tri = Delaunay(plane)
        except (QhullError, ValueError):
            continue  # collinear: no faces
        acc = 0.0
        for f in tri.simplices:
            a, b, c = plane[f[0]], plane[f[1]], plane[f[2]]
            if (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]) < 0:
                f = f[[0, 2, 1]]
            w = np.nonzero(f == 0)[0]
            if w.size == 0:
                continue
            X0 = int(w[0])
            acc += _star_term(plane[f[X0]], plane[f[(X0 + 1) % 3]], plane[f[(X0 + 2) % 3]])
        A[i] = acc if np.isfinite(acc) else 0.0
    return A, T


# ---------------- antipodal method ----------------
def winding_number_antipodal(V, F, O, x0=None, nthreads=0):
    """WindingNumberAntipodalScene restated (boundary term + brute-force signed
    ray crossings).  Returns (W, number of boundary segments)."""
    V = _c(V, np.float64).reshape(-1, 3)
    F = _c(F, np.int32).reshape(-1, 3)
    O = _c(O, np.float64).reshape(-1, 3)
    W = np.zeros(O.shape[0])
    ns = C.c_int64(0)
    xd = None if x0 is None else _c(x0, np.float64)
    rc = lib().oracle_winding_number_antipodal(V, V.shape[0], F, F.shape[0], None if xd is None else xd.ctypes.data,
                                               O, O.shape[0], W, C.byref(ns), nthreads)
    if rc != 0:
        raise MemoryError("oracle antipodal failed")
    return W, ns.value
