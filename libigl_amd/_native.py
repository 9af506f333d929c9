"""ctypes binding of libgwn.so (the HIP library).  No fallback: if the library
is missing or a call fails, this raises.  The symbols bound here are exactly
the ones include/gwn.h declares."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgwn.so")

GWN_OK, GWN_ERR_INVALID, GWN_ERR_HIP, GWN_ERR_NOMEM, GWN_ERR_NODEVICE = 0, 1, 2, 3, 4
TREE_SAH4, TREE_LBVH = 0, 1
SIGN_PSEUDONORMAL, SIGN_WINDING_NUMBER, SIGN_DEFAULT, SIGN_UNSIGNED, SIGN_FAST_WINDING_NUMBER = 0, 1, 2, 3, 4
PREC_FP32, PREC_MIXED, PREC_FP64 = 0, 1, 2

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
vp = C.c_void_p


class SoupInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_int64), ("num_triangles", C.c_int64), ("num_nodes", C.c_int64),
                ("num_slots", C.c_int64), ("order", C.c_int32), ("tree_kind", C.c_int32), ("depth", C.c_int32),
                ("reserved", C.c_int32), ("device_bytes", C.c_int64), ("bbox", C.c_float * 6)]


class CurveInfo(C.Structure):
    _fields_ = [("num_points", C.c_int64), ("num_segments", C.c_int64), ("num_nodes", C.c_int64),
                ("num_slots", C.c_int64), ("order", C.c_int32), ("tree_kind", C.c_int32), ("depth", C.c_int32),
                ("reserved", C.c_int32), ("device_bytes", C.c_int64), ("bbox", C.c_float * 4)]


class MixedParams(C.Structure):
    _fields_ = [("near_dist", C.c_float), ("ambiguous_band", C.c_float)]


class RequestInfo(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("num_queries", C.c_int64), ("mesh", C.c_int32), ("status", C.c_int32),
                ("t_uploaded", C.c_double), ("t_device", C.c_double)]


# name -> (restype, argtypes); mirrors include/gwn.h
SIGNATURES = {
    "gwn_last_error": (C.c_char_p, []),
    "gwn_device_count": (C.c_int, []),
    "gwn_version": (C.c_char_p, []),
    "gwn_profile_enable": (C.c_int, [C.c_int]),
    "gwn_profile_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "gwn_profile_reset": (None, []),
    "gwn_profile_copy": (C.c_int, [vp, vp, C.c_int64, vp]),
    "gwn_exact_winding_number": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, f64p, C.c_int64, f64p]),
    "gwn_exact_winding_number_device": (C.c_int, [vp, C.c_int64, vp, C.c_int64, vp, C.c_int64, vp, vp]),
    "gwn_exact_winding_number_bruteforce": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, f64p, C.c_int64, f64p]),
    "gwn_exact_tree_build": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, C.POINTER(vp)]),
    "gwn_exact_tree_free": (None, [vp]),
    "gwn_exact_tree_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gwn_exact_tree_nodes": (C.c_int, [vp, i64p]),
    "gwn_exact_tree_query": (C.c_int, [vp, f64p, C.c_int64, f64p]),
    "gwn_exact_tree_query_device": (C.c_int, [vp, vp, C.c_int64, vp, vp]),
    "gwn_exact_tree_host_nodes": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, vp, C.c_int64,
                                            C.POINTER(C.c_int64)]),
    "gwn_soup_build": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]),
    "gwn_soup_build_f32": (C.c_int, [f32p, C.c_int64, i32p, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]),
    "gwn_soup_free": (None, [vp]),
    "gwn_soup_get_info": (C.c_int, [vp, C.POINTER(SoupInfo)]),
    "gwn_soup_query": (C.c_int, [vp, f64p, C.c_int64, C.c_float, C.c_int, f64p]),
    "gwn_soup_query_device": (C.c_int, [vp, vp, C.c_int64, C.c_float, C.c_int, vp, vp]),
    "gwn_soup_set_mixed_params": (C.c_int, [vp, C.POINTER(MixedParams)]),
    "gwn_soup_query_device_noalloc": (C.c_int, [vp, vp, C.c_int64, C.c_float, C.c_int, vp, vp]),
    "gwn_soup_noalloc_max_queries": (C.c_int64, []),
    "gwn_sah4_tree": (C.c_int, [f32p, C.c_int64, i32p, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64)]),
    "gwn_soup_copy_nodes": (C.c_int, [vp, u32p]),
    "gwn_soup_copy_boxdata": (C.c_int, [vp, f32p]),
    "gwn_cloud_build": (C.c_int, [f32p, vp, vp, C.c_int64, C.c_int, C.POINTER(vp)]),
    "gwn_cloud_build_from_octree": (C.c_int, [f32p, f32p, f32p, C.c_int64, i32p, C.c_int64, i64p, i32p, C.c_int,
                                              vp, vp, vp, C.POINTER(vp)]),
    "gwn_cloud_free": (None, [vp]),
    "gwn_cloud_num_cells": (C.c_int, [vp, C.POINTER(C.c_int64)]),
    "gwn_cloud_copy_octree": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "gwn_cloud_copy_expansion": (C.c_int, [vp, vp, vp, vp]),
    "gwn_cloud_query": (C.c_int, [vp, f32p, C.c_int64, C.c_double, f32p]),
    "gwn_cloud_query_device": (C.c_int, [vp, vp, C.c_int64, C.c_double, vp, vp]),
    "gwn_antipodal_build": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, vp, C.POINTER(vp)]),
    "gwn_antipodal_free": (None, [vp]),
    "gwn_antipodal_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), vp]),
    "gwn_antipodal_query": (C.c_int, [vp, f64p, C.c_int64, f64p]),
    "gwn_antipodal_query_device": (C.c_int, [vp, vp, C.c_int64, vp, vp]),
    "gwn_winding_number_antipodal": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, f64p, C.c_int64, f64p]),
    "gwn_curve_build": (C.c_int, [f32p, C.c_int64, i32p, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]),
    "gwn_curve_free": (None, [vp]),
    "gwn_curve_get_info": (C.c_int, [vp, C.POINTER(CurveInfo)]),
    "gwn_curve_query": (C.c_int, [vp, f32p, C.c_int64, C.c_float, f32p]),
    "gwn_curve_query_device": (C.c_int, [vp, vp, C.c_int64, C.c_float, vp, vp]),
    "gwn_curve_copy_nodes": (C.c_int, [vp, u32p]),
    "gwn_curve_copy_boxdata": (C.c_int, [vp, f32p]),
    "gwn_sah4_tree_2d": (C.c_int, [f32p, C.c_int64, i32p, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64)]),
    "gwn_knn_columns": (C.c_int, [C.c_int64, C.c_int64]),
    "gwn_knn": (C.c_int, [f64p, C.c_int64, f64p, C.c_int64, C.c_int64, vp, C.c_int64, i32p]),
    "gwn_knn_device": (C.c_int, [vp, C.c_int64, vp, C.c_int64, C.c_int64, vp, C.c_int64, vp, vp]),
    "gwn_point_areas": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, vp, f64p, vp]),
    "gwn_point_areas_device": (C.c_int, [vp, C.c_int64, vp, C.c_int64, vp, vp, vp, vp]),
    "gwn_cloud_build_estimated_areas": (C.c_int, [f64p, f64p, C.c_int64, C.c_int, vp, C.POINTER(vp)]),
    "gwn_signed_distance": (C.c_int, [vp, f64p, C.c_int64, C.c_int, C.c_double, C.c_double, f64p, i32p, f64p]),
    "gwn_signed_distance_device": (C.c_int, [vp, vp, C.c_int64, C.c_int, C.c_double, C.c_double, vp, vp, vp, vp]),
    "gwn_signed_distance_with_normals": (C.c_int, [vp, f64p, C.c_int64, C.c_int, C.c_double, C.c_double, f64p, i32p,
                                                   f64p, vp]),
    "gwn_signed_distance_with_normals_device": (C.c_int, [vp, vp, C.c_int64, C.c_int, C.c_double, C.c_double, vp, vp,
                                                          vp, vp, vp]),
    "gwn_exact_winding_number_2d": (C.c_int, [f64p, C.c_int64, i32p, C.c_int64, f64p, C.c_int64, f64p]),
    "gwn_read_triangle_mesh": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int64),
                                         C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.c_int64)]),
    "gwn_free": (None, [vp]),
    "gwn_server_add_mesh_file": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int32)]),
    "gwn_server_create": (C.c_int, [C.c_int, C.c_int64, C.POINTER(vp)]),
    "gwn_server_destroy": (None, [vp]),
    "gwn_server_add_mesh": (C.c_int, [vp, f64p, C.c_int64, i32p, C.c_int64, C.c_int, C.c_int,
                                      C.POINTER(C.c_int32)]),
    "gwn_server_submit": (C.c_int, [vp, C.c_int32, vp, C.c_int64, C.c_float, C.c_int, vp, C.POINTER(C.c_int64)]),
    "gwn_server_wait": (C.c_int, [vp, C.c_int64]),
    "gwn_server_drain": (C.c_int, [vp]),
    "gwn_server_request_info": (C.c_int, [vp, C.c_int64, C.POINTER(RequestInfo)]),
    "gwn_server_memory": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gwn_server_now": (C.c_double, [vp]),
    "gwn_server_request_bytes": (C.c_int64, [C.c_int64]),
    "gwn_server_reserve": (C.c_int, [vp, C.c_int64]),
}

_lib = None


class GWNError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gwn error {code}: {msg}")
        self.code = code


def lib():
    """Load libgwn.so; raises if it is not built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgwn.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != GWN_OK:
        msg = lib().gwn_last_error()
        raise GWNError(rc, msg.decode() if msg else "")
    return rc
