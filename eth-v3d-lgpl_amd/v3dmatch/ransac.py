"""Robust relative orientation from match lists (computeRobustOrientationMLE,
V3D/Apps/SfM/reconstruction_common.cpp:139-340) over the C ABI of libv3dmatch.so
(include/v3dm_ransac.h): host code in double precision, no GPU.

Correspondences are (n, 4) arrays (left u, left v, right u, right v) in pixels; K1, K2 the
3x3 intrinsics; poses follow x2 ~ R x1 + T for normalised points, E = [T]x R,
F = K2^-T E K1^-1.
"""
import ctypes as C

import numpy as np

from ._native import V3DMError, load_library


class Orientation(C.Structure):
    _fields_ = [("essential", C.c_double * 9), ("fundamental", C.c_double * 9), ("rotation", C.c_double * 9),
                ("translation", C.c_double * 3), ("error", C.c_double), ("nInliers", C.c_int),
                ("nSamplesDrawn", C.c_int), ("nHypotheses", C.c_int)]


def _check(rc, what):
    if rc:
        raise V3DMError(rc, what)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def sampson_errors(F, corrs):
    """sampsonEpipolarError of every correspondence under F (reconstruction_common.h:588-608)."""
    F = _f64(F, (3, 3))
    c = _f64(corrs).reshape(-1, 4)
    out = np.empty(c.shape[0], np.float64)
    _check(load_library().v3dm_sampson_errors(F.ctypes.data, c.ctypes.data, c.shape[0], out.ctypes.data),
           "sampson_errors")
    return out


def essentials_from_five(x1, y1, x2, y2):
    """The real essential matrices (k, 3, 3) through five normalised correspondences."""
    a = [_f64(v, (5,)) for v in (x1, y1, x2, y2)]
    Es = np.empty((10, 3, 3), np.float64)
    n = C.c_int(0)
    _check(load_library().v3dm_essentials_from_five(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                                    a[3].ctypes.data, Es.ctypes.data, C.byref(n)),
           "essentials_from_five")
    return Es[: n.value].copy()


def relative_pose_from_essential(E, x1, y1, x2, y2):
    """(R, T, ok) for E and normalised correspondences (v3d_poseutilities.cpp:1263-1291)."""
    E = _f64(E, (3, 3))
    a = [_f64(v) for v in (x1, y1, x2, y2)]
    R = np.empty((3, 3), np.float64)
    T = np.empty(3, np.float64)
    ok = C.c_int(0)
    _check(load_library().v3dm_relative_pose_from_essential(E.ctypes.data, a[0].shape[0], a[0].ctypes.data,
                                                            a[1].ctypes.data, a[2].ctypes.data, a[3].ctypes.data,
                                                            R.ctypes.data, T.ctypes.data, C.byref(ok)),
           "relative_pose_from_essential")
    return R, T, bool(ok.value)


def refine_relative_pose(corrs, K1, K2, R, T, sampson_inlier_threshold):
    """refineRelativePose (v3d_poseutilities.cpp:1063-1099); returns the refined (R, T)."""
    c = _f64(corrs).reshape(-1, 4)
    R = _f64(R, (3, 3)).copy()
    T = _f64(T, (3,)).copy()
    _check(load_library().v3dm_refine_relative_pose(c.ctypes.data, c.shape[0], _f64(K1, (3, 3)).ctypes.data,
                                                    _f64(K2, (3, 3)).ctypes.data, R.ctypes.data, T.ctypes.data,
                                                    C.c_double(sampson_inlier_threshold)),
           "refine_relative_pose")
    return R, T


def robust_orientation(corrs, K1, K2, inlier_threshold, n_samples=1000, report_inliers=True, refine=True, seed=1):
    """computeRobustOrientationMLE: dict with essential, fundamental, rotation, translation,
    error, inliers (indices), samples drawn and hypotheses scored."""
    c = _f64(corrs).reshape(-1, 4)
    K1 = _f64(K1, (3, 3))
    K2 = _f64(K2, (3, 3))
    res = Orientation()
    inl = np.empty(max(c.shape[0], 1), np.int32)
    _check(load_library().v3dm_robust_orientation(c.ctypes.data, c.shape[0], K1.ctypes.data, K2.ctypes.data,
                                                  C.c_double(inlier_threshold), int(n_samples), int(bool(report_inliers)),
                                                  int(bool(refine)), C.c_uint64(seed), C.byref(res), inl.ctypes.data),
           "robust_orientation (at least 5 correspondences and invertible K1, K2 are required)")
    return {"essential": np.array(res.essential).reshape(3, 3), "fundamental": np.array(res.fundamental).reshape(3, 3),
            "rotation": np.array(res.rotation).reshape(3, 3), "translation": np.array(res.translation),
            "error": res.error, "inliers": inl[: res.nInliers].copy() if report_inliers else None,
            "n_inliers": res.nInliers, "samples": res.nSamplesDrawn, "hypotheses": res.nHypotheses}
