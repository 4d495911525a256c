"""ctypes binding of the oracle libraries -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by the smoke check and by the CPU-baseline leg of bench.py, never by
the matcher library itself.  See v3dm_oracle.c for the semantics and their citations.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_cpu = None


def _declare(lib):
    vp, i, f = C.c_void_p, C.c_int, C.c_float
    sig = {
        "v3do_decide": (i, [f, f, i, f, f]),
        "v3do_exact_dot": (f, [vp, vp]),
        "v3do_scores": (None, [vp, i, vp, i, vp]),
        "v3do_top2_scan": (None, [vp, i, vp, i, vp, vp, vp]),
        "v3do_column_top2": (None, [vp, i, i, vp, vp, vp]),
        "v3do_row_top2": (None, [vp, i, i, vp, vp, vp]),
        "v3do_column_maxima": (None, [vp, i, i, f, f, vp]),
        "v3do_row_maxima": (None, [vp, i, i, f, f, vp]),
        "v3do_mutual": (i, [vp, i, vp, i, vp, i]),
        "v3do_maxima": (None, [vp, i, vp, i, f, f, vp, vp]),
        "v3do_find_putative_matches": (i, [vp, i, vp, i, f, f, vp, i]),
        "v3do_round32": (i, [i]),
        "v3do_effective_count": (i, [i, i]),
        "v3do_mask_epipolar": (None, [vp, i, i, f, vp, vp, vp]),
        "v3do_mask_homography": (None, [vp, i, i, f, vp, vp, vp]),
        "v3do_guided": (i, [vp, i, i, i, vp, f, vp, vp, f, f, vp, i]),
        "v3do_voctree_leaf": (C.c_int32, [vp, i, i, vp, i]),
        "v3do_voctree_leaves": (None, [vp, i, i, vp, i, vp, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def use_library(path):
    """Point this module at another build of the oracle (planted-mistake variants)."""
    global _lib
    _lib = _declare(C.CDLL(path))


def _load():
    global _lib
    if _lib is None:
        p = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(p):
            raise RuntimeError(f"{p} missing: run oracle/build.py")
        lib = C.CDLL(p)
        _declare(lib)
        _lib = lib
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def decide(best, nxt, arg, min_score=0.9, min_ratio=0.8):
    return _load().v3do_decide(best, nxt, arg, min_score, min_ratio)


def exact_dot(a, b):
    a, b = _f32(a), _f32(b)
    return _load().v3do_exact_dot(a.ctypes.data, b.ctypes.data)


def scores(L, R):
    """Exact score matrix S[j, i] = fl32(sum_k double(L_ik) double(R_jk)), shape (nR, nL)."""
    L, R = _f32(L), _f32(R)
    S = np.empty((R.shape[0], L.shape[0]), dtype=np.float32)
    _load().v3do_scores(L.ctypes.data, L.shape[0], R.ctypes.data, R.shape[0], S.ctypes.data)
    return S


def top2_scan(Q, Cand):
    """Per query: (best, second, argbest) over candidates in ascending order, exact scores."""
    Q, Cand = _f32(Q), _f32(Cand)
    n = Q.shape[0]
    b = np.empty(n, np.float32)
    s = np.empty(n, np.float32)
    a = np.empty(n, np.int32)
    _load().v3do_top2_scan(Q.ctypes.data, n, Cand.ctypes.data, Cand.shape[0], b.ctypes.data, s.ctypes.data,
                           a.ctypes.data)
    return b, s, a


def column_top2(S):
    S = _f32(S)
    nr, nc = S.shape
    b, s, a = np.empty(nc, np.float32), np.empty(nc, np.float32), np.empty(nc, np.int32)
    _load().v3do_column_top2(S.ctypes.data, nr, nc, b.ctypes.data, s.ctypes.data, a.ctypes.data)
    return b, s, a


def row_top2(S):
    S = _f32(S)
    nr, nc = S.shape
    b, s, a = np.empty(nr, np.float32), np.empty(nr, np.float32), np.empty(nr, np.int32)
    _load().v3do_row_top2(S.ctypes.data, nr, nc, b.ctypes.data, s.ctypes.data, a.ctypes.data)
    return b, s, a


def column_maxima(S, min_score=0.9, min_ratio=0.8):
    S = _f32(S)
    out = np.empty(S.shape[1], np.int32)
    _load().v3do_column_maxima(S.ctypes.data, S.shape[0], S.shape[1], min_score, min_ratio, out.ctypes.data)
    return out


def row_maxima(S, min_score=0.9, min_ratio=0.8):
    S = _f32(S)
    out = np.empty(S.shape[0], np.int32)
    _load().v3do_row_maxima(S.ctypes.data, S.shape[0], S.shape[1], min_score, min_ratio, out.ctypes.data)
    return out


def mutual(col_max, row_max):
    cm = np.ascontiguousarray(col_max, np.int32)
    rm = np.ascontiguousarray(row_max, np.int32)
    cap = max(1, min(len(cm), len(rm)))
    out = np.empty((cap, 2), np.int32)
    n = _load().v3do_mutual(cm.ctypes.data, len(cm), rm.ctypes.data, len(rm), out.ctypes.data, cap)
    return out[:n].copy()


def maxima(L, R, min_score=0.9, min_ratio=0.8):
    """(colMax[nL], rowMax[nR]) from descriptors with exact scores."""
    L, R = _f32(L), _f32(R)
    cm = np.empty(max(L.shape[0], 1), np.int32)
    rm = np.empty(max(R.shape[0], 1), np.int32)
    _load().v3do_maxima(L.ctypes.data, L.shape[0], R.ctypes.data, R.shape[0], min_score, min_ratio,
                        cm.ctypes.data, rm.ctypes.data)
    return cm[: L.shape[0]], rm[: R.shape[0]]


def find_putative_matches(L, R, min_score=0.9, min_ratio=0.8):
    """Exact match set of already-rounded feature sets, (n, 2) int32 sorted by left."""
    L, R = _f32(L), _f32(R)
    cap = max(1, min(L.shape[0], R.shape[0]))
    out = np.empty((cap, 2), np.int32)
    n = _load().v3do_find_putative_matches(L.ctypes.data, L.shape[0], R.ctypes.data, R.shape[0], min_score,
                                           min_ratio, out.ctypes.data, cap)
    return out[:n].copy()


def round32(n):
    return _load().v3do_round32(n)


def effective_count(n, n_max):
    return _load().v3do_effective_count(n, n_max)


def mask_epipolar(S, sqr_thr, F, left_xy, right_xy):
    S = _f32(S).copy()
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    lx, rx = _f32(left_xy), _f32(right_xy)
    _load().v3do_mask_epipolar(S.ctypes.data, S.shape[0], S.shape[1], sqr_thr, F.ctypes.data, lx.ctypes.data,
                               rx.ctypes.data)
    return S


def mask_homography(S, sqr_thr, H, left_xy, right_xy):
    S = _f32(S).copy()
    H = np.ascontiguousarray(H, np.float64).reshape(9)
    lx, rx = _f32(left_xy), _f32(right_xy)
    _load().v3do_mask_homography(S.ctypes.data, S.shape[0], S.shape[1], sqr_thr, H.ctypes.data, lx.ctypes.data,
                                 rx.ctypes.data)
    return S


def guided(S, kind, M, thr, left_xy, right_xy, min_score=0.9, min_ratio=0.8):
    """Guided matching on a stored score matrix S (nR, nL), modified in place."""
    assert S.dtype == np.float32 and S.flags.c_contiguous
    M = np.ascontiguousarray(M, np.float64).reshape(9)
    lx, rx = _f32(left_xy), _f32(right_xy)
    nR, nL = S.shape
    cap = max(1, min(nL, nR))
    out = np.empty((cap, 2), np.int32)
    n = _load().v3do_guided(S.ctypes.data, nL, nR, kind, M.ctypes.data, thr, lx.ctypes.data, rx.ctypes.data,
                            min_score, min_ratio, out.ctypes.data, cap)
    return out[:n].copy()


def voctree_node_count(branch, depth):
    return sum(branch ** (d + 1) for d in range(depth))


def voctree_leaves(tree, branch, depth, keys, global_ids=False):
    """findBestLeafIndex of every key (v3d_voctree.h:63-99); global_ids: node ids as the
    GPU traversal writes them (cu:784-791)."""
    tree = _f32(tree).reshape(-1, 128)
    assert tree.shape[0] == voctree_node_count(branch, depth)
    keys = _f32(keys).reshape(-1, 128)
    out = np.empty(keys.shape[0], np.int32)
    _load().v3do_voctree_leaves(tree.ctypes.data, branch, depth, keys.ctypes.data, keys.shape[0], out.ctypes.data,
                                1 if global_ids else 0)
    return out


# ------------------------------------------------------------------ CPU baseline
def _load_cpu():
    global _cpu
    if _cpu is None:
        p = os.path.join(HERE, "libcpubase.so")
        if not os.path.exists(p):
            raise RuntimeError(f"{p} missing: run oracle/build.py")
        lib = C.CDLL(p)
        lib.v3dcpu_find_putative_matches.restype = C.c_int
        lib.v3dcpu_find_putative_matches.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float,
                                                     C.c_float, C.c_void_p, C.c_int, C.c_int]
        lib.v3dcpu_max_threads.restype = C.c_int
        lib.v3dcpu_max_threads.argtypes = []
        _cpu = lib
    return _cpu


def cpu_baseline_find(L, R, min_score=0.9, min_ratio=0.8, threads=0):
    L, R = _f32(L), _f32(R)
    cap = max(1, min(L.shape[0], R.shape[0]))
    out = np.empty((cap, 2), np.int32)
    n = _load_cpu().v3dcpu_find_putative_matches(L.ctypes.data, L.shape[0], R.ctypes.data, R.shape[0],
                                                 min_score, min_ratio, out.ctypes.data, cap, int(threads))
    return out[:n].copy()


def cpu_max_threads():
    return _load_cpu().v3dcpu_max_threads()
