"""Synthetic SIFT-like descriptor sets (ctypes binding of libv3dsynth.so).

Generator definition (eth-v3d-lgpl_amd/csrc/v3dm_synth.c): Gamma(0.5,1) components,
L2 normalise, clip at 0.2, renormalise, quantise round(512 x) clamped to 255 (uint8),
dequantise as SIFT_Feature::normalizeDescriptor (V3D/Base/v3d_feature.h:30-38).
Right sets mix noisy copies of left descriptors (sigma = 0.05, clamped at 0,
re-quantised) with fresh distractors, randomly permuted.
"""
import ctypes as C
import os

import numpy as np

from ._native import V3DMError, lib_path

BASE_SEED = 0x5EED_E7A3D
SIGMA = 0.05

_lib = None


def _load():
    global _lib
    if _lib is None:
        p = lib_path("libv3dsynth.so")
        if not os.path.exists(p):
            raise V3DMError(5, f"{p} is missing: build it with eth-v3d-lgpl_amd/build.py")
        lib = C.CDLL(p)
        lib.v3ds_seed.restype = C.c_uint64
        lib.v3ds_seed.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        lib.v3ds_dequantise.restype = None
        lib.v3ds_dequantise.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        lib.v3ds_fresh.restype = None
        lib.v3ds_fresh.argtypes = [C.c_uint64, C.c_int, C.c_void_p]
        lib.v3ds_make_pair.restype = C.c_int
        lib.v3ds_make_pair.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.v3ds_batch_sizes.restype = None
        lib.v3ds_batch_sizes.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = lib
    return _lib


def seed(base, config, pair, side):
    return int(_load().v3ds_seed(base, config, pair, side))


def fresh(seed_value, n):
    q = np.empty((n, 128), dtype=np.uint8)
    _load().v3ds_fresh(C.c_uint64(seed_value), int(n), q.ctypes.data)
    return q


def dequantise(q):
    """uint8 (n,128) -> float32 (n,128), as SIFT_Feature::normalizeDescriptor."""
    q = np.ascontiguousarray(q, dtype=np.uint8).reshape(-1, 128)
    out = np.empty(q.shape, dtype=np.float32)
    _load().v3ds_dequantise(q.ctypes.data, q.shape[0], out.ctypes.data)
    return out


def make_pair(nL, nR, copy_frac=0.4, config=1, pair=0, base=BASE_SEED, sigma=SIGMA):
    """One image pair: (qL uint8 (nL,128), qR uint8 (nR,128), truth int32 (nR,))."""
    qL = np.empty((nL, 128), dtype=np.uint8)
    qR = np.empty((nR, 128), dtype=np.uint8)
    truth = np.empty(nR, dtype=np.int32)
    _load().v3ds_make_pair(C.c_uint64(base), config, pair, int(nL), int(nR), float(copy_frac), float(sigma),
                           qL.ctypes.data, qR.ctypes.data, truth.ctypes.data)
    return qL, qR, truth


def batch_sizes(n_pairs, lo=500, hi=16000, frac_lo=0.3, frac_hi=0.6, config=3, base=BASE_SEED):
    """Per-pair (nL, nR, copy_frac) of an uneven batch: n = floor(exp(U(ln lo, ln hi)))."""
    nL = np.empty(n_pairs, dtype=np.int32)
    nR = np.empty(n_pairs, dtype=np.int32)
    fr = np.empty(n_pairs, dtype=np.float64)
    _load().v3ds_batch_sizes(C.c_uint64(base), config, int(n_pairs), int(lo), int(hi), float(frac_lo),
                             float(frac_hi), nL.ctypes.data, nR.ctypes.data, fr.ctypes.data)
    return nL, nR, fr


def make_batch(n_pairs, lo=500, hi=16000, frac_lo=0.3, frac_hi=0.6, config=3, base=BASE_SEED):
    """An uneven batch of image pairs in the layout of the batch API: pair p matches image
    2p (left) against image 2p + 1 (right).  Returns (counts int32 (2P,), uint8 descriptors
    (sum(counts), 128), left int32 (P,), right int32 (P,), copy fractions (P,))."""
    nL, nR, fr = batch_sizes(n_pairs, lo, hi, frac_lo, frac_hi, config=config, base=base)
    counts = np.empty(2 * n_pairs, dtype=np.int32)
    counts[0::2] = nL
    counts[1::2] = nR
    desc = np.empty((int(counts.sum()), 128), dtype=np.uint8)
    row = 0
    for p in range(n_pairs):
        qL, qR, _ = make_pair(int(nL[p]), int(nR[p]), float(fr[p]), config=config, pair=p, base=base)
        desc[row:row + nL[p]] = qL
        row += int(nL[p])
        desc[row:row + nR[p]] = qR
        row += int(nR[p])
    left = np.arange(0, 2 * n_pairs, 2, dtype=np.int32)
    return counts, desc, left, left + 1, fr


def two_view_scene(n_in, n_out, seed=0, noise_px=0.5, focal=800.0, size=(640, 480)):
    """Correspondences (n_in + n_out, 4) of a synthetic two-view scene in pixels (left u, v,
    right u, v), shuffled: n_in points in front of both cameras with Gaussian pixel noise
    and n_out uniform outliers.  Returns (corrs, K, R, T, inlier mask); x2 ~ R x1 + T."""
    rng = np.random.default_rng(seed)
    K = np.array([[focal, 0.0, size[0] / 2.0], [0.0, focal, size[1] / 2.0], [0.0, 0.0, 1.0]])
    w = np.array([0.05, -0.2, 0.03]) + 0.02 * rng.standard_normal(3)
    th = np.linalg.norm(w)
    k = w / th
    S = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    R = np.eye(3) + np.sin(th) * S + (1.0 - np.cos(th)) * S @ S
    T = np.array([1.0, 0.1, 0.2]) + 0.05 * rng.standard_normal(3)
    T /= np.linalg.norm(T)
    X = np.c_[rng.uniform(-3, 3, (n_in, 2)), rng.uniform(4, 12, n_in)]
    u1 = X @ K.T
    u2 = (X @ R.T + T) @ K.T
    u1 = u1[:, :2] / u1[:, 2:3] + rng.normal(0, noise_px, (n_in, 2))
    u2 = u2[:, :2] / u2[:, 2:3] + rng.normal(0, noise_px, (n_in, 2))
    o = rng.uniform([0, 0, 0, 0], [size[0], size[1], size[0], size[1]], (n_out, 4))
    corrs = np.r_[np.c_[u1, u2], o]
    perm = rng.permutation(len(corrs))
    return corrs[perm], K, R, T, perm < n_in
