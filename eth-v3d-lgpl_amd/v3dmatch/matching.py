"""Python mirror of V3D_CUDA::SIFT_Matching over the C ABI (include/v3dm.h).

Method names, argument meaning and defaults follow V3D/CUDA/v3d_cudamatching.h:15-67:
allocate / deallocate / setLeftFeatures / setRightFeatures / findPutativeMatches
(minScore=0.9, minRatio=0.8; minRatio < 0 disables the ratio test) /
runGuidedMatching / runGuidedHomographyMatching / enableRoundingWarning /
enableReductionPath.  Where the reference fills a std::vector<pair<int,int>> the Python
methods return a list of (left, right) tuples sorted by left index.  Errors raise
V3DMError (the reference exits the process, cu:12-27).
"""
import ctypes as C

import numpy as np

from ._native import Stats, V3DMError, load_library

PREC_FP32 = 0
PREC_BF16 = 1
PREC_FP8 = 2
PREC_LEARNED = 3
FALLBACK_THRESHOLD = 0
FALLBACK_BINOMIAL = 1


def _f32_rows(desc, n):
    a = np.ascontiguousarray(desc, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(-1, 128)
    if a.ndim != 2 or a.shape[1] != 128:
        raise ValueError("descriptors must have shape (n, 128)")
    if a.shape[0] < n:
        raise ValueError(f"descriptor array has {a.shape[0]} rows, nFeatures = {n}")
    return a


def _xy(pos, n):
    a = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 2)
    if a.shape[0] < n:
        raise ValueError(f"position array has {a.shape[0]} rows, need {n}")
    return a


class SIFT_Matching:
    """Putative matcher on one GPU (V3D_CUDA::SIFT_Matching, v3d_cudamatching.h:15)."""

    def __init__(self, device=0):
        self._lib = load_library()
        h = C.c_void_p()
        rc = self._lib.v3dm_create(C.byref(h), int(device))
        if rc != 0:
            raise V3DMError(rc, self._lib.v3dm_last_error(None).decode())
        self._h = h
        self._nL = 0
        self._nR = 0
        self._nMaxL = 0
        self._nMaxR = 0
        self._warn = True
        self._reduce = False

    # -- life cycle -----------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.v3dm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise V3DMError(rc, self._lib.v3dm_last_error(self._h).decode())

    # -- reference API ---------------------------------------------------------------
    def enableRoundingWarning(self, flag=True):
        self._warn = bool(flag)
        self._check(self._lib.v3dm_set_flags(self._h, int(self._warn), int(self._reduce)))

    def enableReductionPath(self, flag=True):
        self._reduce = bool(flag)
        self._check(self._lib.v3dm_set_flags(self._h, int(self._warn), int(self._reduce)))

    def allocate(self, nMaxLeftFeatures, nMaxRightFeatures=None, useGuidedMatching=False):
        """allocate(nMaxLeft, nMaxRight, useGuided) (cu:439) or allocate(nMax, useGuided) (h:34)."""
        if isinstance(nMaxRightFeatures, bool):
            useGuidedMatching, nMaxRightFeatures = nMaxRightFeatures, None
        if nMaxRightFeatures is None:
            nMaxRightFeatures = nMaxLeftFeatures
        self._check(self._lib.v3dm_allocate(self._h, int(nMaxLeftFeatures), int(nMaxRightFeatures),
                                            int(bool(useGuidedMatching))))
        self._nMaxL = int(nMaxLeftFeatures) - (int(nMaxLeftFeatures) & 31)
        self._nMaxR = int(nMaxRightFeatures) - (int(nMaxRightFeatures) & 31)

    def deallocate(self):
        self._check(self._lib.v3dm_deallocate(self._h))

    def setLeftFeatures(self, nFeatures, descriptors):
        a = _f32_rows(descriptors, nFeatures) if nFeatures > 0 else None
        self._check(self._lib.v3dm_set_left(self._h, int(nFeatures), a.ctypes.data if a is not None else None))
        self._nL = min(int(nFeatures) - (int(nFeatures) & 31), self._nMaxL)

    def setRightFeatures(self, nFeatures, descriptors):
        a = _f32_rows(descriptors, nFeatures) if nFeatures > 0 else None
        self._check(self._lib.v3dm_set_right(self._h, int(nFeatures), a.ctypes.data if a is not None else None))
        self._nR = min(int(nFeatures) - (int(nFeatures) & 31), self._nMaxR)

    def findPutativeMatches(self, minScore=0.9, minRatio=0.8):
        return [tuple(p) for p in self.find_putative_matches_array(minScore, minRatio).tolist()]

    def runGuidedMatching(self, F, distanceThreshold, leftPositions, rightPositions, minScore=0.9, minRatio=0.8):
        return self._guided(self._lib.v3dm_run_guided_F, F, distanceThreshold, leftPositions, rightPositions,
                            minScore, minRatio)

    def runGuidedHomographyMatching(self, H, distanceThreshold, leftPositions, rightPositions, minScore=0.9,
                                    minRatio=0.8):
        return self._guided(self._lib.v3dm_run_guided_H, H, distanceThreshold, leftPositions, rightPositions,
                            minScore, minRatio)

    # -- array forms and extensions --------------------------------------------------
    def _cap(self):
        return max(1, min(max(self._nL, 0), max(self._nR, 0)))

    def _pairs_out(self, fn, *args):
        cap = self._cap()
        out = np.empty((cap, 2), dtype=np.int32)
        n = C.c_int(0)
        self._check(fn(self._h, *args, out.ctypes.data, cap, C.byref(n)))
        return out[: n.value].copy()

    def find_putative_matches_array(self, minScore=0.9, minRatio=0.8):
        """Matches as an (n, 2) int32 array of (left, right), sorted by left index."""
        return self._pairs_out(self._lib.v3dm_find_putative_matches, C.c_float(minScore), C.c_float(minRatio))

    def _guided(self, fn, M, thr, lpos, rpos, minScore, minRatio):
        return [tuple(p) for p in self._guided_array(fn, M, thr, lpos, rpos, minScore, minRatio).tolist()]

    def _guided_array(self, fn, M, thr, lpos, rpos, minScore, minRatio):
        m = np.ascontiguousarray(np.asarray(M, dtype=np.float64).reshape(3, 3))
        lp = _xy(lpos, self._nL)
        rp = _xy(rpos, self._nR)
        return self._pairs_out(fn, m.ctypes.data, C.c_float(thr), lp.ctypes.data, rp.ctypes.data,
                               C.c_float(minScore), C.c_float(minRatio))

    def run_guided_array(self, homography, M, distanceThreshold, leftPositions, rightPositions, minScore=0.9,
                         minRatio=0.8):
        """runGuidedHomographyMatching (homography=True) or runGuidedMatching with the matches
        as an (n, 2) int32 array instead of a list of tuples."""
        fn = self._lib.v3dm_run_guided_H if homography else self._lib.v3dm_run_guided_F
        return self._guided_array(fn, M, distanceThreshold, leftPositions, rightPositions, minScore, minRatio)

    def set_precision(self, mode):
        self._check(self._lib.v3dm_set_precision(self._h, int(mode)))

    def set_certification(self, eps_rel):
        """Error-bound scale of certified decisions (0 = every query re-checked exactly)."""
        self._check(self._lib.v3dm_set_certification(self._h, C.c_float(eps_rel)))

    def set_learned_topk(self, k):
        self._check(self._lib.v3dm_set_learned_topk(self._h, int(k)))

    def set_learned_margin(self, margin_rel):
        """Learned mode: candidates more than margin_rel * scale^2 below the best code score
        are not rescored (negative: off)."""
        self._check(self._lib.v3dm_set_learned_margin(self._h, C.c_float(margin_rel)))

    def set_fallback(self, samples=128, min_recall=0.995):
        self._check(self._lib.v3dm_set_fallback(self._h, int(samples), C.c_float(min_recall)))

    def set_fallback_rule(self, rule):
        """FALLBACK_THRESHOLD (default: the plain form when the sampled recall is below
        min_recall) or FALLBACK_BINOMIAL (a one-sided binomial test at 5%)."""
        self._check(self._lib.v3dm_set_fallback_rule(self._h, int(rule)))

    def debug_mutual_wait(self, spin_limit=-1, never_ready=False):
        """Test hook of the mutual check's wait-ran-out path (v3dm_debug_mutual_wait):
        debug_mutual_wait(8, True) makes the next calls fail; debug_mutual_wait() restores."""
        self._check(self._lib.v3dm_debug_mutual_wait(self._h, int(spin_limit), int(bool(never_ready))))

    @staticmethod
    def fallback_allowed_misses(accepted, min_recall=0.995, rule=0):
        """Misses a sampled check of `accepted` exact matches tolerates before the plain
        form is run: recall (accepted - m) / accepted not below min_recall (rule 0), or a
        one-sided binomial test at 5% (rule 1)."""
        fn = load_library().v3dm_fallback_allowed_misses_binomial if rule == 1 else \
            load_library().v3dm_fallback_allowed_misses
        return int(fn(int(accepted), C.c_float(min_recall)))

    def set_derived_cache(self, enable=True):
        self._check(self._lib.v3dm_set_derived_cache(self._h, int(bool(enable))))

    def set_stream(self, stream_ptr):
        self._check(self._lib.v3dm_set_stream(self._h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def set_left_u8(self, q):
        q = np.ascontiguousarray(q, dtype=np.uint8).reshape(-1, 128)
        self._check(self._lib.v3dm_set_left_u8(self._h, q.shape[0], q.ctypes.data))
        self._nL = min(q.shape[0] - (q.shape[0] & 31), self._nMaxL)

    def set_right_u8(self, q):
        q = np.ascontiguousarray(q, dtype=np.uint8).reshape(-1, 128)
        self._check(self._lib.v3dm_set_right_u8(self._h, q.shape[0], q.ctypes.data))
        self._nR = min(q.shape[0] - (q.shape[0] & 31), self._nMaxR)

    def set_left_device(self, n, dptr):
        self._check(self._lib.v3dm_set_left_device(self._h, int(n), C.c_void_p(dptr)))
        self._nL = min(int(n) - (int(n) & 31), self._nMaxL)

    def set_right_device(self, n, dptr):
        self._check(self._lib.v3dm_set_right_device(self._h, int(n), C.c_void_p(dptr)))
        self._nR = min(int(n) - (int(n) & 31), self._nMaxR)

    def find_async(self, minScore=0.9, minRatio=0.8):
        self._check(self._lib.v3dm_find_putative_matches_async(self._h, C.c_float(minScore), C.c_float(minRatio)))

    def get_matches(self):
        return self._pairs_out(self._lib.v3dm_get_matches)

    def get_maxima(self):
        cm = np.empty(max(self._nL, 1), dtype=np.int32)
        rm = np.empty(max(self._nR, 1), dtype=np.int32)
        self._check(self._lib.v3dm_get_maxima(self._h, cm.ctypes.data, rm.ctypes.data))
        return cm[: self._nL], rm[: self._nR]

    def time_dominant_kernel(self, reps=20):
        """Average duration (ms) of one launch of the current mode's dominant kernel,
        launched reps times back to back on the last call's inputs."""
        ms = C.c_float(0)
        self._check(self._lib.v3dm_time_dominant_kernel(self._h, int(reps), C.byref(ms)))
        return ms.value

    def score_block(self, mode, n_rows, n_cols):
        """The approximate scores of mode's pass (PREC_FP32 / BF16 / FP8) for the first
        n_rows left x n_cols right features, float32 (n_rows, n_cols)."""
        out = np.empty((max(n_rows, 1), max(n_cols, 1)), dtype=np.float32)
        self._check(self._lib.v3dm_score_block(self._h, int(mode), int(n_rows), int(n_cols), out.ctypes.data))
        return out[:n_rows, :n_cols]

    def get_second_best(self):
        """(colSecond, rowSecond) of the last shortlist-mode call: per feature, the feature
        of the other set with the second best exact score among the rescored candidates."""
        cs = np.empty(max(self._nL, 1), dtype=np.int32)
        rs = np.empty(max(self._nR, 1), dtype=np.int32)
        self._check(self._lib.v3dm_get_second_best(self._h, cs.ctypes.data, rs.ctypes.data))
        return cs[: self._nL], rs[: self._nR]

    def synchronize(self):
        self._check(self._lib.v3dm_synchronize(self._h))

    def set_profiling(self, enable=True):
        self._check(self._lib.v3dm_set_profiling(self._h, int(bool(enable))))

    def stats(self):
        s = Stats()
        self._check(self._lib.v3dm_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    # -- batched matching ----------------------------------------------------------------
    def batch_set_images(self, counts, descriptors):
        """Images of a batch: counts[m] features each, descriptors concatenated row-wise
        (uint8 SIFT descriptors, dequantised on the GPU, or float32)."""
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        d = np.asarray(descriptors)
        if d.dtype == np.uint8:
            d = np.ascontiguousarray(d).reshape(-1, 128)
            fn = self._lib.v3dm_batch_set_images_u8
        else:
            d = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 128)
            fn = self._lib.v3dm_batch_set_images
        if d.shape[0] != int(counts.sum()):
            raise ValueError("descriptor rows do not match the counts")
        self._check(fn(self._h, counts.shape[0], counts.ctypes.data, d.ctypes.data))
        self._bcounts = counts.copy()

    def batch_set_pairs(self, left, right):
        left = np.ascontiguousarray(left, dtype=np.int32)
        right = np.ascontiguousarray(right, dtype=np.int32)
        self._check(self._lib.v3dm_batch_set_pairs(self._h, left.shape[0], left.ctypes.data, right.ctypes.data))
        rounded = self._bcounts - (self._bcounts & 31)
        self._bcap = max(1, int(rounded[left].sum()))
        self._bpairs = left.shape[0]

    def batch_match_async(self, minScore=0.9, minRatio=0.8):
        self._check(self._lib.v3dm_batch_match_async(self._h, C.c_float(minScore), C.c_float(minRatio)))

    def batch_get_matches(self):
        """(offsets int32 (nPairs+1,), pairs int32 (total, 2)) of the last batch run."""
        offs = np.empty(self._bpairs + 1, dtype=np.int32)
        out = np.empty((self._bcap, 2), dtype=np.int32)
        n = C.c_int(0)
        self._check(self._lib.v3dm_batch_get_matches(self._h, offs.ctypes.data, out.ctypes.data, self._bcap,
                                                      C.byref(n)))
        return offs, out[: n.value].copy()

    def batch_match(self, minScore=0.9, minRatio=0.8):
        """Matches of every pair of the batch: a list of (n_p, 2) int32 arrays."""
        self.batch_match_async(minScore, minRatio)
        offs, pairs = self.batch_get_matches()
        return [pairs[offs[p]:offs[p + 1]] for p in range(self._bpairs)]

    def batch_set_tuning(self, autotune=True, mode=0, tiles_per_item=0, grid=0):
        self._check(self._lib.v3dm_batch_set_tuning(self._h, int(bool(autotune)), int(mode), int(tiles_per_item),
                                                     int(grid)))

    def batch_set_tile_shape(self, mode, query_rows=0):
        """The work-group shape of the pass of `mode`'s form: 128 or 256 query rows (plain:
        left features) per 64-candidate tile; 0: chosen by the autotuner."""
        self._check(self._lib.v3dm_batch_set_tile_shape(self._h, int(mode), int(query_rows)))

    def batch_get_tile_shape(self, mode=0):
        q, c = C.c_int(0), C.c_int(0)
        self._check(self._lib.v3dm_batch_get_tile_shape(self._h, int(mode), C.byref(q), C.byref(c)))
        return {"query_rows": q.value, "candidate_rows": c.value}

    def batch_get_tuning(self, mode=0):
        t, g, n, ms = C.c_int(0), C.c_int(0), C.c_int(0), C.c_float(0)
        self._check(self._lib.v3dm_batch_get_tuning(self._h, int(mode), C.byref(t), C.byref(g), C.byref(n),
                                                     C.byref(ms)))
        return {"tiles_per_item": t.value, "grid": g.value, "items": n.value, "pass_ms": ms.value}

    def batch_use_graph(self, enable=True):
        self._check(self._lib.v3dm_batch_use_graph(self._h, int(bool(enable))))

    def batch_get_fallback(self):
        """Per pair of the last learned batch run: 1 if its sampled check sent it to the
        plain form."""
        out = np.zeros(self._bpairs, dtype=np.int32)
        self._check(self._lib.v3dm_batch_get_fallback(self._h, out.ctypes.data))
        return out

    def batch_get_check(self):
        """Per pair of the last learned batch run: (exact matches among the samples, missed,
        decided by exact FP64 scores, samples), an int32 (nPairs, 4) array."""
        out = np.zeros((self._bpairs, 4), dtype=np.int32)
        self._check(self._lib.v3dm_batch_get_check(self._h, out.ctypes.data))
        return out

    def batch_time_dominant_kernel(self, reps=10):
        """Average ms of one launch of the last batch run's pass kernel (launched reps times
        back to back between one event pair)."""
        ms = C.c_float(0)
        self._check(self._lib.v3dm_batch_time_dominant_kernel(self._h, int(reps), C.byref(ms)))
        return ms.value

    def batch_last_ms(self):
        ms = C.c_float(0)
        self._check(self._lib.v3dm_batch_last_ms(self._h, C.byref(ms)))
        return ms.value

    @property
    def nLeft(self):
        return self._nL

    @property
    def nRight(self):
        return self._nR
