"""Two-view geometry checks for the robust-orientation consumer -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and by the CPU measurement leg of bench.py, never by the library.  Each
function restates in numpy (float64) what the reference computes, citing it, or states a
property that the mathematics fixes (the epipolar and essential-matrix constraints, the
pose of a synthetic scene), so that the library's results are checked against something
it does not share code with:

* sampson_error: sampsonEpipolarError (V3D/Apps/SfM/reconstruction_common.h:588-608).
* shuffle_count: getShuffleCount (V3D/Geometry/v3d_homography.cpp:14-18).
* ransac_n_samples: ransacNSamples(nNeededSamples, outlierRatio, wantedPsucces)
  (reconstruction_common.h:642-650).
* essential_residuals: det E, 2 E E^T E - tr(E E^T) E and the epipolar constraints (the ten
  cubic constraints and the five linear ones that define computeEssentialsFromFiveCorrs'
  solutions, v3d_poseutilities.cpp:1015-1060).
* synthetic_two_view: a scene with known K, R, T (x2 ~ R x1 + T), Gaussian pixel noise
  and uniform outliers; rotation_error_deg / direction_error_deg measure a recovered pose.
"""
import math

import numpy as np


def sampson_error(F, corrs):
    """Element by element in the reference's order, so that the float64 roundings are the
    same: the positions are single precision there (PointMeasurement::pos is a Vector2f,
    V3D/Geometry/v3d_mviewutilities.h:14-17); multiply_A_v_affine (V3D/Math/
    v3d_linearbase.h:1048-1061) starts each sum from the last column, (F p)_r = F[r,2] +
    F[r,0] p_x + F[r,1] p_y, and (F^T q)_c = F[2,c] + F[0,c] q_x + F[1,c] q_y; then the
    numerator and denominator sums left to right."""
    F = np.asarray(F, np.float64).reshape(3, 3)
    c = np.asarray(corrs, np.float64).reshape(-1, 4).astype(np.float32).astype(np.float64)
    px, py, qx, qy = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    fp0 = F[0, 2] + F[0, 0] * px + F[0, 1] * py
    fp1 = F[1, 2] + F[1, 0] * px + F[1, 1] * py
    fp2 = F[2, 2] + F[2, 0] * px + F[2, 1] * py
    ftq0 = F[2, 0] + F[0, 0] * qx + F[1, 0] * qy
    ftq1 = F[2, 1] + F[0, 1] * qx + F[1, 1] * qy
    tmp = qx * fp0 + qy * fp1 + fp2
    return (tmp * tmp) / (fp0 * fp0 + fp1 * fp1 + ftq0 * ftq0 + ftq1 * ftq1)


def shuffle_count(n, n_rounds, min_sample_size):
    """getShuffleCount: the product and quotient in float, the + 0.5 in double (a double
    literal), truncated."""
    q = np.float32(np.float32(n_rounds) * np.float32(min_sample_size)) / np.float32(n)
    return max(1, int(float(np.float32(q)) + 0.5))


def _ieee_log(x):
    return math.log(x) if x > 0.0 else (-math.inf if x == 0.0 else math.nan)


def _ieee_div(a, b):
    if b != 0.0:
        return a / b
    return math.nan if a == 0.0 or math.isnan(a) else math.copysign(math.inf, a) * math.copysign(1.0, b)


def ransac_n_samples(n_needed, outlier_ratio, wanted):
    """ransacNSamples with C's semantics: the C library's log and pow (math.log, float **),
    IEEE division (a zero denominator gives an infinity, so an outlier fraction so close to
    1 that 1 - (1 - r)^5 rounds to 1 gives 0 samples), and std::min / std::max clamping."""
    max_n = 1e8
    if outlier_ratio >= 1.0:
        return int(max_n)
    den = _ieee_log(1.0 - float(1.0 - outlier_ratio) ** float(n_needed))
    n = _ieee_div(_ieee_log(1.0 - wanted), den)
    n = n if n < max_n else max_n  # std::min(maxN, N): NaN gives maxN
    n = n if 0.0 < n else 0.0  # std::max(0.0, N)
    return int(n)


def skew(t):
    return np.array([[0.0, -t[2], t[1]], [t[2], 0.0, -t[0]], [-t[1], t[0], 0.0]])


def essential_from_pose(R, T):
    return skew(T) @ R


def essential_residuals(E, x1, y1, x2, y2):
    """(|det E|, |2 E E^T E - tr(E E^T) E|_max, max |x2^T E x1|) for unit-norm E."""
    E = np.asarray(E, np.float64) / np.linalg.norm(E)
    EEt = E @ E.T
    cubic = 2.0 * EEt @ E - np.trace(EEt) * E
    p = np.c_[x1, y1, np.ones(len(x1))]
    q = np.c_[x2, y2, np.ones(len(x2))]
    epi = np.abs((q @ E * p).sum(1)).max()
    return abs(np.linalg.det(E)), np.abs(cubic).max(), epi


def rotation(w):
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    if th == 0.0:
        return np.eye(3)
    K = skew(w / th)
    return np.eye(3) + np.sin(th) * K + (1.0 - np.cos(th)) * K @ K


def synthetic_two_view(rng, n_in, n_out, K, R, T, noise_px, depth=(4.0, 12.0), size=(640, 480)):
    """Correspondences (n, 4) in pixels, shuffled, and the inlier mask."""
    X = np.c_[rng.uniform(-3, 3, (n_in, 2)), rng.uniform(depth[0], depth[1], n_in)]
    u1 = X @ K.T
    u1 = u1[:, :2] / u1[:, 2:3]
    X2 = X @ R.T + T
    u2 = X2 @ K.T
    u2 = u2[:, :2] / u2[:, 2:3]
    u1 = u1 + rng.normal(0, noise_px, u1.shape)
    u2 = u2 + rng.normal(0, noise_px, u2.shape)
    o = rng.uniform([0, 0, 0, 0], [size[0], size[1], size[0], size[1]], (n_out, 4))
    corrs = np.r_[np.c_[u1, u2], o]
    perm = rng.permutation(len(corrs))
    return corrs[perm], perm < n_in


def rotation_error_deg(Ra, Rb):
    c = (np.trace(np.asarray(Ra).T @ np.asarray(Rb)) - 1.0) / 2.0
    return float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))


def direction_error_deg(ta, tb):
    ta = np.asarray(ta) / np.linalg.norm(ta)
    tb = np.asarray(tb) / np.linalg.norm(tb)
    return float(np.degrees(np.arccos(np.clip(ta @ tb, -1.0, 1.0))))


# ---------------------------------------------------------------------------------------
# computeRobustOrientationMLE (V3D/Apps/SfM/reconstruction_common.cpp:139-340) restated
# step by step, with its building blocks passed in (the tests pass the reference's own,
# compiled from its sources into oracle/_ref).  Arithmetic in the reference's order: 3x3
# products accumulated from zero over k (multiply_A_B, V3D/Math/v3d_linearbase.h:1294-1310),
# the inverse of invertedMatrix (v3d_linear.h:558-575), normalised points stored as floats
# (multiply_A_v_projective into a Vector2f), errors summed in index order.  The shuffles are
# the library's generator (splitmix64 + Fisher-Yates from the end), restated so that both
# sides draw the same samples; the reference draws them with std::random_shuffle.

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def below(self, n):
        return ((self.next() >> 32) * n) >> 32


def _mm(a, b):
    out = [[0.0] * 3 for _ in range(3)]
    for r in range(3):
        for c in range(3):
            acc = 0.0
            for k in range(3):
                acc += a[r][k] * b[k][c]
            out[r][c] = acc
    return out


def _inv3(A):
    a, b, c = (float(v) for v in A[0])
    d, e, f = (float(v) for v in A[1])
    g, h, i = (float(v) for v in A[2])
    det = a * e * i + b * f * g + c * d * h - c * e * g - b * d * i - a * f * h
    s = 1.0 / det
    adj = [[e * i - f * h, c * h - b * i, b * f - c * e],
           [f * g - d * i, a * i - c * g, c * d - a * f],
           [d * h - e * g, b * g - a * h, a * e - b * d]]
    return [[v * s for v in row] for row in adj]


def _f32(x):
    return float(np.float32(x))


def _normalise_f32(iK, u, v):
    w = _f32(iK[2][2])
    w = _f32(w + iK[2][0] * u)
    w = _f32(w + iK[2][1] * v)
    d = []
    for r in range(2):
        x = _f32(iK[r][2])
        x = _f32(x + iK[r][0] * u)
        x = _f32(x + iK[r][1] * v)
        d.append(float(np.float32(x) / np.float32(w)))
    return d


def _cross_matrix(t):
    return [[0.0, -t[2], t[1]], [t[2], 0.0, -t[0]], [-t[1], t[0], 0.0]]


def robust_orientation_mle(corrs, K1, K2, threshold, n_samples, refine, seed, five, pose, refine_pose):
    """five(x1, y1, x2, y2) -> list of 3x3 E; pose(E, x1, y1, x2, y2) -> (R, T, ok);
    refine_pose(pixel corrs, K1, K2, R, T, threshold) -> (R, T).  Returns the result fields of
    RobustOrientationResult plus the samples drawn and the hypotheses scored."""
    c = np.asarray(corrs, np.float64).reshape(-1, 4).astype(np.float32).astype(np.float64)
    N = len(c)
    sqr = threshold * threshold
    iK1, iK2 = _inv3(K1), _inv3(K2)
    iK2t = [[iK2[k][r] for k in range(3)] for r in range(3)]
    nl = [_normalise_f32(iK1, c[i, 0], c[i, 1]) for i in range(N)]
    nr = [_normalise_f32(iK2, c[i, 2], c[i, 3]) for i in range(N)]
    n_shuffles = shuffle_count(N, n_samples, 5)
    min_error = 1e30
    indices = list(range(N))
    rng = SplitMix64(seed)
    drawn = hyps = 0
    res = None

    def score(F):
        d = sampson_error(np.array(F), c)
        return float(np.add.accumulate(np.minimum(d, sqr))[-1]), int((d < sqr).sum()), d

    for _ in range(n_shuffles):
        if drawn > n_samples:
            break
        for i in range(N - 1, 0, -1):
            j = rng.below(i + 1)
            indices[i], indices[j] = indices[j], indices[i]
        for j in range(0, N - 5, 5):
            if drawn > n_samples:
                break
            ix = indices[j:j + 5]
            x1 = np.array([nl[k][0] for k in ix]); y1 = np.array([nl[k][1] for k in ix])
            x2 = np.array([nr[k][0] for k in ix]); y2 = np.array([nr[k][1] for k in ix])
            Es = five(x1, y1, x2, y2)
            drawn += 1
            for E in Es:
                hyps += 1
                E = [[float(v) for v in row] for row in np.asarray(E).reshape(3, 3)]
                fund = _mm(_mm(iK2t, E), iK1)
                cur, n_inl, d = score(fund)
                if not cur < min_error:
                    continue
                R, T, ok = pose(np.array(E), x1, y1, x2, y2)
                if not ok:
                    continue
                if refine:
                    inl = np.flatnonzero(d < sqr)
                    R, T = refine_pose(c[inl], K1, K2, R, T, threshold)
                    E = _mm(_cross_matrix([float(v) for v in T]), [[float(v) for v in row] for row in R])
                    fund = _mm(_mm(iK2t, E), iK1)
                    cur_ref, n_inl, _ = score(fund)
                    cur = min(cur, cur_ref)
                min_error = cur
                res = dict(error=min_error / N, essential=np.array(E), fundamental=np.array(fund),
                           rotation=np.array(R, np.float64).reshape(3, 3), translation=np.array(T, np.float64))
                outlier = 1.0 - float(np.float32(n_inl - 5) / np.float32(N - 5))
                n_samples = min(n_samples, ransac_n_samples(5, outlier, 1.0 - 10e-10))
    out = dict(samples=drawn, hypotheses=hyps)
    if res is not None:
        out.update(res)
        out["inliers"] = np.flatnonzero(sampson_error(res["fundamental"], c) < sqr).astype(np.int32)
    return out
